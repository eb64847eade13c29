/*
 * Registration of the HIP compute backend into QGTL's backend registry
 * (compute_backend.h:439 compute_register_backend, called at load time as
 * COMPUTE_REGISTER_BACKEND does, :523-534).  This program stands in for QGTL's
 * registry with its own compute_register_backend (exported with -rdynamic) and
 * checks that loading libqgtl_mi355x registered the HIP backend once, in the
 * GPU slot, with a complete vtable.  No GPU work is done.
 */
#include <stdio.h>

#include "qgtl_compute_backend.h"

static const ComputeBackendInfo *g_seen;
static int g_calls;

ComputeResult compute_register_backend(const ComputeBackendInfo *info)
{
    g_seen = info;
    g_calls++;
    return COMPUTE_SUCCESS;
}

int main(void)
{
    if (g_calls != 1 || g_seen != qgtl_hip_compute_backend_info()) {
        fprintf(stderr, "FAIL: registered %d times\n", g_calls);
        return 1;
    }
    const ComputeBackendOps *o = g_seen->ops;
    const void *slots[] = {(void *)o->init, (void *)o->cleanup, (void *)o->probe, (void *)o->get_capabilities,
                           (void *)o->alloc, (void *)o->free, (void *)o->memcpy, (void *)o->memset,
                           (void *)o->create_stream, (void *)o->destroy_stream, (void *)o->synchronize_stream,
                           (void *)o->create_event, (void *)o->destroy_event, (void *)o->record_event,
                           (void *)o->wait_event, (void *)o->quantum_unitary, (void *)o->quantum_normalize,
                           (void *)o->quantum_tensor_contract, (void *)o->quantum_gradient,
                           (void *)o->quantum_inner_product, (void *)o->quantum_expectation, (void *)o->barrier,
                           (void *)o->broadcast, (void *)o->allreduce, (void *)o->scatter, (void *)o->gather,
                           (void *)o->allgather, (void *)o->reduce_scatter, (void *)o->execute,
                           (void *)o->create_plan, (void *)o->destroy_plan, (void *)o->get_metrics,
                           (void *)o->reset_metrics};
    for (size_t i = 0; i < sizeof slots / sizeof slots[0]; i++)
        if (!slots[i]) {
            fprintf(stderr, "FAIL: slot %zu empty\n", i);
            return 1;
        }
    if (g_seen->type != COMPUTE_BACKEND_CUDA || g_seen->priority != 100 || o != qgtl_hip_compute_backend_ops()) {
        fprintf(stderr, "FAIL: registration record\n");
        return 1;
    }
    printf("ok registered \"%s\" %s, priority %d\n", g_seen->name, g_seen->version, g_seen->priority);
    return 0;
}
