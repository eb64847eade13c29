/*
 * qgtl_io.c -- files for computed QGTs and learned-preconditioner training sets.
 *
 * QGTL declares geometric_tensor_save / geometric_tensor_load
 * (include/quantum_geometric/core/quantum_geometric_tensor.h:144-147) with no
 * implementation and no file format.  Both are exported here over the
 * library's own container (kind "QGTLTEN1"), beside the formats for what the
 * QGT path produces, so that generated QGT datasets round-trip (MLP
 * checkpoints: qgtl_mlp_save / qgtl_mlp_load in qgtl_mlp.c, same container).
 *
 * Container (little-endian, every field fixed width):
 *   char     magic[8]        "QGTLQGT1" | "QGTLSET1" | "QGTLMLP1" | "QGTLTEN1"
 *   uint32   version         1
 *   uint32   header_bytes    size of the kind-specific header that follows
 *   uint64   payload_bytes
 *   uint64   checksum        FNV-1a 64 over the kind header and the payload
 *   kind header, payload
 * A reader checks magic, version, sizes (against the header's own counts) and
 * the checksum before it hands anything back.
 */
#define _POSIX_C_SOURCE 200809L
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "qgtl_api.h"
#include "qgtl_internal.h"

#define QGTL_IO_VERSION 1u

typedef struct {
    char magic[8];
    uint32_t version;
    uint32_t header_bytes;
    uint64_t payload_bytes;
    uint64_t checksum;
} io_prefix_t;

uint64_t qgtl_fnv1a(uint64_t h, const void *data, size_t n)
{
    const unsigned char *p = (const unsigned char *)data;
    for (size_t i = 0; i < n; i++) {
        h ^= p[i];
        h *= 0x100000001b3ull;
    }
    return h;
}

#define FNV_OFFSET 0xcbf29ce484222325ull

/* Write prefix + kind header + payload pieces (count pieces of sizes[i] bytes). */
int qgtl_io_write(const char *path, const char magic[8], const void *hdr, uint32_t hdr_bytes, int count,
                  const void *const *pieces, const size_t *sizes)
{
    if (!path) return qgtl_set_error(QGT_ERROR_INVALID_ARGUMENT, "null path");
    io_prefix_t pre;
    memcpy(pre.magic, magic, 8);
    pre.version = QGTL_IO_VERSION;
    pre.header_bytes = hdr_bytes;
    pre.payload_bytes = 0;
    uint64_t h = qgtl_fnv1a(FNV_OFFSET, hdr, hdr_bytes);
    for (int i = 0; i < count; i++) {
        pre.payload_bytes += sizes[i];
        h = qgtl_fnv1a(h, pieces[i], sizes[i]);
    }
    pre.checksum = h;
    FILE *f = fopen(path, "wb");
    if (!f) return qgtl_set_error(QGT_ERROR_IO_ERROR, "cannot open %s for writing", path);
    int ok = fwrite(&pre, sizeof pre, 1, f) == 1 && fwrite(hdr, hdr_bytes, 1, f) == 1;
    for (int i = 0; ok && i < count; i++)
        if (sizes[i] && fwrite(pieces[i], sizes[i], 1, f) != 1) ok = 0;
    if (fclose(f) != 0) ok = 0;
    if (!ok) return qgtl_set_error(QGT_ERROR_IO_ERROR, "write to %s failed", path);
    return 0;
}

/* Read and check the prefix and the kind header (hdr_bytes expected); leaves f at the payload. */
int qgtl_io_open(const char *path, const char magic[8], void *hdr, uint32_t hdr_bytes, uint64_t *payload_bytes,
                 uint64_t *checksum, FILE **out)
{
    if (!path || !out) return qgtl_set_error(QGT_ERROR_INVALID_ARGUMENT, "null argument");
    FILE *f = fopen(path, "rb");
    if (!f) return qgtl_set_error(QGT_ERROR_IO_ERROR, "cannot open %s", path);
    io_prefix_t pre;
    if (fread(&pre, sizeof pre, 1, f) != 1) {
        fclose(f);
        return qgtl_set_error(QGT_ERROR_IO_ERROR, "%s: truncated header", path);
    }
    if (memcmp(pre.magic, magic, 8) != 0) {
        fclose(f);
        return qgtl_set_error(QGT_ERROR_VALIDATION_FAILED, "%s: not a %.8s file", path, magic);
    }
    if (pre.version != QGTL_IO_VERSION || pre.header_bytes != hdr_bytes) {
        fclose(f);
        return qgtl_set_error(QGT_ERROR_INCOMPATIBLE, "%s: version %u / header %u not supported", path, pre.version,
                              pre.header_bytes);
    }
    if (fread(hdr, hdr_bytes, 1, f) != 1) {
        fclose(f);
        return qgtl_set_error(QGT_ERROR_IO_ERROR, "%s: truncated header", path);
    }
    *payload_bytes = pre.payload_bytes;
    *checksum = pre.checksum;
    *out = f;
    return 0;
}

/* Read the payload pieces and check the total size and the checksum (hdr included). */
int qgtl_io_read_payload(FILE *f, const char *path, const void *hdr, uint32_t hdr_bytes, uint64_t payload_bytes,
                         uint64_t checksum, int count, void *const *pieces, const size_t *sizes)
{
    uint64_t want = 0;
    for (int i = 0; i < count; i++) want += sizes[i];
    if (want != payload_bytes) {
        fclose(f);
        return qgtl_set_error(QGT_ERROR_VALIDATION_FAILED, "%s: payload %llu bytes, header implies %llu", path,
                              (unsigned long long)payload_bytes, (unsigned long long)want);
    }
    uint64_t h = qgtl_fnv1a(FNV_OFFSET, hdr, hdr_bytes);
    for (int i = 0; i < count; i++) {
        if (sizes[i] && fread(pieces[i], sizes[i], 1, f) != 1) {
            fclose(f);
            return qgtl_set_error(QGT_ERROR_IO_ERROR, "%s: truncated payload", path);
        }
        h = qgtl_fnv1a(h, pieces[i], sizes[i]);
    }
    const int extra = fgetc(f) != EOF;
    fclose(f);
    if (extra) return qgtl_set_error(QGT_ERROR_VALIDATION_FAILED, "%s: trailing bytes", path);
    if (h != checksum) return qgtl_set_error(QGT_ERROR_VALIDATION_FAILED, "%s: checksum mismatch", path);
    return 0;
}

/* ---- QGT ---- */
typedef struct {
    int32_t n, P, has_theta, reserved;
} qgt_hdr_t;

int qgtl_qgt_save(const char *path, int n, int P, const double *theta, const double *Q)
{
    if (n < 1 || n > QGTL_MAX_QUBITS || P < 1 || !Q)
        return qgtl_set_error(QGT_ERROR_INVALID_ARGUMENT, "qgtl_qgt_save: bad arguments");
    const qgt_hdr_t h = {n, P, theta != NULL, 0};
    const void *pieces[2];
    size_t sizes[2];
    int c = 0;
    if (theta) { pieces[c] = theta; sizes[c++] = sizeof(double) * (size_t)P; }
    pieces[c] = Q;
    sizes[c++] = sizeof(double) * 2 * (size_t)P * (size_t)P;
    return qgtl_io_write(path, "QGTLQGT1", &h, sizeof h, c, pieces, sizes);
}

int qgtl_qgt_file_info(const char *path, int *n, int *P, int *has_theta)
{
    qgt_hdr_t h;
    uint64_t pb = 0, ck = 0;
    FILE *f = NULL;
    int rc = qgtl_io_open(path, "QGTLQGT1", &h, sizeof h, &pb, &ck, &f);
    if (rc) return rc;
    fclose(f);
    if (n) *n = h.n;
    if (P) *P = h.P;
    if (has_theta) *has_theta = h.has_theta;
    return 0;
}

int qgtl_qgt_load(const char *path, int P, double *theta, double *Q)
{
    if (P < 1 || !Q) return qgtl_set_error(QGT_ERROR_INVALID_ARGUMENT, "qgtl_qgt_load: bad arguments");
    qgt_hdr_t h;
    uint64_t pb = 0, ck = 0;
    FILE *f = NULL;
    int rc = qgtl_io_open(path, "QGTLQGT1", &h, sizeof h, &pb, &ck, &f);
    if (rc) return rc;
    if (h.P != P || h.P < 1 || h.n < 1 || h.n > QGTL_MAX_QUBITS) {
        fclose(f);
        return qgtl_set_error(QGT_ERROR_DIMENSION_MISMATCH, "%s holds P = %d, caller expects %d", path, h.P, P);
    }
    double *th = NULL;
    if (h.has_theta) {
        th = theta ? theta : (double *)malloc(sizeof(double) * (size_t)P);
        if (!th) { fclose(f); return qgtl_set_error(QGT_ERROR_NO_MEMORY, "out of host memory"); }
    }
    void *pieces[2];
    size_t sizes[2];
    int c = 0;
    if (h.has_theta) { pieces[c] = th; sizes[c++] = sizeof(double) * (size_t)P; }
    pieces[c] = Q;
    sizes[c++] = sizeof(double) * 2 * (size_t)P * (size_t)P;
    rc = qgtl_io_read_payload(f, path, &h, sizeof h, pb, ck, c, pieces, sizes);
    if (th && th != theta) free(th);
    return rc;
}

/* ---- learned-preconditioner training set ---- */
typedef struct {
    int32_t S, F, P, reserved;
} set_hdr_t;

int qgtl_dataset_save(const char *path, int S, int F, int P, const float *X, const float *As)
{
    if (S < 1 || F < 1 || P < 1 || !X || !As)
        return qgtl_set_error(QGT_ERROR_INVALID_ARGUMENT, "qgtl_dataset_save: bad arguments");
    const set_hdr_t h = {S, F, P, 0};
    const void *pieces[2] = {X, As};
    const size_t sizes[2] = {sizeof(float) * (size_t)S * (size_t)F, sizeof(float) * (size_t)S * (size_t)P * (size_t)P};
    return qgtl_io_write(path, "QGTLSET1", &h, sizeof h, 2, pieces, sizes);
}

int qgtl_dataset_file_info(const char *path, int *S, int *F, int *P)
{
    set_hdr_t h;
    uint64_t pb = 0, ck = 0;
    FILE *f = NULL;
    int rc = qgtl_io_open(path, "QGTLSET1", &h, sizeof h, &pb, &ck, &f);
    if (rc) return rc;
    fclose(f);
    if (S) *S = h.S;
    if (F) *F = h.F;
    if (P) *P = h.P;
    return 0;
}

int qgtl_dataset_load(const char *path, int S, int F, int P, float *X, float *As)
{
    if (S < 1 || F < 1 || P < 1 || !X || !As)
        return qgtl_set_error(QGT_ERROR_INVALID_ARGUMENT, "qgtl_dataset_load: bad arguments");
    set_hdr_t h;
    uint64_t pb = 0, ck = 0;
    FILE *f = NULL;
    int rc = qgtl_io_open(path, "QGTLSET1", &h, sizeof h, &pb, &ck, &f);
    if (rc) return rc;
    if (h.S != S || h.F != F || h.P != P) {
        fclose(f);
        return qgtl_set_error(QGT_ERROR_DIMENSION_MISMATCH, "%s holds S=%d F=%d P=%d", path, h.S, h.F, h.P);
    }
    void *pieces[2] = {X, As};
    const size_t sizes[2] = {sizeof(float) * (size_t)S * (size_t)F, sizeof(float) * (size_t)S * (size_t)P * (size_t)P};
    return qgtl_io_read_payload(f, path, &h, sizeof h, pb, ck, 2, pieces, sizes);
}

/* ---- QGTL's generic tensor container (quantum_geometric_tensor.h:144-147) ----
 * Kind header: type, rank, flags (bit 0 symmetric, 1 unitary, 2 hermitian),
 * hardware, mem_type, dimension, num_spins, total_elements; payload: the rank
 * dimensions (uint64) and total_elements components (ComplexFloat). */
typedef struct {
    int32_t type, flags, hardware, mem_type;
    uint64_t rank, dimension, num_spins, total_elements;
} tensor_hdr_t;

#define QGTL_TENSOR_MAX_RANK 64

qgt_error_t geometric_tensor_save(const quantum_geometric_tensor_t *t, const char *filename)
{
    if (!t || !filename || t->rank == 0 || t->rank > QGTL_TENSOR_MAX_RANK || !t->dimensions)
        return (qgt_error_t)qgtl_set_error(QGT_ERROR_INVALID_ARGUMENT, "geometric_tensor_save: bad arguments");
    size_t total = 1;
    for (size_t i = 0; i < t->rank; i++) {
        if (t->dimensions[i] == 0 || total > ((size_t)1 << 40) / t->dimensions[i])
            return (qgt_error_t)qgtl_set_error(QGT_ERROR_INVALID_DIMENSION, "geometric_tensor_save: bad dimensions");
        total *= t->dimensions[i];
    }
    /* QGTL's geometric_tensor_create sets total_elements to the product of the dimensions */
    if (t->total_elements != total || !t->components)
        return (qgt_error_t)qgtl_set_error(QGT_ERROR_INVALID_STATE, "geometric_tensor_save: components do not match "
                                           "the dimensions");
    const tensor_hdr_t h = {(int32_t)t->type, (t->is_symmetric ? 1 : 0) | (t->is_unitary ? 2 : 0) |
                                                  (t->is_hermitian ? 4 : 0),
                            (int32_t)t->hardware, (int32_t)t->mem_type, t->rank, t->dimension, t->num_spins, total};
    uint64_t dims[QGTL_TENSOR_MAX_RANK];
    for (size_t i = 0; i < t->rank; i++) dims[i] = t->dimensions[i];
    const void *pieces[2] = {dims, t->components};
    const size_t sizes[2] = {sizeof(uint64_t) * t->rank, sizeof(ComplexFloat) * total};
    const int rc = qgtl_io_write(filename, "QGTLTEN1", &h, sizeof h, 2, pieces, sizes);
    return (qgt_error_t)rc;
}

qgt_error_t geometric_tensor_load(quantum_geometric_tensor_t **tensor, const char *filename)
{
    if (!tensor || !filename)
        return (qgt_error_t)qgtl_set_error(QGT_ERROR_INVALID_ARGUMENT, "geometric_tensor_load: bad arguments");
    *tensor = NULL;
    tensor_hdr_t h;
    uint64_t pb = 0, ck = 0;
    FILE *f = NULL;
    int rc = qgtl_io_open(filename, "QGTLTEN1", &h, sizeof h, &pb, &ck, &f);
    if (rc) return (qgt_error_t)rc;
    if (h.rank == 0 || h.rank > QGTL_TENSOR_MAX_RANK || h.total_elements == 0 ||
        h.total_elements > ((uint64_t)1 << 40) ||
        pb != sizeof(uint64_t) * h.rank + sizeof(ComplexFloat) * h.total_elements) {
        fclose(f);
        return (qgt_error_t)qgtl_set_error(QGT_ERROR_VALIDATION_FAILED, "%s: inconsistent tensor header", filename);
    }
    const size_t total = (size_t)h.total_elements;
    const size_t aligned = (total + 7) & ~(size_t)7;   /* QGTL's create pads to 8 elements */
    quantum_geometric_tensor_t *t = (quantum_geometric_tensor_t *)aligned_alloc(64, (sizeof *t + 63) / 64 * 64);
    uint64_t *dims64 = (uint64_t *)malloc(sizeof(uint64_t) * h.rank);
    size_t *dims = (size_t *)malloc(sizeof(size_t) * h.rank);
    ComplexFloat *comps = (ComplexFloat *)aligned_alloc(64, (aligned * sizeof(ComplexFloat) + 63) / 64 * 64);
    if (!t || !dims64 || !dims || !comps) {
        fclose(f);
        free(t); free(dims64); free(dims); free(comps);
        return (qgt_error_t)qgtl_set_error(QGT_ERROR_NO_MEMORY, "out of host memory");
    }
    void *pieces[2] = {dims64, comps};
    const size_t sizes[2] = {sizeof(uint64_t) * h.rank, sizeof(ComplexFloat) * total};
    rc = qgtl_io_read_payload(f, filename, &h, sizeof h, pb, ck, 2, pieces, sizes);
    size_t prod = 1;
    for (uint64_t i = 0; !rc && i < h.rank; i++) {
        dims[i] = (size_t)dims64[i];
        if (dims[i] == 0 || prod > total / dims[i]) rc = QGT_ERROR_VALIDATION_FAILED;
        else prod *= dims[i];
    }
    if (!rc && prod != total) rc = QGT_ERROR_VALIDATION_FAILED;
    free(dims64);
    if (rc) {
        free(t); free(dims); free(comps);
        if (rc == QGT_ERROR_VALIDATION_FAILED) qgtl_set_error(rc, "%s: dimensions do not match the components", filename);
        return (qgt_error_t)rc;
    }
    for (size_t i = total; i < aligned; i++) comps[i] = (ComplexFloat){0.f, 0.f};
    memset(t, 0, sizeof *t);
    t->type = (geometric_tensor_type_t)h.type;
    t->dimensions = dims;
    t->rank = (size_t)h.rank;
    t->dimension = (size_t)h.dimension;
    t->components = comps;
    t->num_spins = (size_t)h.num_spins;
    t->auxiliary_data = NULL;
    t->is_symmetric = (h.flags & 1) != 0;
    t->is_unitary = (h.flags & 2) != 0;
    t->is_hermitian = (h.flags & 4) != 0;
    t->hardware = (HardwareType)h.hardware;
    t->mem_type = (qgt_memory_type_t)h.mem_type;
    t->total_elements = total;
    t->aligned_elements = aligned;
    *tensor = t;
    return QGT_SUCCESS;
}
