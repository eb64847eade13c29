// Shared definitions for the HIP translation units of libqgtl_mi355x.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "qgtl_hip_abi.h"

namespace qgtl {

hipStream_t stream();
int set_error(hipError_t e, const char *what);
int set_error_msg(int code, const char *msg);

// Grouped launches: the item whose unit range holds u, i.e. the largest i
// with items[i].unit_begin <= u (items sorted, items[0].unit_begin = 0).
template <typename Item>
__device__ __forceinline__ int find_item(const Item *__restrict__ items, int nitems, uint32_t u) {
    int lo = 0, hi = nitems - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (items[mid].unit_begin <= u) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// A value every lane of the wave holds (e.g. loaded from a grouped launch's item
// table) moved into scalar registers, so the code using it stays scalar.
__device__ __forceinline__ uint32_t uniform(uint32_t v) { return (uint32_t)__builtin_amdgcn_readfirstlane((int)v); }
__device__ __forceinline__ int32_t uniform(int32_t v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ __forceinline__ uint64_t uniform(uint64_t v) {
    return ((uint64_t)uniform((uint32_t)(v >> 32)) << 32) | uniform((uint32_t)v);
}

// A device address held as an integer (grouped launch item tables) as a pointer
// the compiler knows to be global memory, so accesses through it become
// global_load/store (scalar base + 32-bit offsets) rather than flat ones.
template <class T>
__device__ __forceinline__ T *global_ptr(uint64_t a) {
    typedef __attribute__((address_space(1))) T gT;
    return (T *)(gT *)a;
}

}  // namespace qgtl

// Map a HIP status to the library's error codes and remember the message.
#define QGTL_HIP_CHECK(call)                                                   \
    do {                                                                       \
        hipError_t _e = (call);                                                \
        if (_e != hipSuccess) return qgtl::set_error(_e, #call);               \
    } while (0)

#define QGTL_LAUNCH_CHECK(what)                                                \
    do {                                                                       \
        hipError_t _e = hipGetLastError();                                     \
        if (_e != hipSuccess) return qgtl::set_error(_e, what);                \
    } while (0)
