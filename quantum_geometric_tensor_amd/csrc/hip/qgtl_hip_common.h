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

// Two fp16 pieces of sa = 2^15 x and sb: hi = the value rounded to 11
// significant bits (nearest, ties away from zero: add half an ulp to the bit
// pattern, mask the 13 low mantissa bits), exact in fp16 for |value| >= 2^-14;
// lo = fp16(value - hi), the difference being exact in fp32.  So
// |value - hi - lo| <= 2^-11 |value - hi| <= 2^-22 |value|, the bound of two
// round-to-nearest fp16 conversions, at about 4 VALU operations per value (each
// pair's two conversions are one v_cvt_pk_f16_f32).  Returns the packed pairs
// (sa in the low half).
typedef float qgtl_f32x2 __attribute__((ext_vector_type(2)));
typedef _Float16 qgtl_f16x2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ void split_f16_pair(float sa, float sb, uint32_t &hi, uint32_t &lo) {
    const float ra = __uint_as_float((__float_as_uint(sa) + 0x1000u) & 0xffffe000u);
    const float rb = __uint_as_float((__float_as_uint(sb) + 0x1000u) & 0xffffe000u);
    const qgtl_f32x2 h32 = {ra, rb}, l32 = {sa - ra, sb - rb};
    hi = __builtin_bit_cast(uint32_t, __builtin_convertvector(h32, qgtl_f16x2));
    lo = __builtin_bit_cast(uint32_t, __builtin_convertvector(l32, qgtl_f16x2));
}

constexpr float kSplitScale = 32768.0f;          // 2^15: |x| <= 1 keeps 2^15 x below the fp16 maximum

// A device address held as an integer (grouped launch item tables) as a pointer
// the compiler knows to be global memory, so accesses through it become
// global_load/store (scalar base + 32-bit offsets) rather than flat ones.
template <class T>
__device__ __forceinline__ T *global_ptr(uint64_t a) {
    typedef __attribute__((address_space(1))) T gT;
    return (T *)(gT *)a;
}

// The same for data that no kernel writes while the launch runs (descriptor
// tables written before it): the constant address space, so uniform reads
// through it become scalar loads into SGPRs.
template <class T>
__device__ __forceinline__ const T *const_ptr(uint64_t a) {
    typedef __attribute__((address_space(4))) const T cT;
    return (const T *)(cT *)a;
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
