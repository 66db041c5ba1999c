// Shared host helpers of the HIP library: the per-thread error message, launch checks
// and the curl-term table of the Yee update.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <string>

#include "../../include/tfdtd.h"

// The global-memory view of a device pointer.  Pointers that kernels read from descriptor tables
// in device memory are generic, and generic loads and stores compile to flat instructions (which
// also count against the LDS wait counter); reading them through a global pointer gives global
// instructions.  Kernel-argument pointers are inferred global already.
template <typename T>
__device__ __forceinline__ __attribute__((address_space(1))) T* global(T* p) {
    return (__attribute__((address_space(1))) T*)p;
}

namespace tfdtd_detail {
inline thread_local std::string last_error;
}

inline int tfdtd_fail(int code, const char* fmt, ...) {
    char buffer[512];
    va_list args;
    va_start(args, fmt);
    vsnprintf(buffer, sizeof buffer, fmt, args);
    va_end(args);
    tfdtd_detail::last_error = buffer;
    return code;
}

inline int tfdtd_check_launch(const char* what) {
    hipError_t status = hipGetLastError();
    if (status != hipSuccess) return tfdtd_fail(TFDTD_EHIP, "%s: %s", what, hipGetErrorString(status));
    return TFDTD_OK;
}

// Curl terms in the reference's order (torchfdtd/boundaries.py:21-23):
// t : (derivative axis, input component, output component, sign)
// 0 : (0, 2, 1, -1)   1 : (0, 1, 2, +1)   2 : (1, 2, 0, +1)
// 3 : (1, 0, 2, -1)   4 : (2, 1, 0, -1)   5 : (2, 0, 1, +1)
__host__ __device__ constexpr int term_axis(int t) { return t / 2; }
__host__ __device__ constexpr int term_in(int t) { return (t == 0 || t == 2) ? 2 : (t == 1 || t == 4) ? 1 : 0; }
__host__ __device__ constexpr int term_out(int t) { return (t == 2 || t == 4) ? 0 : (t == 0 || t == 5) ? 1 : 2; }
__host__ __device__ constexpr int term_sign(int t) { return (t == 0 || t == 3 || t == 4) ? -1 : 1; }

// Place CPML segment k (target rows [lo, hi) of the term differentiating `comp` along
// `axis`) in slot 2*term or 2*term+1.  slot_lo/slot_hi describe the filled slots
// (lo == hi: empty).  Returns the slot, or -1 with the error recorded.
inline int tfdtd_assign_slot(int axis, int comp, int lo, int hi, const int dims[3], int k, int slot_lo[12], int slot_hi[12]) {
    int term = -1;
    for (int t = 0; t < 6; ++t)
        if (term_axis(t) == axis && term_in(t) == comp) term = t;
    if (term < 0) {
        tfdtd_fail(TFDTD_EINVAL, "segment %d: no curl term differentiates component %d along axis %d", k, comp, axis);
        return -1;
    }
    if (lo < 0 || hi <= lo || hi > dims[axis]) {
        tfdtd_fail(TFDTD_EINVAL, "segment %d: rows [%d,%d) outside the axis", k, lo, hi);
        return -1;
    }
    int slot = 2 * term;
    if (slot_hi[slot] != slot_lo[slot]) {
        ++slot;
        if (slot_hi[slot] != slot_lo[slot]) {
            tfdtd_fail(TFDTD_EINVAL, "segment %d: more than two slabs for one curl term", k);
            return -1;
        }
        if (lo < slot_hi[slot - 1] && slot_lo[slot - 1] < hi) {
            tfdtd_fail(TFDTD_EINVAL, "segment %d overlaps another slab of its term", k);
            return -1;
        }
    }
    slot_lo[slot] = lo;
    slot_hi[slot] = hi;
    return slot;
}
