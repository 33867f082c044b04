// Shared device helpers of the pyramid kernels (pyramid.hip, fused.hip).
#pragma once

#include "enbl_device.h"

namespace enbl {

// Two adjacent samples (2j, 2j+1) of a row as one vector access; `both`
// false loads/stores the first only.  Callers keep rows 2-element aligned.
// fetch() always reads the whole pair and leaves it packed (Pair); split()
// unpacks it where it is used, so a fetch issued ahead of its use stays in
// flight.  fetch() at the last odd column reads one element past the row,
// which pyramid planes hold as pitch padding.
template <class P> struct Vec2;
template <> struct Vec2<int16_t> {
    typedef uint32_t Pair;
    __device__ static __forceinline__ Pair fetch(const int16_t* p) { return *gmem(reinterpret_cast<const uint32_t*>(p)); }
    __device__ static __forceinline__ void split(Pair v, int32_t& a, int32_t& b) {
        a = (int16_t)(v & 0xffffu);
        b = (int16_t)(v >> 16);
    }
    __device__ static __forceinline__ void load(const int16_t* p, bool both, int32_t& a, int32_t& b) {
        if (both) {
            const uint32_t v = *gmem(reinterpret_cast<const uint32_t*>(p));
            a = (int16_t)(v & 0xffffu);
            b = (int16_t)(v >> 16);
        } else {
            a = *gmem(p);
            b = 0;
        }
    }
    __device__ static __forceinline__ void store(int16_t* p, bool both, int16_t a, int16_t b) {
        if (both) {
            *gmem(reinterpret_cast<uint32_t*>(p)) = (uint32_t)(uint16_t)a | ((uint32_t)(uint16_t)b << 16);
        } else {
            *gmem(p) = a;
        }
    }
};
template <> struct Vec2<int32_t> {
    typedef v2i32 Pair;
    __device__ static __forceinline__ Pair fetch(const int32_t* p) { return *gmem(reinterpret_cast<const v2i32*>(p)); }
    __device__ static __forceinline__ void split(Pair v, int32_t& a, int32_t& b) {
        a = v.x;
        b = v.y;
    }
    __device__ static __forceinline__ void load(const int32_t* p, bool both, int32_t& a, int32_t& b) {
        if (both) {
            const v2i32 v = *gmem(reinterpret_cast<const v2i32*>(p));
            a = v.x;
            b = v.y;
        } else {
            a = *gmem(p);
            b = 0;
        }
    }
    __device__ static __forceinline__ void store(int32_t* p, bool both, int32_t a, int32_t b) {
        if (both) {
            *gmem(reinterpret_cast<v2i32*>(p)) = v2i32{a, b};
        } else {
            *gmem(p) = a;
        }
    }
};
template <> struct Vec2<double> {
    typedef v2f64 Pair;
    __device__ static __forceinline__ Pair fetch(const double* p) { return *gmem(reinterpret_cast<const v2f64*>(p)); }
    __device__ static __forceinline__ void split(Pair v, double& a, double& b) {
        a = v.x;
        b = v.y;
    }
    __device__ static __forceinline__ void load(const double* p, bool both, double& a, double& b) {
        if (both) {
            const v2f64 v = *gmem(reinterpret_cast<const v2f64*>(p));
            a = v.x;
            b = v.y;
        } else {
            a = *gmem(p);
            b = 0.0;
        }
    }
    __device__ static __forceinline__ void store(double* p, bool both, double a, double b) {
        if (both) {
            *gmem(reinterpret_cast<v2f64*>(p)) = v2f64{a, b};
        } else {
            *gmem(p) = a;
        }
    }
};

__device__ __forceinline__ int wrap_index(int k, int n) {
    k %= n;
    return k < 0 ? k + n : k;
}

// Horizontal weight sum behind the reference's SKIPSM_* scale constants for
// destination column X (src/pyramid.h:1199-1411).  A 1-pixel-wide source
// always takes the no-wraparound constants.
__device__ __forceinline__ int expand_hsum(int X, int sw, bool wrap, bool dst_even) {
    const int j = X >> 1;
    if (sw == 1) return (X & 1) ? 4 : 6;
    if ((X & 1) == 0) {
        if (!wrap) return 6 + (j > 0 ? 1 : 0) + (j < sw - 1 ? 1 : 0);
        if (dst_even) return 8;
        return 6 + (j == 0 ? 4 : 1) + (j == sw - 1 ? 4 : 1);
    }
    if (!wrap) return 4 + (j < sw - 1 ? 4 : 0);
    return 8;
}

__device__ __forceinline__ int expand_vsum(int Y, int sh) {
    const int i = Y >> 1;
    if ((Y & 1) == 0) return 6 + (i > 0 ? 1 : 0) + (i < sh - 1 ? 1 : 0);
    return 1 + (i < sh - 1 ? 1 : 0);
}


}  // namespace enbl
