// Device-side arithmetic contract shared by all enblend/enfuse kernels.
//
// The pyramid types follow src/numerictraits.h:168-189 of the reference:
//   Int16 pyramid / Int32 SKIPSM accumulator (8-bit images),
//   Int32 pyramid / Int32 SKIPSM accumulator (16-bit images, masks),
//   double pyramid / double accumulator (float images).
// Integer accumulators wrap (two's complement) exactly like the reference's
// Int32 arithmetic on x86-64; double operations keep the reference's
// evaluation order and every file is compiled with -ffp-contract=off, so the
// double path reproduces the reference's rounding step by step.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "logexp.h"

namespace enbl {

// Every buffer the kernels touch is device (global) memory.  Loads and stores
// through these views compile to global instructions (SGPR base + lane
// offset) even where the pointer itself was read from a job table, instead of
// flat instructions, whose completion the scalar-load waits also cover.
template <class T> __device__ __forceinline__ const __attribute__((address_space(1))) T* gmem(const T* p) {
    return (const __attribute__((address_space(1))) T*)p;
}
template <class T> __device__ __forceinline__ __attribute__((address_space(1))) T* gmem(T* p) {
    return (__attribute__((address_space(1))) T*)p;
}

// Raw buffer accesses through a per-row resource: a lane whose byte offset
// is kBufSkip (or any offset at or past the resource's size) reads zeros and
// its store is dropped, with no branch and no memory traffic.  Kernels use
// them where every iteration of a loop must issue the same memory
// instructions (vmcnt waits then count exactly and never wait for stores).
constexpr uint32_t kBufSkip = 0x80000000u;
__device__ __forceinline__ __amdgpu_buffer_rsrc_t buf_rsrc(const void* base, uint32_t bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), 0, (int)bytes, 0x00020000);
}

// Native vectors for the wide accesses (HIP's int2/uint3 classes cannot be
// read through an address-space-qualified pointer).
typedef int32_t v2i32 __attribute__((ext_vector_type(2)));
typedef int32_t v3i32 __attribute__((ext_vector_type(3)));
typedef int32_t v4i32 __attribute__((ext_vector_type(4)));
typedef double v2f64 __attribute__((ext_vector_type(2)));

constexpr int kWave = 64;
// Lanes 1..62 of a wave own output columns; lanes 0 and 63 only supply the
// horizontal halo through cross-lane shuffles.
constexpr int kOwned = kWave - 2;

struct AccI {
    typedef int32_t T;
    __device__ static __forceinline__ T add(T a, T b) { return (T)((uint32_t)a + (uint32_t)b); }
    __device__ static __forceinline__ T sub(T a, T b) { return (T)((uint32_t)a - (uint32_t)b); }
    __device__ static __forceinline__ T mul(T a, int k) { return (T)((uint32_t)a * (uint32_t)k); }
    // C++ truncating division by a positive divisor.  The filter weight sums
    // are powers of two away from borders and alpha edges (256, 64, 16, ...):
    // those take a shift with the round-toward-zero correction instead of the
    // general division sequence.
    __device__ static __forceinline__ T div(T a, int d) {
        if ((d & (d - 1)) == 0) {
            const int s = __builtin_ctz((unsigned)d);
            return (a + ((a >> 31) & (d - 1))) >> s;
        }
        return a / d;
    }
    __device__ static __forceinline__ T zero() { return 0; }
};

struct AccD {
    typedef double T;
    __device__ static __forceinline__ T add(T a, T b) { return a + b; }
    __device__ static __forceinline__ T sub(T a, T b) { return a - b; }
    __device__ static __forceinline__ T mul(T a, int k) { return a * (double)k; }
    __device__ static __forceinline__ T div(T a, int d) { return a / (double)d; }
    __device__ static __forceinline__ T zero() { return 0.0; }
};

template <class P> struct Pix;
template <> struct Pix<int16_t> {
    typedef AccI A;
    typedef int32_t S;
    __device__ static __forceinline__ S load(int16_t v) { return v; }
    // Store of an int into Int16 through VIGRA's accessor: static_cast (wrap).
    __device__ static __forceinline__ int16_t wrap(S v) { return (int16_t)(uint16_t)(uint32_t)v; }
    // NumericTraits<Int16>::fromPromote(int): clamp.
    __device__ static __forceinline__ int16_t sat(S v) { return (int16_t)min(32767, max(-32768, v)); }
};
template <> struct Pix<int32_t> {
    typedef AccI A;
    typedef int32_t S;
    __device__ static __forceinline__ S load(int32_t v) { return v; }
    __device__ static __forceinline__ int32_t wrap(S v) { return v; }
    __device__ static __forceinline__ int32_t sat(S v) { return v; }
};
template <> struct Pix<double> {
    typedef AccD A;
    typedef double S;
    __device__ static __forceinline__ S load(double v) { return v; }
    __device__ static __forceinline__ double wrap(S v) { return v; }
    __device__ static __forceinline__ double sat(S v) { return v; }
};

// NumericTraits<T>::fromRealPromote: v -/+ 0.5 then truncate, clamped to T.
// NaN follows the reference as built for x86-64 (cvttsd2si): INT_MIN for a
// 32-bit target, 0 for a 16-bit target (low half of INT_MIN).
template <class T> __device__ T from_real(double v);
// Evaluated without branches: t = v -/+ 0.5 as in the reference, then
// clamped to the target range (a t beyond the range belongs to a v beyond it,
// and truncation inside the range is unchanged by the clamp), then truncated.
template <> __device__ __forceinline__ int32_t from_real<int32_t>(double v) {
    const double t = v < 0.0 ? v - 0.5 : v + 0.5;
    const int32_t r = (int32_t)fmin(fmax(t, -2147483648.0), 2147483647.0);
    return v != v ? INT32_MIN : r;
}
template <> __device__ __forceinline__ int16_t from_real<int16_t>(double v) {
    const double t = v < 0.0 ? v - 0.5 : v + 0.5;
    const int32_t r = (int32_t)fmin(fmax(t, -32768.0), 32767.0);
    return v != v ? (int16_t)0 : (int16_t)r;
}
template <> __device__ __forceinline__ double from_real<double>(double v) { return v; }

// Cross-lane moves within a wave: lane_up gives lane L the value of lane
// L-1, lane_down the value of lane L+1, as DPP wave shifts (wave_shr:1,
// wave_shl:1) folded into a VALU move instead of an LDS permute.  The lane
// without a source (0 for lane_up, 63 for lane_down) receives 0; the kernels
// only read these values in halo lanes.  64-bit values move as two halves.
__device__ __forceinline__ int32_t lane_up(int32_t v) { return __builtin_amdgcn_update_dpp(0, v, 0x138, 0xf, 0xf, false); }
__device__ __forceinline__ int32_t lane_down(int32_t v) { return __builtin_amdgcn_update_dpp(0, v, 0x130, 0xf, 0xf, false); }
__device__ __forceinline__ double lane_up(double v) {
    const uint64_t b = __double_as_longlong(v);
    const uint32_t lo = (uint32_t)lane_up((int32_t)(uint32_t)b), hi = (uint32_t)lane_up((int32_t)(uint32_t)(b >> 32));
    return __longlong_as_double((long long)(((uint64_t)hi << 32) | lo));
}
__device__ __forceinline__ double lane_down(double v) {
    const uint64_t b = __double_as_longlong(v);
    const uint32_t lo = (uint32_t)lane_down((int32_t)(uint32_t)b), hi = (uint32_t)lane_down((int32_t)(uint32_t)(b >> 32));
    return __longlong_as_double((long long)(((uint64_t)hi << 32) | lo));
}

// The correctly rounded quotient v / 255.0 for 0 <= v < 2^24 (every value a
// normalised fixed-point mask pyramid holds): one reciprocal product and one
// fused residual correction, equal to the division on that whole range
// (checked exhaustively by the test suite).
__device__ __forceinline__ double quot255(int32_t v) {
    const double x = (double)v;
    const double r = 1.0 / 255.0;
    const double q0 = x * r;
    const double e = __fma_rn(-q0, 255.0, x);
    return __fma_rn(e, r, q0);
}

// Counter-based dither draw: one 32-bit hash per pixel keyed by (seed, x, y)
// (two rounds of the murmur3 32-bit finaliser); channel c of the pixel takes
// bits [10c, 10c + frac - 2] as its offset on the reference's integer grid
// (r & half_m1) + quarter (src/fixmath.h:246-264).  frac <= 11, c <= 2.
__device__ __forceinline__ uint32_t fmix32(uint32_t h) {
    h ^= h >> 16;
    h *= 0x85EBCA6Bu;
    h ^= h >> 13;
    h *= 0xC2B2AE35u;
    return h ^ (h >> 16);
}
__device__ __forceinline__ uint32_t dither_hash(uint64_t seed, int x, int y) {
    const uint32_t s = (uint32_t)seed ^ (uint32_t)(seed >> 32);
    return fmix32(s ^ fmix32(((uint32_t)y * 0x9E3779B1u) ^ (uint32_t)x));
}
__device__ __forceinline__ int dither_draw(uint32_t h, int c, int frac) {
    return (int)((h >> (10 * c)) & ((1u << (frac - 1)) - 1u)) + (1 << (frac - 2));
}
__device__ __forceinline__ int dither_rand(uint64_t seed, int x, int y, int c, int frac) {
    return dither_draw(dither_hash(seed, x, y), c, frac);
}

}  // namespace enbl
