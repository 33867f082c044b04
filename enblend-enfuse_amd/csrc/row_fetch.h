// Raw-row fetches of the streaming passes (fused.hip, reduce_stream.h):
// buffer-resource loads and stores, the row sources of a pass and the
// wave-uniform helpers.  A fetch issues its loads and keeps the words as
// loaded; the caller unpacks them where they are used, so loads issued a
// step ahead stay in flight meanwhile.
#pragma once

#include <type_traits>

#include "enbl_device.h"
#include "pyr_common.h"

namespace enbl {

// ---- buffer accesses ------------------------------------------------------
template <class T> __device__ __forceinline__ T bload(__amdgpu_buffer_rsrc_t r, uint32_t off);
template <> __device__ __forceinline__ uint32_t bload<uint32_t>(__amdgpu_buffer_rsrc_t r, uint32_t off) {
    return (uint32_t)__builtin_amdgcn_raw_buffer_load_b32(r, off, 0, 0);
}
template <> __device__ __forceinline__ v2i32 bload<v2i32>(__amdgpu_buffer_rsrc_t r, uint32_t off) {
    return __builtin_bit_cast(v2i32, __builtin_amdgcn_raw_buffer_load_b64(r, off, 0, 0));
}
template <> __device__ __forceinline__ v3i32 bload<v3i32>(__amdgpu_buffer_rsrc_t r, uint32_t off) {
    return __builtin_bit_cast(v3i32, __builtin_amdgcn_raw_buffer_load_b96(r, off, 0, 0));
}
template <> __device__ __forceinline__ v2f64 bload<v2f64>(__amdgpu_buffer_rsrc_t r, uint32_t off) {
    return __builtin_bit_cast(v2f64, __builtin_amdgcn_raw_buffer_load_b128(r, off, 0, 0));
}
__device__ __forceinline__ uint32_t bload_u16(__amdgpu_buffer_rsrc_t r, uint32_t off) {
    return (uint32_t)__builtin_amdgcn_raw_buffer_load_b16(r, off, 0, 0);
}
__device__ __forceinline__ uint32_t bload_u8(__amdgpu_buffer_rsrc_t r, uint32_t off) {
    return (uint32_t)__builtin_amdgcn_raw_buffer_load_b8(r, off, 0, 0);
}

template <class P> struct BPair;   // pair (2j, 2j+1) and single-sample buffer stores
template <> struct BPair<int16_t> {
    typedef uint32_t Pair;
    __device__ static __forceinline__ Pair fetch(__amdgpu_buffer_rsrc_t r, uint32_t off) { return bload<uint32_t>(r, off); }
    __device__ static __forceinline__ void store2(__amdgpu_buffer_rsrc_t r, uint32_t off, int16_t x, int16_t y) {
        __builtin_amdgcn_raw_buffer_store_b32((uint32_t)(uint16_t)x | ((uint32_t)(uint16_t)y << 16), r, off, 0, 0);
    }
    __device__ static __forceinline__ void store1(__amdgpu_buffer_rsrc_t r, uint32_t off, int16_t x) {
        __builtin_amdgcn_raw_buffer_store_b16((unsigned short)x, r, off, 0, 0);
    }
};
template <> struct BPair<int32_t> {
    typedef v2i32 Pair;
    __device__ static __forceinline__ Pair fetch(__amdgpu_buffer_rsrc_t r, uint32_t off) { return bload<v2i32>(r, off); }
    __device__ static __forceinline__ void store2(__amdgpu_buffer_rsrc_t r, uint32_t off, int32_t x, int32_t y) {
        typedef uint32_t u2 __attribute__((ext_vector_type(2)));
        __builtin_amdgcn_raw_buffer_store_b64(u2{(uint32_t)x, (uint32_t)y}, r, off, 0, 0);
    }
    __device__ static __forceinline__ void store1(__amdgpu_buffer_rsrc_t r, uint32_t off, int32_t x) {
        __builtin_amdgcn_raw_buffer_store_b32((uint32_t)x, r, off, 0, 0);
    }
};
template <> struct BPair<double> {
    typedef v2f64 Pair;
    __device__ static __forceinline__ Pair fetch(__amdgpu_buffer_rsrc_t r, uint32_t off) { return bload<v2f64>(r, off); }
    __device__ static __forceinline__ void store2(__amdgpu_buffer_rsrc_t r, uint32_t off, double x, double y) {
        typedef uint32_t u4 __attribute__((ext_vector_type(4)));
        __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(u4, v2f64{x, y}), r, off, 0, 0);
    }
    __device__ static __forceinline__ void store1(__amdgpu_buffer_rsrc_t r, uint32_t off, double x) {
        typedef uint32_t u2 __attribute__((ext_vector_type(2)));
        __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u2, x), r, off, 0, 0);
    }
};

__device__ __forceinline__ uint32_t round_up2(uint32_t v) { return (v + 1u) & ~1u; }
__device__ __forceinline__ uint32_t round_up4(uint32_t v) { return (v + 3u) & ~3u; }

// Column facts of the pixel pair (2j, 2j+1) a lane handles.
struct LaneCols {
    bool valid;   // j inside the coarse level (else the lane reads zeros)
    bool both;    // 2j+1 inside the fine level
    bool tailw;   // wave-uniform: some lane has valid && !both
    int c;        // 2j
};

// ---- sources of the fine level -------------------------------------------
// fetch() issues the loads of fine row r for this lane's pixel pair through
// a buffer resource over that row (row_ok false: a zero-size resource, so the
// same instructions issue and read nothing) and keeps the words as loaded;
// unpack() converts them where the row is used, an iteration later.  A pair
// whose second pixel is past the row end reads that pixel as padding of the
// row's last aligned word (never past the word holding the row's last byte)
// or not at all.

// Pyramid planes (pitch padded to an even element count, so the pair at the
// last odd column reads padding for its second element).
template <class P, int C> struct PlaneSrc {
    typedef typename Pix<P>::S S;
    typedef BPair<P> B;
    struct Raw { typename B::Pair v[C]; };
    template <class Args>
    __device__ static __forceinline__ void fetch(const Args& a, int fw, int r, bool row_ok, const LaneCols& L, Raw& R) {
        const P* row = static_cast<const P*>(a.src) + (long)r * a.spitch;
        const uint32_t bytes = row_ok ? round_up2((uint32_t)fw) * (uint32_t)sizeof(P) : 0u;
        const uint32_t off = L.valid ? (uint32_t)L.c * (uint32_t)sizeof(P) : kBufSkip;
#pragma unroll
        for (int ch = 0; ch < C; ++ch) R.v[ch] = B::fetch(buf_rsrc(row + ch * a.splane, bytes), off);
    }
    template <class Args>
    __device__ static __forceinline__ void unpack(const Args&, const Raw& R, S (&x0)[C], S (&x1)[C]) {
#pragma unroll
        for (int ch = 0; ch < C; ++ch) Vec2<P>::split(R.v[ch], x0[ch], x1[ch]);
    }
};

// Interleaved integer image, v << frac (src/fixmath.h:197-202).  WIDE: the
// rows are aligned for whole-word pair loads (4 bytes for 16-bit samples, 2
// bytes for one 8-bit channel); otherwise one load per sample.
template <class I, class P, int C, bool WIDE> struct IntImageSrc {
    typedef typename Pix<P>::S S;
    static constexpr bool RGB16 = WIDE && sizeof(I) == 2 && C == 3;   // 12 bytes = 3 words per pair
    static constexpr bool ONE = WIDE && C == 1;                       // 2 or 4 bytes per pair
    static constexpr int NW = RGB16 ? 3 : (ONE ? 1 : 2 * C);
    struct Raw { uint32_t w[NW]; };
    template <class Args>
    __device__ static __forceinline__ void fetch(const Args& a, int fw, int r, bool row_ok, const LaneCols& L, Raw& R) {
        const I* row = static_cast<const I*>(a.src) + (long)r * a.spitch;
        const uint32_t rb = (uint32_t)fw * C * (uint32_t)sizeof(I);
        const uint32_t off = L.valid ? (uint32_t)L.c * C * (uint32_t)sizeof(I) : kBufSkip;
        if constexpr (RGB16) {
            const __amdgpu_buffer_rsrc_t rs = buf_rsrc(row, row_ok ? round_up4(rb) : 0u);
            if (L.tailw) {
                // a last-column pair has its 6 bytes in words 0-1; word 2 is past the row
                const v2i32 lo = bload<v2i32>(rs, off);
                R.w[0] = (uint32_t)lo.x;
                R.w[1] = (uint32_t)lo.y;
                R.w[2] = bload<uint32_t>(rs, L.both ? off + 8u : kBufSkip);
            } else {
                const v3i32 v = bload<v3i32>(rs, off);
                R.w[0] = (uint32_t)v.x;
                R.w[1] = (uint32_t)v.y;
                R.w[2] = (uint32_t)v.z;
            }
        } else if constexpr (ONE && sizeof(I) == 2) {
            R.w[0] = bload<uint32_t>(buf_rsrc(row, row_ok ? round_up4(rb) : 0u), off);
        } else if constexpr (ONE) {
            R.w[0] = bload_u16(buf_rsrc(row, row_ok ? round_up2(rb) : 0u), off);
        } else {
            const __amdgpu_buffer_rsrc_t rs = buf_rsrc(row, row_ok ? rb : 0u);
            const uint32_t off1 = L.valid && L.both ? off + C * (uint32_t)sizeof(I) : kBufSkip;
#pragma unroll
            for (int k = 0; k < C; ++k) {
                if constexpr (sizeof(I) == 2) {
                    R.w[k] = bload_u16(rs, off + 2u * k);
                    R.w[C + k] = bload_u16(rs, off1 + 2u * k);
                } else {
                    R.w[k] = bload_u8(rs, off + k);
                    R.w[C + k] = bload_u8(rs, off1 + k);
                }
            }
        }
    }
    template <class Args>
    __device__ static __forceinline__ void unpack(const Args& a, const Raw& R, S (&x0)[C], S (&x1)[C]) {
        uint32_t v[2 * C];
        if constexpr (RGB16) {
            v[0] = R.w[0] & 0xffffu; v[1] = R.w[0] >> 16;
            v[2] = R.w[1] & 0xffffu; v[3] = R.w[1] >> 16;
            v[4] = R.w[2] & 0xffffu; v[5] = R.w[2] >> 16;
        } else if constexpr (ONE) {
            constexpr int B = 8 * sizeof(I);
            v[0] = R.w[0] & ((1u << B) - 1u);
            v[1] = (R.w[0] >> B) & ((1u << B) - 1u);
        } else {
#pragma unroll
            for (int k = 0; k < 2 * C; ++k) v[k] = R.w[k];
        }
#pragma unroll
        for (int ch = 0; ch < C; ++ch) {
            x0[ch] = (S)(P)((P)v[ch] << a.frac);
            x1[ch] = (S)(P)((P)v[C + ch] << a.frac);
        }
    }
};

// Mask level 0 formed from the image's weight plane (src, float, pitch
// spitch) and the norm plane: the normalisation (src/enfuse.h:1529-1534) with
// the fixed-point conversion (src/fixmath.h:188-194), as normalize_kernel
// computes it; hard masks are converted as they are.
__device__ __forceinline__ double norm_mask_value(float mv, float nv, float fallback, int hard, double scale)
{
    const float v = hard ? mv : nv > 0.0 ? (255.0f * mv) / nv : fallback;
    return (double)v * scale;
}
template <class M> struct NormMaskSrc {
    typedef int32_t S;
    struct Raw { v2i32 w, n; };
    template <class Args>
    __device__ static __forceinline__ void fetch(const Args& a, int fw, int r, bool row_ok, const LaneCols& L, Raw& R) {
        const NormReduceArgs& na = static_cast<const NormReduceArgs&>(a);
        const float* wrow = static_cast<const float*>(a.src) + (long)r * a.spitch;
        const float* nrow = na.norm + (long)r * na.npitch;
        const uint32_t bytes = row_ok ? round_up2((uint32_t)fw) * 4u : 0u;
        const uint32_t off = L.valid ? (uint32_t)L.c * 4u : kBufSkip;
        R.w = bload<v2i32>(buf_rsrc(wrow, bytes), off);
        R.n = bload<v2i32>(buf_rsrc(nrow, bytes), off);
    }
    template <class Args>
    __device__ static __forceinline__ void unpack(const Args& a, const Raw& R, S (&x0)[1], S (&x1)[1]) {
        const NormReduceArgs& na = static_cast<const NormReduceArgs&>(a);
        x0[0] = (S)from_real<M>(norm_mask_value(__int_as_float(R.w.x), __int_as_float(R.n.x), na.fallback, na.hard, na.scale));
        x1[0] = (S)from_real<M>(norm_mask_value(__int_as_float(R.w.y), __int_as_float(R.n.y), na.fallback, na.hard, na.scale));
    }
};

// Interleaved float image, 1 + log1p(v) or 1 / (1 - v) (src/fixmath.h:165-186).
template <int C> struct FloatImageSrc {
    typedef double S;
    struct Raw { float f[2 * C]; };
    __device__ static __forceinline__ double conv(float f) {
        const double v = (double)f;
        return float_to_log_pyramid(v);
    }
    template <class Args>
    __device__ static __forceinline__ void fetch(const Args& a, int fw, int r, bool row_ok, const LaneCols& L, Raw& R) {
        const float* row = static_cast<const float*>(a.src) + (long)r * a.spitch;
        const __amdgpu_buffer_rsrc_t rs = buf_rsrc(row, row_ok ? (uint32_t)fw * C * 4u : 0u);
        const uint32_t off = L.valid ? (uint32_t)L.c * C * 4u : kBufSkip;
        const uint32_t off1 = L.valid && L.both ? off + C * 4u : kBufSkip;
        if constexpr (C == 3) {
            const v3i32 p0 = bload<v3i32>(rs, off), p1 = bload<v3i32>(rs, off1);
            R.f[0] = __int_as_float(p0.x); R.f[1] = __int_as_float(p0.y); R.f[2] = __int_as_float(p0.z);
            R.f[3] = __int_as_float(p1.x); R.f[4] = __int_as_float(p1.y); R.f[5] = __int_as_float(p1.z);
        } else {
#pragma unroll
            for (int k = 0; k < C; ++k) {
                R.f[k] = __uint_as_float(bload<uint32_t>(rs, off + 4u * k));
                R.f[C + k] = __uint_as_float(bload<uint32_t>(rs, off1 + 4u * k));
            }
        }
    }
    template <class Args>
    __device__ static __forceinline__ void unpack(const Args&, const Raw& R, S (&x0)[C], S (&x1)[C]) {
#pragma unroll
        for (int ch = 0; ch < C; ++ch) {
            x0[ch] = conv(R.f[ch]);
            x1[ch] = conv(R.f[C + ch]);
        }
    }
};

// The alpha bytes of a pair: one 16-bit load when the rows are 2-byte
// aligned (the byte after a last-column pair shares its 2-byte word),
// otherwise one load per byte.
template <bool WIDE> struct AlphaPair {
    uint32_t w[WIDE ? 1 : 2];
    __device__ __forceinline__ void fetch(const uint8_t* row, bool row_ok, int fw, const LaneCols& L) {
        if constexpr (WIDE) {
            w[0] = bload_u16(buf_rsrc(row, row_ok ? round_up2((uint32_t)fw) : 0u), L.valid ? (uint32_t)L.c : kBufSkip);
        } else {
            const __amdgpu_buffer_rsrc_t rs = buf_rsrc(row, row_ok ? (uint32_t)fw : 0u);
            w[0] = bload_u8(rs, L.valid ? (uint32_t)L.c : kBufSkip);
            w[1] = bload_u8(rs, L.valid && L.both ? (uint32_t)L.c + 1u : kBufSkip);
        }
    }
    __device__ __forceinline__ bool first() const { return WIDE ? (w[0] & 0xffu) != 0 : w[0] != 0; }
    __device__ __forceinline__ bool second() const { return WIDE ? (w[0] >> 8) != 0 : w[1] != 0; }
};

// C++ truncating division of an integer accumulator by 2^s.
__device__ __forceinline__ int32_t div_pow2(int32_t a, int s) { return (a + ((a >> 31) & ((1 << s) - 1))) >> s; }

__device__ __forceinline__ bool wave_all(bool p) { return __builtin_amdgcn_ballot_w64(!p) == 0; }

template <int N> using Phase = std::integral_constant<int, N>;

}  // namespace enbl
