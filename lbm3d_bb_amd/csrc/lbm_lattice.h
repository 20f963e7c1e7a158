// lbm_lattice.h -- D3Q19 tables and population storage formats for gfx950.
//
// Direction order is the reference's (src/Utilities.jl:950-960) minus one, so a
// population index d here is the reference's d+1 and bbOpp maps to kOpp.
#pragma once

#include <hip/hip_runtime.h>
#include <hip/hip_fp16.h>
#include <stdint.h>

#include "lbm_types.h"

namespace lbm {

constexpr int Q = 19;
constexpr int kCx[Q] = {1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 0, 0, 1, -1, 1, -1, 0, 0, 0};
constexpr int kCy[Q] = {0, 0, 1, -1, 0, 0, 1, -1, 0, 0, 1, -1, -1, 1, 0, 0, 1, -1, 0};
constexpr int kCz[Q] = {0, 0, 0, 0, 1, -1, 0, 0, 1, -1, 1, -1, 0, 0, -1, 1, -1, 1, 0};
constexpr int kOpp[Q] = {1, 0, 3, 2, 5, 4, 7, 6, 9, 8, 11, 10, 13, 12, 15, 14, 17, 16, 18};

// c_a(d) for a runtime direction d (loops that are deliberately not unrolled):
// component + 1 packed in two bits per direction, so no table is read.
constexpr uint64_t kCPack[3] = {0x1522522552ull, 0x1258252525ull, 0x1885225255ull};
__host__ __device__ inline int c_at(int a, int d) { return (int)((kCPack[a] >> (2 * d)) & 3u) - 1; }

// Weights in the compute type, each the correctly rounded value of 1/18, 1/36 or
// 1/3 (the reference stores Float32.(1//18) etc., src/Utilities.jl:954-958).
template <typename C>
__host__ __device__ constexpr C weight(int d) {
    return d < 6 ? C(1) / C(18) : (d < 18 ? C(1) / C(36) : C(1) / C(3));
}

// Sums of the FP32 weights the shifted formats store against: rho_w = sum_d w_d
// and s2 = sum_d w_d c_dx^2, 1 and 1/3 for the exact weights; collide_dev uses
// them so that its equilibrium has exactly the cell's mass and momentum.
constexpr double kRhoW19 = 6.0 * (double)(1.0f / 18.0f) + 12.0 * (double)(1.0f / 36.0f) + (double)(1.0f / 3.0f);
constexpr double kS2_19 = 2.0 * (double)(1.0f / 18.0f) + 8.0 * (double)(1.0f / 36.0f);

// ---------------------------------------------------------------------------
// Stochastic rounding of the 16-bit formats' step stores.  A step's increment
// of a population can be far below half an ulp of its 16-bit deviation (a body
// force g adds 3 w_d g ~ 2e-7 per step at g = 1e-6, while half an FP16 ulp of
// a deviation of 1e-2 is ~2e-6): rounding to nearest then drops it every step
// and the flow stops accelerating.  Rounding up with a probability equal to
// the dropped fraction of an ulp keeps every increment in expectation.  The
// random bits come from a hash of the cell index and the step's seed (one per
// cell, or per pair of cells), spread over the directions by one multiply each,
// so a run stays bitwise reproducible.  (Starts and copies keep
// round-to-nearest.)
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint32_t sr_cell(int64_t cell, uint32_t seed) {
    uint32_t x = (uint32_t)cell ^ seed;
    x ^= x >> 16;
    x *= 0x7feb352du;
    x ^= x >> 15;
    x *= 0x846ca68bu;
    x ^= x >> 16;
    return x;
}

// the random word of direction d from the cell's hash h (d is a compile-time
// constant in the unrolled store loops, so the multiplier folds)
__device__ __forceinline__ uint32_t sr_dir(uint32_t h, int d) {
    uint32_t x = h * (0x9e3779b9u * (2u * (uint32_t)d + 1u) | 1u);
    return x ^ (x >> 15);
}

// FP16: s + a uniform fraction of an FP16 ulp, truncated: the 13 mantissa
// bits FP16 drops from FP32 get random bits added before the round-toward-zero
// conversion (v_cvt_pkrtz_f16_f32), so s rounds up with probability equal to
// its dropped fraction.  Two values, 13 random bits each from r and r >> 16.
// (gfx950's v_cvt_sr_f16_f32, two per pair, measured 2 % slower on the FP16
// step than these adds and the one packed conversion.)
__device__ __forceinline__ uint32_t f16x2_sr(float a, float b, uint32_t r) {
    const float ta = __uint_as_float(__float_as_uint(a) + (r & 0x1fffu));
    const float tb = __uint_as_float(__float_as_uint(b) + ((r >> 16) & 0x1fffu));
    const auto h = __builtin_amdgcn_cvt_pkrtz(ta, tb);
    uint32_t out;
    __builtin_memcpy(&out, &h, 4);
    return out;
}

// BF16: gfx950's v_cvt_sr_bf16_f32 adds a field of the random word r (its
// high 16 bits) to the 16 bits BF16 drops and truncates, so a value rounds
// away from zero with probability equal to its dropped fraction of an ulp
// (measured on the device: the rate matches the fraction for hashed words).
// One instruction per value, where the integer form (NaN test, add, shift)
// took three: the BF16 step measured 3 % faster.  Two values take r and r
// rotated by 16 bits, so their fields are disjoint.  bf16_sr_lo converts a
// into the low half (the high half: old's), bf16_sr_hi b into the high half
// of lo (the low half kept).
using bf16x2_t = __bf16 __attribute__((ext_vector_type(2)));

__device__ __forceinline__ uint32_t bf16_sr_lo(uint32_t old, float a, uint32_t r) {
    return __builtin_bit_cast(uint32_t,
                              __builtin_amdgcn_cvt_sr_bf16_f32(__builtin_bit_cast(bf16x2_t, old), a, r, false));
}
__device__ __forceinline__ uint32_t bf16_sr_hi(uint32_t lo, float b, uint32_t r) {
    return __builtin_bit_cast(uint32_t, __builtin_amdgcn_cvt_sr_bf16_f32(__builtin_bit_cast(bf16x2_t, lo), b, r, true));
}
__device__ __forceinline__ uint32_t bf16x2_sr(float a, float b, uint32_t r) {
    return bf16_sr_hi(bf16_sr_lo(r, a, r), b, __builtin_amdgcn_alignbit(r, r, 16));
}

// ---------------------------------------------------------------------------
// Storage formats.  Each exposes the stored type T, the compute type C and
// load/store of one population.  The shifted 16-bit formats keep
// (f_d - w_d) k: the deviation from the rest state at rho0 = 1, scaled by a
// power of two (Shift), so the 16-bit mantissa spends its bits on the
// deviation instead of on w_d itself and small deviations stay out of FP16's
// subnormal range.
// ---------------------------------------------------------------------------
//
// kVec is the number of consecutive x cells one thread of the bulk pass owns:
// 2 for the 16-bit formats, so each direction moves 4 bytes per lane (the
// +-1 shifted pulls are unaligned 4-byte loads, which gfx950 serves
// directly); 1 for FP32 and FP64.  load2/store2 move the pair (p[0], p[1]).
// kDev: the populations live in registers as deviations delta_d = f_d - w_d
// (the shifted formats), which the step kernels collide in that form.
struct StoreF64 {
    using T = double;
    using C = double;
    static constexpr bool kDev = false;
    static constexpr int kBytes = 8;
    static constexpr int kVec = 1;
    __device__ static inline C load(const T* p, int, const Shift&) { return *p; }
    __device__ static inline C load_pick(const T* p, int, int, bool, const Shift&) { return *p; }
    __device__ static inline void store(T* p, int, C v, const Shift&) { *p = v; }
    __device__ static inline void load2(const T* p, int d, C& a, C& b, const Shift& sh) {
        a = load(p, d, sh);
        b = load(p + 1, d, sh);
    }
    __device__ static inline void store2(T* p, int d, C a, C b, const Shift& sh) {
        store(p, d, a, sh);
        store(p + 1, d, b, sh);
    }
    __device__ static inline void store_sr(T* p, int d, C v, const Shift& sh, uint32_t) { store(p, d, v, sh); }
    __device__ static inline void store2_sr(T* p, int d, C a, C b, const Shift& sh, uint32_t) {
        store2(p, d, a, b, sh);
    }
};

struct StoreF32 {
    using T = float;
    using C = float;
    static constexpr bool kDev = false;
    static constexpr int kBytes = 4;
    static constexpr int kVec = 1;
    __device__ static inline C load(const T* p, int, const Shift&) { return *p; }
    __device__ static inline C load_pick(const T* p, int, int, bool, const Shift&) { return *p; }
    // populations are written once per step and read by the next: stream them
    __device__ static inline void store(T* p, int, C v, const Shift&) { __builtin_nontemporal_store(v, p); }
    __device__ static inline void load2(const T* p, int d, C& a, C& b, const Shift& sh) {
        a = load(p, d, sh);
        b = load(p + 1, d, sh);
    }
    __device__ static inline void store2(T* p, int d, C a, C b, const Shift& sh) {
        store(p, d, a, sh);
        store(p + 1, d, b, sh);
    }
    __device__ static inline void store_sr(T* p, int d, C v, const Shift& sh, uint32_t) { store(p, d, v, sh); }
    __device__ static inline void store2_sr(T* p, int d, C a, C b, const Shift& sh, uint32_t) {
        store2(p, d, a, b, sh);
    }
};

struct StoreF16S {
    using T = __half;
    using C = float;
    static constexpr bool kDev = true;
    static constexpr int kBytes = 2;
    static constexpr int kVec = 2;
    // loads return the deviation delta = f - w (exact: a power-of-two scale)
    __device__ static inline C load(const T* p, int, const Shift& sh) { return __half2float(*p) * sh.ik; }
    __device__ static inline C load_pick(const T* p, int, int, bool, const Shift& sh) {
        return __half2float(*p) * sh.ik;
    }
    // __float2half_rn: round to nearest even (not the round-toward-zero of the
    // packed v_cvt_pkrtz form, which would bias the deviations).
    __device__ static inline void store(T* p, int d, C v, const Shift& sh) {
        const __half h = __float2half_rn((v - weight<float>(d)) * sh.k);
        __builtin_nontemporal_store(__half_as_ushort(h), reinterpret_cast<unsigned short*>(p));
    }
    __device__ static inline void load2(const T* p, int, C& a, C& b, const Shift& sh) {
        uint32_t v;
        __builtin_memcpy(&v, p, 4);
        a = __half2float(__ushort_as_half((unsigned short)(v & 0xffffu))) * sh.ik;
        b = __half2float(__ushort_as_half((unsigned short)(v >> 16))) * sh.ik;
    }
    __device__ static inline void store2(T* p, int d, C a, C b, const Shift& sh) {
        const uint32_t lo = __half_as_ushort(__float2half_rn((a - weight<float>(d)) * sh.k));
        const uint32_t hi = __half_as_ushort(__float2half_rn((b - weight<float>(d)) * sh.k));
        __builtin_nontemporal_store(lo | (hi << 16), reinterpret_cast<uint32_t*>(p));
    }
    // step stores take the deviation itself (what the step kernels hold)
    __device__ static inline void store_sr(T* p, int, C dv, const Shift& sh, uint32_t r) {
        const uint32_t h = f16x2_sr(dv * sh.k, 0.0f, r);
        __builtin_nontemporal_store((unsigned short)(h & 0xffffu), reinterpret_cast<unsigned short*>(p));
    }
    __device__ static inline void store2_sr(T* p, int, C a, C b, const Shift& sh, uint32_t r) {
        const uint32_t h = f16x2_sr(a * sh.k, b * sh.k, r);
        __builtin_nontemporal_store(h, reinterpret_cast<uint32_t*>(p));
    }
};

struct StoreBF16S {
    using T = uint16_t;
    using C = float;
    static constexpr bool kDev = true;
    static constexpr int kBytes = 2;
    static constexpr int kVec = 2;
    __device__ static inline float bf16_to_f32(uint16_t h) { return __uint_as_float(uint32_t(h) << 16); }
    __device__ static inline uint16_t f32_to_bf16_rne(float f) {
        uint32_t u = __float_as_uint(f);
        if ((u & 0x7fffffffu) > 0x7f800000u) return uint16_t((u >> 16) | 0x40u);  // quiet NaN
        u += 0x7fffu + ((u >> 16) & 1u);
        return uint16_t(u >> 16);
    }
    __device__ static inline C load(const T* p, int, const Shift&) { return bf16_to_f32(*p); }
    __device__ static inline C load_pick(const T* p, int, int, bool, const Shift&) { return bf16_to_f32(*p); }
    __device__ static inline void store(T* p, int d, C v, const Shift& sh) {
        __builtin_nontemporal_store(f32_to_bf16_rne(v - weight<float>(d)), p);
    }
    __device__ static inline void load2(const T* p, int, C& a, C& b, const Shift&) {
        uint32_t v;
        __builtin_memcpy(&v, p, 4);
        a = __uint_as_float(v << 16);
        b = __uint_as_float(v & 0xffff0000u);
    }
    __device__ static inline void store2(T* p, int d, C a, C b, const Shift& sh) {
        const uint32_t lo = f32_to_bf16_rne(a - weight<float>(d));
        const uint32_t hi = f32_to_bf16_rne(b - weight<float>(d));
        __builtin_nontemporal_store(lo | (hi << 16), reinterpret_cast<uint32_t*>(p));
    }
    __device__ static inline void store_sr(T* p, int, C dv, const Shift&, uint32_t r) {
        __builtin_nontemporal_store((uint16_t)(bf16_sr_lo(r, dv, r) & 0xffffu), p);
    }
    __device__ static inline void store2_sr(T* p, int, C a, C b, const Shift&, uint32_t r) {
        __builtin_nontemporal_store(bf16x2_sr(a, b, r), reinterpret_cast<uint32_t*>(p));
    }
};

}  // namespace lbm
