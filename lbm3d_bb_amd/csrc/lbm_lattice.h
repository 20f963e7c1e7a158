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
struct StoreF64 {
    using T = double;
    using C = double;
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
};

struct StoreF32 {
    using T = float;
    using C = float;
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
};

struct StoreF16S {
    using T = __half;
    using C = float;
    static constexpr int kBytes = 2;
    static constexpr int kVec = 2;
    __device__ static inline C load(const T* p, int d, const Shift& sh) {
        return __fmaf_rn(__half2float(*p), sh.ik, weight<float>(d));
    }
    // direction d1 if `second`, else d0 (both compile-time: the reference
    // value is a select, not an indexed read)
    __device__ static inline C load_pick(const T* p, int d0, int d1, bool second, const Shift& sh) {
        return __fmaf_rn(__half2float(*p), sh.ik, second ? weight<float>(d1) : weight<float>(d0));
    }
    // __float2half_rn: round to nearest even (not the round-toward-zero of the
    // packed v_cvt_pkrtz form, which would bias the deviations).
    __device__ static inline void store(T* p, int d, C v, const Shift& sh) {
        const __half h = __float2half_rn((v - weight<float>(d)) * sh.k);
        __builtin_nontemporal_store(__half_as_ushort(h), reinterpret_cast<unsigned short*>(p));
    }
    __device__ static inline void load2(const T* p, int d, C& a, C& b, const Shift& sh) {
        uint32_t v;
        __builtin_memcpy(&v, p, 4);
        a = __fmaf_rn(__half2float(__ushort_as_half((unsigned short)(v & 0xffffu))), sh.ik, weight<float>(d));
        b = __fmaf_rn(__half2float(__ushort_as_half((unsigned short)(v >> 16))), sh.ik, weight<float>(d));
    }
    __device__ static inline void store2(T* p, int d, C a, C b, const Shift& sh) {
        const uint32_t lo = __half_as_ushort(__float2half_rn((a - weight<float>(d)) * sh.k));
        const uint32_t hi = __half_as_ushort(__float2half_rn((b - weight<float>(d)) * sh.k));
        __builtin_nontemporal_store(lo | (hi << 16), reinterpret_cast<uint32_t*>(p));
    }
};

struct StoreBF16S {
    using T = uint16_t;
    using C = float;
    static constexpr int kBytes = 2;
    static constexpr int kVec = 2;
    __device__ static inline float bf16_to_f32(uint16_t h) { return __uint_as_float(uint32_t(h) << 16); }
    __device__ static inline uint16_t f32_to_bf16_rne(float f) {
        uint32_t u = __float_as_uint(f);
        if ((u & 0x7fffffffu) > 0x7f800000u) return uint16_t((u >> 16) | 0x40u);  // quiet NaN
        u += 0x7fffu + ((u >> 16) & 1u);
        return uint16_t(u >> 16);
    }
    __device__ static inline C load(const T* p, int d, const Shift& sh) { return bf16_to_f32(*p) + weight<float>(d); }
    __device__ static inline C load_pick(const T* p, int d0, int d1, bool second, const Shift& sh) {
        return bf16_to_f32(*p) + (second ? weight<float>(d1) : weight<float>(d0));
    }
    __device__ static inline void store(T* p, int d, C v, const Shift& sh) {
        __builtin_nontemporal_store(f32_to_bf16_rne(v - weight<float>(d)), p);
    }
    __device__ static inline void load2(const T* p, int d, C& a, C& b, const Shift& sh) {
        uint32_t v;
        __builtin_memcpy(&v, p, 4);
        a = __uint_as_float(v << 16) + weight<float>(d);
        b = __uint_as_float(v & 0xffff0000u) + weight<float>(d);
    }
    __device__ static inline void store2(T* p, int d, C a, C b, const Shift& sh) {
        const uint32_t lo = f32_to_bf16_rne(a - weight<float>(d));
        const uint32_t hi = f32_to_bf16_rne(b - weight<float>(d));
        __builtin_nontemporal_store(lo | (hi << 16), reinterpret_cast<uint32_t*>(p));
    }
};

}  // namespace lbm
