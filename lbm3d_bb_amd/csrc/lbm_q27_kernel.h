// lbm_q27_kernel.h -- the D3Q27 step kernel and its launch templates, shared by
// lbm_q27.hip (simple bounce-back) and lbm_q27_ibb.hip (interpolated).
//
// The D3Q27 lattice (src/Utilities.jl:962-976): the same time
// step as the D3Q19 path (pull stream of the previous post-collision state,
// faces with precedence, moving bounce-back with the wall-velocity
// correction, fresh-node refill, momentum exchange, collision of the next
// step) with the reference's D3Q27 velocities, weights and opposite table.
//
// One thread per cell, one launch per step.  A cell away from the non-periodic
// velocity / anti-bounce-back faces and from every body (the usual case) only
// pulls its 27 populations and collides.  A cell with a face rule or a solid
// neighbour also reads its own 27 populations for the previous-step moments,
// and its rule directions load the cell's own f*[opp(d)] in place of the pull
// (the address is selected, the load is unconditional).  FP64, FP32 and
// shifted 16-bit storage.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>

#include "lbm_device.h"
#include "lbm_q27.h"

namespace lbm {
namespace q27 {

constexpr int QQ = kQ;

// bbOpp (src/Utilities.jl:976)
__host__ __device__ constexpr int opp(int d) { return d == QQ - 1 ? d : (d ^ 1); }

// w (src/Utilities.jl:967-973): 8/27 rest, 2/27 face, 1/54 edge, 1/216 corner
template <typename C>
__device__ inline C w27(int d) {
    const int n = (kX[d] != 0) + (kY[d] != 0) + (kZ[d] != 0);
    return n == 0 ? C(8) / C(27) : (n == 1 ? C(2) / C(27) : (n == 2 ? C(1) / C(54) : C(1) / C(216)));
}

// The random word of direction d for the 16-bit D3Q27 stores, from the cell's
// hash h (sr_cell): a 24-bit multiply by a per-direction odd constant (full
// rate, where a 32-bit multiply runs at a quarter) and a shift-xor that brings
// product bits 15-31 into the low bits the FP16 rounding takes (BF16 takes the
// high bits).
__device__ __forceinline__ uint32_t sr_dir27(uint32_t h, int d) {
    const uint32_t x = __umul24(h, ((0x9e3779u * (2u * (uint32_t)d + 1u)) & 0xffffffu) | 1u);
    return x ^ (x >> 15);
}

// 16-bit population storage on D3Q27: the deviation from the rest state
// f_d - w_d (rho0 = 1) with this lattice's weights, times 2^12 in FP16 or as is
// in BF16; the step kernels hold and collide the deviations (kDev), loads
// return them and step stores take them (stochastic rounding, lbm_lattice.h);
// the start-state store takes f and rounds to nearest.
struct StoreF16S27 {
    using T = __half;
    using C = float;
    static constexpr bool kDev = true;
    __device__ static inline C load(const T* p, int, const Shift& sh) { return __half2float(*p) * sh.ik; }
    __device__ static inline void store(T* p, int d, C v, const Shift& sh) {
        const __half h = __float2half_rn((v - w27<float>(d)) * sh.k);
        __builtin_nontemporal_store(__half_as_ushort(h), reinterpret_cast<unsigned short*>(p));
    }
    __device__ static inline void store_sr(T* p, int, C dv, const Shift& sh, uint32_t r) {
        __builtin_nontemporal_store((unsigned short)sr_bits(dv, sh, r), reinterpret_cast<unsigned short*>(p));
    }
    // the 16 bits store_sr writes, and the deviation they hold
    __device__ static inline uint32_t sr_bits(C dv, const Shift& sh, uint32_t r) {
        return f16x2_sr(dv * sh.k, 0.0f, r) & 0xffffu;
    }
    __device__ static inline C dec(uint32_t h, const Shift& sh) {
        return __half2float(__ushort_as_half((unsigned short)(h & 0xffffu))) * sh.ik;
    }
    // a two-cell word: prep gives cell a's rounding input (the scaled
    // deviation with its random bits added, as f16x2_sr), pack_sr adds cell
    // b's and converts both (low half: a); each half equals what store_sr
    // writes with the same random word
    __device__ static inline uint32_t prep(C dv, const Shift& sh, uint32_t r) {
        return __float_as_uint(dv * sh.k) + (r & 0x1fffu);
    }
    __device__ static inline uint32_t pack_sr(uint32_t a, C dv, const Shift& sh, uint32_t r) {
        const auto h = __builtin_amdgcn_cvt_pkrtz(__uint_as_float(a), __uint_as_float(prep(dv, sh, r)));
        uint32_t out;
        __builtin_memcpy(&out, &h, 4);
        return out;
    }
};

struct StoreBF16S27 {
    using T = uint16_t;
    using C = float;
    static constexpr bool kDev = true;
    __device__ static inline C load(const T* p, int, const Shift&) { return StoreBF16S::bf16_to_f32(*p); }
    __device__ static inline void store(T* p, int d, C v, const Shift&) {
        __builtin_nontemporal_store(StoreBF16S::f32_to_bf16_rne(v - w27<float>(d)), p);
    }
    __device__ static inline void store_sr(T* p, int, C dv, const Shift&, uint32_t r) {
        __builtin_nontemporal_store((uint16_t)(bf16_sr_lo(r, dv, r) & 0xffffu), p);
    }
    __device__ static inline uint32_t sr_bits(C dv, const Shift&, uint32_t r) { return bf16_sr_lo(r, dv, r) & 0xffffu; }
    __device__ static inline C dec(uint32_t h, const Shift&) { return __uint_as_float(h << 16); }
    __device__ static inline uint32_t prep(C dv, const Shift&, uint32_t r) { return bf16_sr_lo(r, dv, r); }
    __device__ static inline uint32_t pack_sr(uint32_t lo, C dv, const Shift&, uint32_t r) {
        return bf16_sr_hi(lo, dv, r);
    }
};

// Sums of the FP32 D3Q27 weights (rho_w = sum w_d, s2 = sum w_d c_dx^2; 1 and
// 1/3 for the exact weights), as for D3Q19 (collide_dev, lbm_device.h).
constexpr double kRhoW27 = (double)(8.0f / 27.0f) + 6.0 * (double)(2.0f / 27.0f) + 12.0 * (double)(1.0f / 54.0f) +
                           8.0 * (double)(1.0f / 216.0f);
constexpr double kS2_27 = 2.0 * (double)(2.0f / 27.0f) + 8.0 * (double)(1.0f / 54.0f) + 8.0 * (double)(1.0f / 216.0f);

// rest level of direction d in the register form of S (w_d for deviations)
template <class S>
__device__ inline double rest27(int d) {
    if constexpr (S::kDev) return (double)w27<float>(d);
    return 0.0;
}

template <typename C>
__device__ inline void moments27(const C f[QQ], C& rho, C& jx, C& jy, C& jz) {
    rho = C(0);
    jx = jy = jz = C(0);
#pragma unroll
    for (int d = 0; d < QQ; ++d) {
        rho += f[d];
        if (kX[d] > 0) jx += f[d];
        if (kX[d] < 0) jx -= f[d];
        if (kY[d] > 0) jy += f[d];
        if (kY[d] < 0) jy -= f[d];
        if (kZ[d] > 0) jz += f[d];
        if (kZ[d] < 0) jz -= f[d];
    }
}

// Previous-step rho, u of a cell from its post-collision populations: the
// collision keeps rho and adds F to the momentum, so u = j*/rho - g/2.  For the
// shifted formats rho is the sum of the deviations and rho_out receives
// rho - 1, what their rules take.
template <class S, bool FORCED>
__device__ inline void prev_moments(typename S::C rho, typename S::C jx, typename S::C jy, typename S::C jz,
                                    const Phys<typename S::C>& p, typename S::C& rho_out, typename S::C u[3],
                                    typename S::C& uu2) {
    using C = typename S::C;
    if constexpr (S::kDev) {
        rho_out = C(kRhoW27 - 1.0) + rho;
        rho = C(kRhoW27) + rho;
    } else {
        rho_out = rho;
    }
    u[0] = jx / rho;
    u[1] = jy / rho;
    u[2] = jz / rho;
    if (FORCED) {
        u[0] -= C(0.5) * p.g[0];
        u[1] -= C(0.5) * p.g[1];
        u[2] -= C(0.5) * p.g[2];
    }
    uu2 = u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
}

// BGK (src/Collision.jl:232-240) / TRT (src/Collision.jl:136-170) with the Guo
// source, as for D3Q19: u = (sum f c + F/2) / rho, S split into even and odd
// parts scaled by (1 - omega_pm / 2).
template <typename C, int COLL, bool FORCED>
__device__ inline void collide27(C f[QQ], const Phys<C>& p) {
    C rho, jx, jy, jz;
    moments27<C>(f, rho, jx, jy, jz);
    const C inv = C(1) / rho;
    C ux = jx * inv, uy = jy * inv, uz = jz * inv;
    C Fx = C(0), Fy = C(0), Fz = C(0), uF = C(0);
    if (FORCED) {
        ux += C(0.5) * p.g[0];
        uy += C(0.5) * p.g[1];
        uz += C(0.5) * p.g[2];
        Fx = rho * p.g[0];
        Fy = rho * p.g[1];
        Fz = rho * p.g[2];
        uF = ux * Fx + uy * Fy + uz * Fz;
    }
    const C u2 = ux * ux + uy * uy + uz * uz;
    const C wp = p.wp;
    if (COLL == LBM_TRT) {
        const C wm = p.wm;
        const C kp = C(1) - C(0.5) * wp, km = C(1) - C(0.5) * wm;
#pragma unroll
        for (int d = 0; d < QQ - 1; d += 2) {
            const C w = w27<C>(d);
            const C cu = C(kX[d]) * ux + C(kY[d]) * uy + C(kZ[d]) * uz;
            const C feq_p = w * rho * (C(1) + C(4.5) * cu * cu - C(1.5) * u2);
            const C feq_m = w * rho * (C(3) * cu);
            const C fp = C(0.5) * (f[d] + f[d + 1]);
            const C fm = C(0.5) * (f[d] - f[d + 1]);
            C fpn = fp - wp * (fp - feq_p);
            C fmn = fm - wm * (fm - feq_m);
            if (FORCED) {
                const C cF = C(kX[d]) * Fx + C(kY[d]) * Fy + C(kZ[d]) * Fz;
                fpn += kp * (w * (C(9) * cu * cF - C(3) * uF));
                fmn += km * (w * (C(3) * cF));
            }
            f[d] = fpn + fmn;
            f[d + 1] = fpn - fmn;
        }
        const C w0 = w27<C>(QQ - 1);
        const C feq0 = w0 * rho * (C(1) - C(1.5) * u2);
        C r = f[QQ - 1] - wp * (f[QQ - 1] - feq0);
        if (FORCED) r += kp * (w0 * (-C(3) * uF));
        f[QQ - 1] = r;
    } else {
        const C kp = C(1) - C(0.5) * wp;
#pragma unroll
        for (int d = 0; d < QQ; ++d) {
            const C w = w27<C>(d);
            const C cu = C(kX[d]) * ux + C(kY[d]) * uy + C(kZ[d]) * uz;
            const C feq = w * rho * (C(1) + C(3) * cu + C(4.5) * cu * cu - C(1.5) * u2);
            C r = f[d] - wp * (f[d] - feq);
            if (FORCED) {
                const C cF = C(kX[d]) * Fx + C(kY[d]) * Fy + C(kZ[d]) * Fz;
                r += kp * (w * (C(3) * cF + C(9) * cu * cF - C(3) * uF));
            }
            f[d] = r;
        }
    }
}

// collide27 on deviations delta = f - w for the shifted formats: the same
// normalised equilibrium and source as collide_dev (lbm_device.h) with the
// D3Q27 weights, so the collision keeps the cell's mass and momentum exactly.
template <int COLL, bool FORCED>
__device__ inline void collide27_dev(float f[QQ], const Phys<float>& p) {
    constexpr float rw = (float)kRhoW27, irw = (float)(1.0 / kRhoW27);
    constexpr float c3 = (float)(1.0 / kS2_27), a2 = (float)(1.5 * kRhoW27 / kS2_27),
                    a9 = (float)(3.0 * kRhoW27 / kS2_27);
    float sd, jx, jy, jz;
    moments27<float>(f, sd, jx, jy, jz);
    const float rho = rw + sd;
    const float rm = sd * irw, rh = 1.0f + rm;
    const float inv = 1.0f / rho;
    float ux = jx * inv, uy = jy * inv, uz = jz * inv;
    float Fx = 0.f, Fy = 0.f, Fz = 0.f, uF = 0.f;
    if (FORCED) {
        ux += 0.5f * p.g[0];
        uy += 0.5f * p.g[1];
        uz += 0.5f * p.g[2];
        Fx = rho * p.g[0];
        Fy = rho * p.g[1];
        Fz = rho * p.g[2];
        uF = ux * Fx + uy * Fy + uz * Fz;
    }
    const float u2 = ux * ux + uy * uy + uz * uz;
    const float mx = rho * ux, my = rho * uy, mz = rho * uz;
    const float wp = p.wp;
    const float kp = 1.0f - 0.5f * wp;
    if (COLL == LBM_TRT) {
        const float wm = p.wm;
        const float km = 1.0f - 0.5f * wm;
        // The pair (a, b) = (f_d, f_opp) relaxes its symmetric part with wp and
        // its antisymmetric part with wm:
        //   f_d = A1 a + B1 b + (P + M),  f_opp = B1 a + A1 b + (P - M),
        //   A1 = 1 - (wp + wm) / 2,  B1 = (wm - wp) / 2,
        //   P = wp dp (+ kp Sp),  M = wm dm (+ km Sm),
        // the same update as ((a+b)/2 relaxed to dp) +- ((a-b)/2 relaxed to dm).
        const float A1 = 1.0f - 0.5f * (wp + wm), B1 = 0.5f * (wm - wp);
#pragma unroll
        for (int d = 0; d < QQ - 1; d += 2) {
            const float w = w27<float>(d);
            const float cu = float(kX[d]) * ux + float(kY[d]) * uy + float(kZ[d]) * uz;
            const float cm = float(kX[d]) * mx + float(kY[d]) * my + float(kZ[d]) * mz;
            float P = (wp * w) * (rm + rh * (a2 * cu * cu - 1.5f * u2));
            float M = (wm * w * c3) * cm;
            if (FORCED) {
                const float cF = float(kX[d]) * Fx + float(kY[d]) * Fy + float(kZ[d]) * Fz;
                P += (kp * w) * (a9 * cu * cF - 3.0f * uF);
                M += (km * w * c3) * cF;
            }
            const float a = f[d], b = f[d + 1];
            f[d] = fmaf(A1, a, fmaf(B1, b, P + M));
            f[d + 1] = fmaf(B1, a, fmaf(A1, b, P - M));
        }
        const float w0 = w27<float>(QQ - 1);
        const float d0 = w0 * (rm - rh * (1.5f * u2));
        float r = f[QQ - 1] - wp * (f[QQ - 1] - d0);
        if (FORCED) r += kp * (w0 * (-3.0f * uF));
        f[QQ - 1] = r;
    } else {
#pragma unroll
        for (int d = 0; d < QQ; ++d) {
            const float w = w27<float>(d);
            const float cu = float(kX[d]) * ux + float(kY[d]) * uy + float(kZ[d]) * uz;
            const float cm = float(kX[d]) * mx + float(kY[d]) * my + float(kZ[d]) * mz;
            const float de = w * (rm + rh * (a2 * cu * cu - 1.5f * u2) + c3 * cm);
            float r = f[d] - wp * (f[d] - de);
            if (FORCED) {
                const float cF = float(kX[d]) * Fx + float(kY[d]) * Fy + float(kZ[d]) * Fz;
                r += kp * (w * (c3 * cF + a9 * cu * cF - 3.0f * uF));
            }
            f[d] = r;
        }
    }
}

// Equilibrium in the reference's form w rho (1 + 3cu + 4.5cu^2 - 1.5u^2)
// (src/Initialization.jl:37-42)
template <typename C>
__device__ inline C feq27(int d, C r, C ux, C uy, C uz) {
    const C cu = C(kX[d]) * ux + C(kY[d]) * uy + C(kZ[d]) * uz;
    const C u2 = ux * ux + uy * uy + uz * uz;
    return w27<C>(d) * r * (C(1) + cu * C(3) + C(0.5) * (cu * cu) * C(9) - C(0.5) * u2 * C(3));
}

// Cell coordinates with 32-bit divisions (cell counts stay below 2^31).
__device__ inline void xyz_of(int64_t cell, const Grid& g, int& x, int& y, int& z) {
    const uint32_t c = (uint32_t)cell, nx = (uint32_t)g.nx, ny = (uint32_t)g.ny;
    const uint32_t t = c / nx;
    x = (int)(c - t * nx);
    const uint32_t zz = t / ny;
    y = (int)(t - zz * ny);
    z = (int)zz;
}

// A cell whose neighbourhood may hold a solid node of body B: every neighbour
// of a cell farther than B.ext from the centre (on some axis) lies outside.
__device__ inline bool near_body(const DevBody& B, int x, int y, int z, const Grid& g) {
    if (B.reach < 0.0f) return true;  // explicit flags: no analytic extent
    const float dx = min_image((float)x + 0.5f - B.c[0], g.per & 1, g.nx);
    const float dy = min_image((float)y + 0.5f - B.c[1], g.per & 2, g.ny);
    const float dz = min_image((float)z + 0.5f - B.c[2], g.per & 4, g.nz);
    return fabsf(dx) <= B.ext[0] && fabsf(dy) <= B.ext[1] && fabsf(dz) <= B.ext[2];
}

// Faces with a velocity (bit f of vel) or anti-bounce-back (abb) rule, as bit
// masks: a select chain over the face table would be turned back into an
// indexed read of the kernel argument, which lives in private memory.
struct RuleFaces {
    uint32_t vel, abb;
};

__device__ inline RuleFaces rule_faces_of(const Grid& g) {
    RuleFaces r{0u, 0u};
#pragma unroll
    for (int f = 0; f < 6; ++f) {
        r.vel |= (g.face_kind[f] == LBM_FACE_VELOCITY ? 1u : 0u) << f;
        r.abb |= (g.face_kind[f] == LBM_FACE_PRESSURE_ABB ? 1u : 0u) << f;
    }
    return r;
}

// Distance from the point c to the cell centres x + 1/2, x in [xa, xb], along
// an axis of length L (minimum image on a periodic axis).
__device__ inline float span_dist(int xa, int xb, float c, bool per, int L) {
    float best = 3.0e38f;
    for (int k = per ? -1 : 0; k <= (per ? 1 : 0); ++k) {
        const float q = c + (float)(k * L);
        const float d = fmaxf(fmaxf((float)xa + 0.5f - q, q - ((float)xb + 0.5f)), 0.0f);
        best = fminf(best, d);
    }
    return best;
}

// Whether a cell of this block comes within reach of a body (near_body), row
// by row (a block spans at most 256 / nx + 2 rows) with the block's x span in
// each: a block is near only when one of its cells is, and the split step's
// host-computed boxes contain every such cell.
__device__ inline bool block_near(uint32_t block, const DevBody* __restrict__ bnew, const DevBody* __restrict__ bold,
                                  int nb, const Grid& g) {
    if (nb <= 0) return false;
    // 32-bit arithmetic: cell counts stay below 2^31
    const uint32_t c0 = block * (uint32_t)kBlock;
    const uint32_t c1 = (c0 + kBlock < (uint32_t)g.N ? c0 + kBlock : (uint32_t)g.N) - 1u;
    const uint32_t nx = (uint32_t)g.nx, ny = (uint32_t)g.ny;
    for (uint32_t r = c0 / nx; r <= c1 / nx; ++r) {
        const float py = (float)(r % ny) + 0.5f, pz = (float)(r / ny) + 0.5f;
        const int xa = r == c0 / nx ? (int)(c0 - r * nx) : 0;
        const int xb = r == c1 / nx ? (int)(c1 - r * nx) : g.nx - 1;
        for (int b = 0; b < nb; ++b) {
            for (int k = 0; k < 2; ++k) {
                if (k == 1 && bold == nullptr) break;
                const DevBody& B = k == 0 ? bnew[b] : bold[b];
                if (B.reach < 0.0f) return true;
                const float dy = min_image(py - B.c[1], g.per & 2, g.ny);
                const float dz = min_image(pz - B.c[2], g.per & 4, g.nz);
                if (fabsf(dy) <= B.ext[1] && fabsf(dz) <= B.ext[2] &&
                    span_dist(xa, xb, B.c[0], g.per & 1, g.nx) <= B.ext[0])
                    return true;
            }
        }
    }
    return false;
}

// Pull sources x - c of a cell for the offsets c = -1, 0, 1 on each axis
// (index c + 1), as cell-index terms: wrapped on periodic axes, clamped onto
// the cell on the others (zero gradient; the face rules replace the rest).
// Bit 3a + (c + 1) of `in`: that source exists (no clamp).  The source of
// direction d is x[1 + cx] + y[1 + cy] + z[1 + cz], two integer adds.
struct Sources {
    int x[3], y[3], z[3];
    uint32_t in;
};

__device__ inline Sources sources_of(int qx, int qy, int qz, const Grid& g) {
    Sources s;
    const int q[3] = {qx, qy, qz};
    const int L[3] = {g.nx, g.ny, g.nz};
    const int sc[3] = {1, g.nx, (int)g.nxy};
    int t[3][3];
    s.in = 0u;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const bool per = (g.per >> a) & 1;
        const bool lo_in = q[a] > 0, hi_in = q[a] < L[a] - 1;
        const int lo = lo_in ? q[a] - 1 : (per ? L[a] - 1 : q[a]);  // source for c = +1
        const int hi = hi_in ? q[a] + 1 : (per ? 0 : q[a]);         // source for c = -1
        t[a][0] = hi * sc[a];
        t[a][1] = q[a] * sc[a];
        t[a][2] = lo * sc[a];
        s.in |= ((hi_in || per) ? 1u : 0u) << (3 * a) | 1u << (3 * a + 1) | ((lo_in || per) ? 1u : 0u) << (3 * a + 2);
    }
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        s.x[k] = t[0][k];
        s.y[k] = t[1][k];
        s.z[k] = t[2][k];
    }
    return s;
}

__device__ inline int src_of(const Sources& s, int d) { return s.x[1 + kX[d]] + s.y[1 + kY[d]] + s.z[1 + kZ[d]]; }

__device__ inline bool src_in(const Sources& s, int d) {
    const uint32_t need = 1u << (1 + kX[d]) | 1u << (3 + 1 + kY[d]) | 1u << (6 + 1 + kZ[d]);
    return (s.in & need) == need;
}

// Face whose rule gives direction d when x - c_d lies beyond non-periodic
// faces: VELOCITY > PRESSURE_ABB > ZERO_GRADIENT, lower axis first
// (velocity_BC!, src/OtherBC.jl:137-166).  -1: zero gradient or inside.
__device__ inline int face_rule_of(const int q0[3], int d, const Grid& g, RuleFaces rf, bool* velocity) {
    const int L[3] = {g.nx, g.ny, g.nz};
    const int cd[3] = {kX[d], kY[d], kZ[d]};
    uint32_t out = 0u;  // bit f: the source lies beyond face f
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const int s = q0[a] - cd[a];
        if (!((g.per >> a) & 1) && (s < 0 || s >= L[a])) out |= 1u << (2 * a + (s >= L[a] ? 1 : 0));
    }
    const uint32_t v = out & rf.vel, b = out & rf.abb;
    *velocity = v != 0u;
    const uint32_t m = v != 0u ? v : b;
    return m != 0u ? __builtin_ctz(m) : -1;
}

// Face rules for direction d whose source lies beyond the chosen face:
// VELOCITY src/OtherBC.jl:143-148, PRESSURE_ABB src/OtherBC.jl:156-165.
// For the shifted formats own_opp is a deviation and rho_prev holds rho - 1
// (prev_moments): the rules in deviation form.
template <class S>
__device__ inline typename S::C velocity_rule(int d, typename S::C own_opp, const typename S::C uf[3],
                                              typename S::C rho_prev) {
    using C = typename S::C;
    const C rho = S::kDev ? C(1) + rho_prev : rho_prev;
    const C cu = C(kX[d]) * (-uf[0]) + C(kY[d]) * (-uf[1]) + C(kZ[d]) * (-uf[2]);
    return own_opp - C(2) * w27<C>(d) * rho * cu * C(3);
}

template <class S>
__device__ inline typename S::C abb_rule(int d, typename S::C own_opp, const typename S::C up[3], typename S::C uu2,
                                         typename S::C rho_prev) {
    using C = typename S::C;
    const C cu = C(kX[d]) * up[0] + C(kY[d]) * up[1] + C(kZ[d]) * up[2];
    if constexpr (S::kDev)
        return -own_opp + C(2) * w27<C>(d) *
                              (rho_prev + (C(1) + rho_prev) * (C(0.5) * cu * cu * C(9) - C(0.5) * uu2 * C(3)));
    return -own_opp + C(2) * w27<C>(d) * rho_prev * (C(1) + C(0.5) * cu * cu * C(9) - C(0.5) * uu2 * C(3));
}

// Directions whose source x - c_d lies beyond face f = 2a + side: c_a = +1 on
// the low face (side 0), c_a = -1 on the high one.
__host__ __device__ constexpr uint32_t face_dirs(int f) {
    uint32_t m = 0u;
    for (int d = 0; d < QQ; ++d) {
        const int c = (f >> 1) == 0 ? kX[d] : ((f >> 1) == 1 ? kY[d] : kZ[d]);
        if (c == ((f & 1) ? -1 : 1)) m |= 1u << d;
    }
    return m;
}

// Rule directions of a cell's faces: the union of face_dirs over the rule
// faces the cell lies on (a direction beyond several faces takes the first
// rule face in the precedence order, which face_rule_of picks).  *single: the
// one rule face of a cell on exactly one, else -1; *abb: the directions whose
// rule is the anti-bounce-back one (the only rule that reads the velocity).
__device__ inline uint32_t face_rule_mask(const int q0[3], const Grid& g, RuleFaces rf, int* single, uint32_t* abb) {
    const int L[3] = {g.nx, g.ny, g.nz};
    uint32_t vm = 0u, am = 0u;
    int n = 0, one = -1;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        if ((g.per >> a) & 1) continue;
#pragma unroll
        for (int side = 0; side < 2; ++side) {
            const int fc = 2 * a + side;
            if (q0[a] != (side ? L[a] - 1 : 0)) continue;
            if ((rf.vel >> fc) & 1u) vm |= face_dirs(fc);
            else if ((rf.abb >> fc) & 1u) am |= face_dirs(fc);
            else continue;
            ++n;
            one = fc;
        }
    }
    *single = n == 1 ? one : -1;
    *abb = am & ~vm;
    return vm | am;
}

// The rule of face FC on its nine directions, for a cell on that face only
// (f[d] holds the cell's own f*[opp(d)] the rule starts from).
template <class S, int FC>
__device__ inline void one_face_rule(typename S::C f[QQ], bool velocity, const typename S::C uf[3],
                                     const typename S::C up[3], typename S::C uu2, typename S::C rho_prev) {
#pragma unroll
    for (int d = 0; d < QQ; ++d) {
        if (!((face_dirs(FC) >> d) & 1u)) continue;
        f[d] = velocity ? velocity_rule<S>(d, f[d], uf, rho_prev) : abb_rule<S>(d, f[d], up, uu2, rho_prev);
    }
}

// In-domain cell x - c_d (periodic wrap), or -1 beyond a non-periodic face.
__device__ inline int64_t in_domain_source(const int q0[3], int d, const Grid& g) {
    const int L[3] = {g.nx, g.ny, g.nz};
    int s[3] = {q0[0] - kX[d], q0[1] - kY[d], q0[2] - kZ[d]};
    bool ok = true;
#pragma unroll
    for (int a = 0; a < 3; ++a) ok = ok && wrap_coord(s[a], L[a], (g.per >> a) & 1);
    return ok ? (int64_t)s[0] + (int64_t)g.nx * s[1] + g.nxy * s[2] : (int64_t)-1;
}

// Arm component on axis a, minimum image on periodic axes.
__device__ inline double wrap_arm(double r, int a, const Grid& g) {
    if ((g.per >> a) & 1) {
        const int L = a == 0 ? g.nx : (a == 1 ? g.ny : g.nz);
        const double h = 0.5 * L;
        if (r > h) r -= L;
        else if (r < -h) r += L;
    }
    return r;
}

// ---------------------------------------------------------------------------
// The split 16-bit step (k27_split, one launch).  A one-cell-per-thread
// 16-bit step issues one 2-byte access per lane and direction: as many memory
// instructions as FP32 for half the bytes.  Pair blocks take two x-neighbours
// per thread with 4-byte loads and stores for every cell off the rows on
// non-periodic y/z faces and outside the bodies' boxes (percell_cell); the
// x-face cells stay in their rows' waves: periodic x wraps with one 2-byte
// load, a zero-gradient face clamps (a value select), a velocity /
// anti-bounce-back face applies its rule from the cell's own f*[opp d] and
// the previous-step moments in the cell's face slot, which whichever blocks
// produced the cell's f* wrote (q27_slot_put).  Per-cell blocks (step_block,
// split = 1) take the face rows and the boxes, enumerated from kernel
// arguments (percell_block) so that no block is launched only to leave.
// ---------------------------------------------------------------------------

// A row on a non-periodic y or z face (the per-cell kernel's).
__device__ inline bool face_row27(int y, int z, const Grid& g) {
    return (!(g.per & 2) && (y == 0 || y == g.ny - 1)) || (!(g.per & 4) && (z == 0 || z == g.nz - 1));
}

// A cell of the per-cell blocks in a split step: on a y/z face row or in a
// body's box (the same answer for both cells of a pair: boxes start on even x
// and end on odd x).
__device__ inline bool percell_cell(int x, int y, int z, const Grid& g, const SplitRows& r) {
    if (face_row27(y, z, g)) return true;
    bool in = false;
    for (int k = 0; k < r.n && !in; ++k) {
        int dy = y - r.y0[k], dz = z - r.z0[k];
        if (dy < 0) dy += g.ny;
        if (dz < 0) dz += g.nz;
        in = (unsigned)(x - r.x0[k]) < (unsigned)r.ex[k] && dy < r.ey[k] && dz < r.ez[k];
    }
    return in;
}

// Slot of the x-face rule cell (x, y, z) on face f (0: x = 0, 1: x = nx - 1).
__device__ inline int x_slot(const Grid& g, int f, int y, int z) { return g.face_base[f] + y + g.ny * z; }

// The previous-step moments that an x-face rule reads, from a cell's
// populations as stored (decoded deviations dv): what the per-cell kernel
// computes from its own loads next step.
// Sum over the 64 lanes of a wave, in a fixed butterfly order (every lane gets it).
template <typename C>
__device__ __forceinline__ C wave_sum_any(C v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// Running sums of moments27 (the same additions in the same order), one
// population at a time.
template <typename C>
__device__ __forceinline__ void moment_add(int d, C v, C& rho, C& jx, C& jy, C& jz) {
    rho += v;
    if (kX[d] > 0) jx += v;
    if (kX[d] < 0) jx -= v;
    if (kY[d] > 0) jy += v;
    if (kY[d] < 0) jy -= v;
    if (kZ[d] > 0) jz += v;
    if (kZ[d] < 0) jz -= v;
}

template <class S, bool FORCED>
__device__ inline void q27_slot_put(const Grid& g, typename S::C rho, typename S::C jx, typename S::C jy,
                                    typename S::C jz, const Phys<typename S::C>& p, int slot) {
    using C = typename S::C;
    C rp, up[3], uu2;
    prev_moments<S, FORCED>(rho, jx, jy, jz, p, rp, up, uu2);
    C* o = reinterpret_cast<C*>(g.side_out) + 4 * (int64_t)slot;
    o[0] = rp;
    o[1] = up[0];
    o[2] = up[1];
    o[3] = up[2];
}

// IBB: the interpolated bounce-back (IBB!, src/SolidBC.jl:253-350) on the links
// instead of the simple one: wall fraction q from the body's level set at the
// cell and at the solid node (FP32, clamped to [kEpsQ, 1 - kEpsQ]); rule 1
// (q < 1/2 with a fluid second node x - c_dl), rule 2 (q >= 1/2, the wall
// term over 2q), else the simple rule; wall velocity and torque arm at x + q c_dl.
template <class S, int COLL, bool FORCED, int MODE, bool IBB>
__device__ __forceinline__ void step_block(const typename S::T* __restrict__ fin, typename S::T* __restrict__ fout,
                                           double* __restrict__ gout, double* __restrict__ rho_out,
                                           double* __restrict__ u_out, const Grid& g, const Phys<typename S::C>& p,
                                           const uint8_t* __restrict__ fl_new, const uint8_t* __restrict__ fl_old,
                                           const DevBody* __restrict__ bnew, const DevBody* __restrict__ bold, int nb,
                                           double* __restrict__ partials, int refill,
                                           const double* __restrict__ face_u, int split, const SplitRows& rows,
                                           uint32_t blk) {
    using C = typename S::C;
    extern __shared__ double lds_dyn[];
    const int64_t cell = (int64_t)blk * kBlock + threadIdx.x;
    // split step: the pair blocks take the cells off the y/z face rows and
    // outside the bodies' boxes (percell_cell, from kernel arguments alone)
    bool mine = true;
    if (MODE == MODE_STEP && split && cell < g.N) {
        int x, y, z;
        xyz_of(cell, g, x, y, z);
        mine = percell_cell(x, y, z, g, rows);
    }
    const int n6 = nb * 6;
    double* s_wave = lds_dyn + (threadIdx.x >> 6) * n6;  // this wave's force slots
    // IBB: the wall fraction of each link direction, one column per thread
    float* lds_q = reinterpret_cast<float*>(lds_dyn + kWaves * n6) + threadIdx.x;
    // Only blocks within reach of a body can hold links: they alone reduce
    // forces and write their partials (the sum visits the same blocks).
    const bool bnear = MODE != MODE_COLLIDE_ONLY && MODE != MODE_GET_RAW && block_near(blk, bnew, bold, nb, g);
    __shared__ uint32_t s_waves_done;
    if (MODE == MODE_STEP && bnear) {
        for (int i = threadIdx.x; i < kWaves * n6; i += kBlock) lds_dyn[i] = 0.0;
        if (threadIdx.x == 0) s_waves_done = 0u;
        __syncthreads();
    }
    ForceAcc acc;
    acc.b0 = -1;
#pragma unroll
    for (int c = 0; c < 6; ++c) acc.a0[c] = 0.0;
    const int64_t N = g.N;
    // Fresh nodes (refill): the mean previous-step rho, u of the neighbours
    // x + c_d that are fluid and not fresh, gathered by the whole wave, one
    // neighbour per lane (two memory round trips per fresh node), summed in a
    // fixed shuffle order.  rf_*: the sums for this lane's node, rf_n the count.
    bool fresh_l = false;
    C rf_r = C(0), rf_x = C(0), rf_y = C(0), rf_z = C(0);
    int rf_n = 0;
    if ((MODE == MODE_STEP || MODE == MODE_GET) && bnear && refill) {
        int x = 0, y = 0, z = 0;
        if (cell < N && mine) {
            xyz_of(cell, g, x, y, z);
            bool nr = false;
            for (int b = 0; b < nb && !nr; ++b)
                nr = near_body(bnew[b], x, y, z, g) || (bold != nullptr && near_body(bold[b], x, y, z, g));
            fresh_l = nr && (fl_new[cell] & kFlagBody) == 0 && (fl_old[cell] & kFlagBody) != 0;
        }
        const int lane = (int)(threadIdx.x & 63u);
        uint64_t fm = __ballot(fresh_l);
        while (fm) {
            const int L = __ffsll((unsigned long long)fm) - 1;
            const int q0[3] = {__shfl(x, L, 64), __shfl(y, L, 64), __shfl(z, L, 64)};
            int64_t nc = -1;
            if (lane < QQ - 1) {
#pragma unroll
                for (int d = 0; d < QQ - 1; ++d)
                    if (d == lane) nc = in_domain_source(q0, opp(d), g);
            }
            bool ok = nc >= 0;
            if (ok) ok = (fl_new[nc] & kFlagBody) == 0 && (fl_old[nc] & kFlagBody) == 0;
            C rr = C(0), uu[3] = {C(0), C(0), C(0)};
            if (ok) {
                C r = C(0), jx = C(0), jy = C(0), jz = C(0);
#pragma unroll
                for (int e = 0; e < QQ; ++e)
                    moment_add<C>(e, S::load(fin + (int64_t)e * g.P + nc, e, g.sh), r, jx, jy, jz);
                C u2;
                prev_moments<S, FORCED>(r, jx, jy, jz, p, rr, uu, u2);
            }
            const C sr = wave_sum_any<C>(rr), sx = wave_sum_any<C>(uu[0]), sy = wave_sum_any<C>(uu[1]),
                    sz = wave_sum_any<C>(uu[2]);
            const int n = __popcll(__ballot(ok));
            if (lane == L) {
                rf_r = sr;
                rf_x = sx;
                rf_y = sy;
                rf_z = sz;
                rf_n = n;
            }
            fm &= fm - 1;
        }
    }
    if (cell < N && mine) {
        C f[QQ];
        if (MODE == MODE_COLLIDE_ONLY || MODE == MODE_GET_RAW) {
#pragma unroll
            for (int d = 0; d < QQ; ++d) f[d] = S::load(fin + (int64_t)d * g.P + cell, d, g.sh);
        } else {
            int x, y, z;
            xyz_of(cell, g, x, y, z);
            const int q0[3] = {x, y, z};
            // rule directions of the faces (fmask); single: the one rule face of the cell
            const RuleFaces rf = rule_faces_of(g);
            int single = -1;
            uint32_t abb_dirs = 0u;
            const uint32_t fmask = face_rule_mask(q0, g, rf, &single, &abb_dirs);
            const Sources src = sources_of(x, y, z, g);
            bool near = false;
            if (bnear) {
                for (int b = 0; b < nb && !near; ++b)
                    near = near_body(bnew[b], x, y, z, g) || (bold != nullptr && near_body(bold[b], x, y, z, g));
            }
            const int self = near ? (fl_new[cell] & kFlagBody) : 0;
            const bool fresh = fresh_l;  // (near && refill && self == 0 && solid in the old flags)
            // links to a solid source (lmask)
            uint32_t lmask = 0u;
            int lbody = -1;       // body of the links (the last one met)
            bool lmixed = false;  // links to more than one body
            if (near && self == 0 && !fresh) {
                // 26 independent flag loads at in-domain (clamped) addresses
#pragma unroll
                for (int d = 0; d < QQ - 1; ++d) {
                    const uint8_t fv = fl_new[src_of(src, d)];
                    const int b = (fv & kFlagBody) - 1;
                    if (src_in(src, d) && b >= 0) {
                        lmask |= 1u << d;
                        lmixed = lmixed || (lbody >= 0 && lbody != b);
                        lbody = b;
                    }
                }
            }
            // IBB: rule of every link (2 bits per direction) and its wall fraction,
            // in a loop that is not unrolled (the level sets are long)
            uint64_t rules = 0u;
            if (IBB && lmask != 0u) {
                const float px = (float)x + 0.5f, py = (float)y + 0.5f, pz = (float)z + 0.5f;
#pragma unroll 1
                for (int d = 0; d < QQ - 1; ++d) {
                    if (!((lmask >> d) & 1u)) continue;
                    // link (x, dl), dl = opp(d): the solid node x + c_dl is the pull source
                    const int64_t sc = in_domain_source(q0, d, g);
                    const int bi = lmixed ? (fl_new[sc] & kFlagBody) - 1 : lbody;
                    const DevBody& B = bnew[bi];
                    const float cx = (float)(-kX[d]), cy = (float)(-kY[d]), cz = (float)(-kZ[d]);
                    const float phb = body_phi(B, px, py, pz, g);
                    const float phs = body_phi(B, px + cx, py + cy, pz + cz, g);
                    const float den = phb - phs;
                    if (!(phb > 0.0f && phs < 0.0f && fabsf(den) >= 1e-8f)) continue;
                    const float q = fminf(fmaxf(phb / den, kEpsQ), 1.0f - kEpsQ);
                    uint64_t rule = 2u;
                    if (q < 0.5f) {
                        const int64_t s2 = in_domain_source(q0, opp(d), g);  // x - c_dl = x + c_d
                        rule = s2 >= 0 && (fl_new[s2] & kFlagBody) == 0 ? 1u : 0u;
                    }
                    if (rule != 0u) {
                        lds_q[d * kBlock] = q;
                        rules |= rule << (2 * d);
                    }
                }
            }
            if (fresh) {
                // equilibrium of the neighbours' mean previous-step rho, u
                // (gathered above by the wave)
                const C sr = rf_r, su = rf_x, sv = rf_y, sw = rf_z;
                const int cnt = rf_n;
                // (shifted formats: sr sums rho - 1, and f holds f_eq - w)
                C r = S::kDev ? C(g.rho0 - 1.0) : C(g.rho0), ux = C(0), uy = C(0), uz = C(0);
                if (cnt > 0) {
                    const C n = C(cnt);
                    r = sr / n;
                    ux = su / n;
                    uy = sv / n;
                    uz = sw / n;
                }
#pragma unroll
                for (int d = 0; d < QQ; ++d) {
                    if constexpr (S::kDev) {
                        const C cu = C(kX[d]) * ux + C(kY[d]) * uy + C(kZ[d]) * uz;
                        const C u2 = ux * ux + uy * uy + uz * uz;
                        f[d] = w27<C>(d) * (r + (C(1) + r) * (cu * C(3) + C(0.5) * (cu * cu) * C(9) - C(0.5) * u2 * C(3)));
                    } else {
                        f[d] = feq27<C>(d, r, ux, uy, uz);
                    }
                }
            } else {
                const uint32_t rmask = fmask | lmask;
                // A cell on one rule face and with no link (an x-face cell shares
                // its wave with the row's bulk cells, a wall row is a wave of them)
                // takes the short path: that face's rule on its nine directions.
                const bool one_face = single >= 0 && lmask == 0u;
                C rho_prev = C(1), up[3] = {C(0), C(0), C(0)}, uu2 = C(0);
                if (rmask) {
                    // previous-step moments from the cell's own f*: rho for the
                    // velocity and link rules, u too for an anti-bounce-back one
                    C r = C(0), jx = C(0), jy = C(0), jz = C(0);
                    if (abb_dirs != 0u) {
#pragma unroll
                        for (int d = 0; d < QQ; ++d) {
                            const C v = S::load(fin + (int64_t)d * g.P + cell, d, g.sh);
                            r += v;
                            if (kX[d] > 0) jx += v;
                            if (kX[d] < 0) jx -= v;
                            if (kY[d] > 0) jy += v;
                            if (kY[d] < 0) jy -= v;
                            if (kZ[d] > 0) jz += v;
                            if (kZ[d] < 0) jz -= v;
                        }
                        prev_moments<S, FORCED>(r, jx, jy, jz, p, rho_prev, up, uu2);
                    } else {
#pragma unroll
                        for (int d = 0; d < QQ; ++d) r += S::load(fin + (int64_t)d * g.P + cell, d, g.sh);
                        rho_prev = S::kDev ? C(kRhoW27 - 1.0) + r : r;  // rho - 1 for the shifted formats
                    }
                }
                // pulls; a rule direction loads the cell's own f*[opp(d)] instead,
                // the value its rule starts from
                if (rmask == 0u) {
#pragma unroll
                    for (int d = 0; d < QQ; ++d) f[d] = S::load(fin + (int64_t)d * g.P + src_of(src, d), d, g.sh);
                } else {
#pragma unroll
                    for (int d = 0; d < QQ; ++d) {
                        const bool rule = (rmask >> d) & 1u;
                        const int64_t off = rule ? (int64_t)opp(d) * g.P + cell : (int64_t)d * g.P + src_of(src, d);
                        f[d] = S::load(fin + off, d, g.sh);
                    }
                }
                if (one_face) {
                    const int fc = single;
                    const C uf[3] = {C(face_u[3 * fc]), C(face_u[3 * fc + 1]), C(face_u[3 * fc + 2])};
                    const bool vel = (rf.vel >> fc) & 1u;
                    switch (fc) {
                        case 0: one_face_rule<S, 0>(f, vel, uf, up, uu2, rho_prev); break;
                        case 1: one_face_rule<S, 1>(f, vel, uf, up, uu2, rho_prev); break;
                        case 2: one_face_rule<S, 2>(f, vel, uf, up, uu2, rho_prev); break;
                        case 3: one_face_rule<S, 3>(f, vel, uf, up, uu2, rho_prev); break;
                        case 4: one_face_rule<S, 4>(f, vel, uf, up, uu2, rho_prev); break;
                        default: one_face_rule<S, 5>(f, vel, uf, up, uu2, rho_prev); break;
                    }
                } else if (rmask) {
                    // wall point of the moving-wall velocity: physical x + c_dl/2 or
                    // the reference's x - c_dl/2 (src/SolidBC.jl:195-196)
                    const double hs = g.wall_point == LBM_WALL_REFERENCE ? -0.5 : 0.5;
                    // kinematics of the link body, read once when every link of the
                    // cell meets one body (no dependent loads inside the loop)
                    double bch[3] = {0.0, 0.0, 0.0}, bv[3] = {0.0, 0.0, 0.0}, bom[3] = {0.0, 0.0, 0.0};
                    if (lmask != 0u && !lmixed) {
                        const DevBody& B = bnew[lbody];
#pragma unroll
                        for (int a = 0; a < 3; ++a) {
                            bch[a] = B.ch[a];
                            bv[a] = B.v[a];
                            bom[a] = B.om[a];
                        }
                    }
#pragma unroll
                    for (int d = 0; d < QQ; ++d) {
                        if (!((rmask >> d) & 1u)) continue;
                        const C own_opp = f[d];
                        const C w = w27<C>(d);
                        if ((fmask >> d) & 1u) {
                            bool vel;
                            const int ch = face_rule_of(q0, d, g, rf, &vel);
                            const C uf[3] = {C(face_u[3 * ch]), C(face_u[3 * ch + 1]), C(face_u[3 * ch + 2])};
                            if (vel) f[d] = velocity_rule<S>(d, own_opp, uf, rho_prev);
                            else f[d] = abb_rule<S>(d, own_opp, up, uu2, rho_prev);
                        } else {
                            // boundary link (x, dl), dl = opp(d), solid node x + c_dl
                            // (SBB!, src/SolidBC.jl:187-208)
                            int bi = lbody;
                            double ch[3] = {bch[0], bch[1], bch[2]}, v[3] = {bv[0], bv[1], bv[2]};
                            double om[3] = {bom[0], bom[1], bom[2]};
                            if (lmixed) {
                                bi = (fl_new[src_of(src, d)] & kFlagBody) - 1;
                                const DevBody& B = bnew[bi];
#pragma unroll
                                for (int a = 0; a < 3; ++a) {
                                    ch[a] = B.ch[a];
                                    v[a] = B.v[a];
                                    om[a] = B.om[a];
                                }
                            }
                            const int cdl[3] = {-kX[d], -kY[d], -kZ[d]};
                            const int rule = IBB ? (int)((rules >> (2 * d)) & 3u) : 0;
                            const double qw = IBB && rule != 0 ? (double)lds_q[d * kBlock] : 0.5;
                            // wall point x + h c_dl: h = q (interpolated) or 1/2 (simple),
                            // mirrored for the reference's sampling point
                            const double hw = rule == 0 ? hs : (g.wall_point == LBM_WALL_REFERENCE ? -qw : qw);
                            double r[3];
#pragma unroll
                            for (int a = 0; a < 3; ++a)
                                r[a] = wrap_arm((double)q0[a] + 0.5 + hw * (double)cdl[a] - ch[a], a, g);
                            const C uwx = C(v[0] + (om[1] * r[2] - om[2] * r[1]));
                            const C uwy = C(v[1] + (om[2] * r[0] - om[0] * r[2]));
                            const C uwz = C(v[2] + (om[0] * r[1] - om[1] * r[0]));
                            const C cuw = C(cdl[0]) * uwx + C(cdl[1]) * uwy + C(cdl[2]) * uwz;
                            const C corr = C(2) * w * (S::kDev ? C(1) + rho_prev : rho_prev) * cuw * C(3);
                            C val = own_opp - corr;
                            if (IBB && rule != 0) {
                                // rule 1: f*_dl at the second fluid node x - c_dl; rule 2: the
                                // cell's own f*_d (SBB/IBB forms of src/SolidBC.jl:321-347)
                                const bool near1 = rule == 1;
                                const int64_t off = near1 ? (int64_t)opp(d) * g.P + src_of(src, opp(d))
                                                          : (int64_t)d * g.P + cell;
                                const C fx = S::load(fin + off, d, g.sh);
                                const C q2 = C(2) * C(qw);
                                val = near1 ? q2 * own_opp + (C(1) - q2) * fx - corr
                                            : (own_opp + (q2 - C(1)) * fx - corr) / q2;
                            }
                            f[d] = val;
                            if (MODE == MODE_STEP) {
                                // momentum exchange (src/MEA.jl:102-107) in FP64, torque arm to the
                                // wall point x + c_dl/2 (simple) or x + q c_dl (interpolated)
                                const double t = S::kDev ? (double)own_opp + (double)val + 2.0 * rest27<S>(d)
                                                         : (double)(own_opp + val);
                                const double P[3] = {t * cdl[0], t * cdl[1], t * cdl[2]};
                                const double ha = rule == 0 ? 0.5 : qw;
                                double ra[3];
#pragma unroll
                                for (int a = 0; a < 3; ++a)
                                    ra[a] = wrap_arm((double)q0[a] + 0.5 + ha * (double)cdl[a] - ch[a], a, g);
                                const double v6[6] = {P[0], P[1], P[2], ra[1] * P[2] - ra[2] * P[1],
                                                      ra[2] * P[0] - ra[0] * P[2], ra[0] * P[1] - ra[1] * P[0]};
                                if (acc.b0 < 0 || acc.b0 == bi) {
                                    acc.b0 = bi;
#pragma unroll
                                    for (int c = 0; c < 6; ++c) acc.a0[c] += v6[c];
                                } else {
                                    wave_slot_add(bi, v6, s_wave);
                                }
                            }
                        }
                    }
                }
            }
        }
        if (MODE == MODE_GET || MODE == MODE_GET_RAW) {
            // f_old and its moments u = sum f c / rho + g/2 (src/ComputeMacro.jl:26-45)
            C rho, jx, jy, jz;
            moments27<C>(f, rho, jx, jy, jz);
            if (S::kDev) rho = C(kRhoW27) + rho;  // the weights carry no momentum
            C ux = jx / rho, uy = jy / rho, uz = jz / rho;
            if (FORCED) {
                ux += C(0.5) * p.g[0];
                uy += C(0.5) * p.g[1];
                uz += C(0.5) * p.g[2];
            }
            if (gout) {
#pragma unroll
                for (int d = 0; d < QQ; ++d) gout[(int64_t)d * N + cell] = (double)f[d] + rest27<S>(d);
            }
            if (rho_out) rho_out[cell] = (double)rho;
            if (u_out) {
                u_out[cell] = (double)ux;
                u_out[N + cell] = (double)uy;
                u_out[2 * N + cell] = (double)uz;
            }
        } else {
            if constexpr (S::kDev) collide27_dev<COLL, FORCED>(f, p);
            else collide27<C, COLL, FORCED>(f, p);
            const uint32_t h = sr_cell(cell, g.sh.seed);
#pragma unroll
            for (int d = 0; d < QQ; ++d) S::store_sr(fout + (int64_t)d * g.P + cell, d, f[d], g.sh, sr_dir27(h, d));
            if constexpr (S::kDev) {
                // split step: the face slot of an x-face rule cell, for the pair blocks' next step
                if (split) {
                    const uint32_t xc = (uint32_t)(cell % g.nx);
                    const RuleFaces rf = rule_faces_of(g);
                    const int fx = xc == 0u && ((rf.vel | rf.abb) & 1u) ? 0
                                   : (xc == (uint32_t)g.nx - 1u && ((rf.vel | rf.abb) & 2u) ? 1 : -1);
                    if (fx >= 0) {
                        const uint32_t row = (uint32_t)(cell / g.nx);
                        const int y = (int)(row % (uint32_t)g.ny), z = (int)(row / (uint32_t)g.ny);
                        C sr = C(0), sx = C(0), sy = C(0), sz = C(0);
#pragma unroll
                        for (int d = 0; d < QQ; ++d)
                            moment_add<C>(d, S::dec(S::sr_bits(f[d], g.sh, sr_dir27(h, d)), g.sh), sr, sx, sy, sz);
                        q27_slot_put<S, FORCED>(g, sr, sx, sy, sz, p, x_slot(g, fx, y, z));
                    }
                }
            }
        }
    }
    if (MODE == MODE_STEP && bnear) {
        // the last wave of the block to finish adds the waves' slots in a fixed
        // order (no barrier: the other waves leave as soon as they are done)
        wave_reduce_to_lds(acc.b0, acc.a0, s_wave);
        __threadfence_block();
        uint32_t done = 0u;
        if ((threadIdx.x & 63u) == 0u) done = atomicAdd(&s_waves_done, 1u);
        done = (uint32_t)__shfl((int)done, 0, 64);
        if (done == (uint32_t)kWaves - 1u) {
            __threadfence_block();
            for (int i = (int)(threadIdx.x & 63u); i < n6; i += 64) {
                double s = lds_dyn[i];
                for (int w = 1; w < kWaves; ++w) s += lds_dyn[w * n6 + i];
                partials[(int64_t)blk * n6 + i] = s;
            }
        }
    }
}

// index of a direction with c_x = +1 among the nine (-1 otherwise)
__host__ __device__ constexpr int xpos_index(int d) {
    int k = 0;
    for (int e = 0; e < QQ; ++e) {
        if (kX[e] <= 0) continue;
        if (e == d) return k;
        ++k;
    }
    return -1;
}

__device__ inline uint32_t load_word(const void* p) {
    uint32_t v;
    __builtin_memcpy(&v, p, 4);
    return v;
}
__device__ inline uint32_t load_half(const void* p) { return *reinterpret_cast<const uint16_t*>(p); }

// x treatment of a wave of pairs
enum { PX_NONE = 0, PX_WRAP = 1, PX_CLAMP = 2, PX_RULE = 3 };

// Element e of a population plane: the plane's base is uniform (scalar
// registers) and the lane's offset a 32-bit byte count (a plane of 16-bit
// populations is below 4 GiB, checked at context creation), so each access
// takes one 32-bit add instead of 64-bit address arithmetic.
__device__ __forceinline__ const char* plane_of(const void* f, int d, const Grid& g) {
    return reinterpret_cast<const char*>(f) + (size_t)d * (size_t)g.P * 2u;
}
__device__ __forceinline__ uint32_t load_at(const char* plane, uint32_t e) { return load_word(plane + 2u * e); }
__device__ __forceinline__ uint32_t load_half_at(const char* plane, uint32_t e) { return load_half(plane + 2u * e); }

template <class S, int COLL, bool FORCED, int XM>
__device__ __forceinline__ void pair_cells(const typename S::T* __restrict__ fin, typename S::T* __restrict__ fout,
                                           const Grid& g, const Phys<float>& p, const double* __restrict__ face_u,
                                           int x0, int y, int z) {
    const int nx = g.nx, nxy = (int)g.nxy;
    const int row = nx * y + nxy * z;
    const bool left = x0 == 0, right = x0 + 2 == nx;
    const RuleFaces rf = rule_faces_of(g);
    const bool rule0 = ((rf.vel | rf.abb) & 1u) != 0u, rule1 = ((rf.vel | rf.abb) & 2u) != 0u;
    const bool rl = XM == PX_RULE && left && rule0;
    const bool rr = XM == PX_RULE && right && rule1;
    // periodic y/z wraps by index arithmetic (rows on non-periodic y/z faces are not here)
    const int wym = (y == 0) ? nxy : 0, wyp = (y == g.ny - 1) ? -nxy : 0;
    const int wzm = (z == 0) ? (int)g.N : 0, wzp = (z == g.nz - 1) ? -(int)g.N : 0;
    float rho_prev = 0.f, up[3] = {0.f, 0.f, 0.f}, uu2 = 0.f;
    if (XM == PX_RULE) {
        // every lane reads a valid slot; only rule lanes use it
        int slot = rule0 ? x_slot(g, 0, y, z) : x_slot(g, 1, y, z);
        if (rr && rule0) slot = x_slot(g, 1, y, z);
        const float4 sv = *reinterpret_cast<const float4*>(reinterpret_cast<const float*>(g.side_in) + 4 * (int64_t)slot);
        rho_prev = sv.x;
        up[0] = sv.y;
        up[1] = sv.z;
        up[2] = sv.w;
        uu2 = up[0] * up[0] + up[1] * up[1] + up[2] * up[2];
    }
    // w[d]: the pulled pair, cell x0 in the low half and x0 + 1 in the high
    // half, as stored (16 bits each); decoded one cell at a time
    uint32_t w[QQ];
#pragma unroll
    for (int d = 0; d < QQ; ++d) {
        int srow = row - (kY[d] * nx + kZ[d] * nxy);
        if (kY[d] > 0) srow += wym;
        if (kY[d] < 0) srow += wyp;
        if (kZ[d] > 0) srow += wzm;
        if (kZ[d] < 0) srow += wzp;
        const char* pl = plane_of(fin, d, g);
        const uint32_t e = (uint32_t)(srow + x0);  // the pair's own x in the source row
        if (kX[d] == 0) {
            w[d] = load_at(pl, e);
        } else if (kX[d] > 0) {
            // sources (x0 - 1, x0); the left pair loads (0, 1) so that no load
            // leaves the row: x = -1 wraps, clamps onto x = 0 or takes the rule
            const uint32_t v = load_at(pl, (XM != PX_NONE && left) ? e : e - 1u);
            if (XM == PX_NONE) {
                w[d] = v;
            } else {
                const uint32_t lo = XM == PX_WRAP ? load_half_at(pl, left ? e + (uint32_t)nx - 1u : e - 1u) : (v & 0xffffu);
                w[d] = left ? (lo | (v << 16)) : ((v & 0xffff0000u) | lo);
            }
        } else {
            // sources (x0 + 1, x0 + 2); the right pair loads (nx - 2, nx - 1)
            const uint32_t v = load_at(pl, (XM != PX_NONE && right) ? e : e + 1u);
            if (XM == PX_NONE) {
                w[d] = v;
            } else {
                const uint32_t hi = XM == PX_WRAP ? load_half_at(pl, right ? e + 1u - (uint32_t)nx + 1u : e + 2u) : (v >> 16);
                w[d] = right ? ((v >> 16) | (hi << 16)) : ((v & 0xffffu) | (hi << 16));
            }
        }
    }
    const uint32_t c0 = (uint32_t)(row + x0);
    float f[QQ];
    uint32_t pa[QQ];  // cell x0's rounding inputs (S::prep), converted with cell x0 + 1's
    // cell x0: decode, face rule (x = 0), collide, rounding inputs
    {
#pragma unroll
        for (int d = 0; d < QQ; ++d) f[d] = S::dec(w[d], g.sh);
        if (XM == PX_RULE && rl) {
            // d leaves x = 0 inward: its rule starts from the cell's own f*[opp d]
            // (loads of this lane alone, once the pulls have arrived)
            const float uf[3] = {(float)face_u[0], (float)face_u[1], (float)face_u[2]};
            const bool vel = (rf.vel & 1u) != 0u;
#pragma unroll
            for (int d = 0; d < QQ; ++d) {
                if (kX[d] <= 0) continue;
                const float o = S::dec(load_half_at(plane_of(fin, opp(d), g), (uint32_t)row), g.sh);
                f[d] = vel ? velocity_rule<S>(d, o, uf, rho_prev) : abb_rule<S>(d, o, up, uu2, rho_prev);
            }
        }
        collide27_dev<COLL, FORCED>(f, p);
        const uint32_t h = sr_cell(c0, g.sh.seed);
#pragma unroll
        for (int d = 0; d < QQ; ++d) pa[d] = S::prep(f[d], g.sh, sr_dir27(h, d));
    }
    // cell x0 + 1: the same with the high halves (face rule at x = nx - 1), then
    // both cells' words
    {
#pragma unroll
        for (int d = 0; d < QQ; ++d) f[d] = S::dec(w[d] >> 16, g.sh);
        if (XM == PX_RULE && rr) {
            // d leaves x = nx - 1 inward: own f*[opp d] there
            const float uf[3] = {(float)face_u[3], (float)face_u[4], (float)face_u[5]};
            const bool vel = (rf.vel & 2u) != 0u;
#pragma unroll
            for (int d = 0; d < QQ; ++d) {
                if (kX[d] >= 0) continue;
                const float o = S::dec(load_half_at(plane_of(fin, opp(d), g), (uint32_t)(row + nx - 1)), g.sh);
                f[d] = vel ? velocity_rule<S>(d, o, uf, rho_prev) : abb_rule<S>(d, o, up, uu2, rho_prev);
            }
        }
        collide27_dev<COLL, FORCED>(f, p);
        const uint32_t h = sr_cell(c0 + 1u, g.sh.seed);
#pragma unroll
        for (int d = 0; d < QQ; ++d) {
            w[d] = S::pack_sr(pa[d], f[d], g.sh, sr_dir27(h, d));
            __builtin_nontemporal_store(w[d], reinterpret_cast<uint32_t*>(const_cast<char*>(plane_of(fout, d, g)) + 2u * c0));
        }
    }
    if (XM == PX_RULE && (rl || rr)) {
        // the rule cell as stored: its face slot
        float sr = 0.f, sx = 0.f, sy = 0.f, sz = 0.f;
#pragma unroll
        for (int d = 0; d < QQ; ++d) moment_add<float>(d, S::dec(rl ? w[d] : (w[d] >> 16), g.sh), sr, sx, sy, sz);
        q27_slot_put<S, FORCED>(g, sr, sx, sy, sz, p, x_slot(g, rl ? 0 : 1, y, z));
    }
}

// Thread t of the pair blocks: cells 2t and 2t + 1 (one row: nx is even),
// unless the per-cell blocks take them (percell_cell).
template <class S, int COLL, bool FORCED>
__device__ __forceinline__ void pair_block(const typename S::T* __restrict__ fin, typename S::T* __restrict__ fout,
                                           const Grid& g, const Phys<float>& p, const double* __restrict__ face_u,
                                           const SplitRows& rows, uint32_t pblk) {
    const int64_t t = (int64_t)pblk * kBlock + threadIdx.x;
    const int64_t c0 = 2 * t;
    if (c0 >= g.N) return;
    int x0, y, z;
    xyz_of(c0, g, x0, y, z);
    if (percell_cell(x0, y, z, g, rows)) return;
    const bool edge = __ballot(x0 == 0 || x0 + 2 == g.nx) != 0ull;
    if (!edge) pair_cells<S, COLL, FORCED, PX_NONE>(fin, fout, g, p, face_u, x0, y, z);
    else if (g.per & 1) pair_cells<S, COLL, FORCED, PX_WRAP>(fin, fout, g, p, face_u, x0, y, z);
    else if (((rule_faces_of(g).vel | rule_faces_of(g).abb) & 3u) != 0u)
        pair_cells<S, COLL, FORCED, PX_RULE>(fin, fout, g, p, face_u, x0, y, z);
    else pair_cells<S, COLL, FORCED, PX_CLAMP>(fin, fout, g, p, face_u, x0, y, z);
}

// The step kernel (one thread per cell, blocks of 256 cells in cell order).
template <class S, int COLL, bool FORCED, int MODE, bool IBB>
__global__ __launch_bounds__(kBlock) void k27_step(const typename S::T* __restrict__ fin, typename S::T* __restrict__ fout,
                                                   double* __restrict__ gout, double* __restrict__ rho_out,
                                                   double* __restrict__ u_out, Grid g, Phys<typename S::C> p,
                                                   const uint8_t* __restrict__ fl_new, const uint8_t* __restrict__ fl_old,
                                                   const DevBody* __restrict__ bnew, const DevBody* __restrict__ bold,
                                                   int nb, double* __restrict__ partials, int refill,
                                                   const double* __restrict__ face_u, int split, SplitRows rows) {
    step_block<S, COLL, FORCED, MODE, IBB>(fin, fout, gout, rho_out, u_out, g, p, fl_new, fl_old, bnew, bold, nb,
                                           partials, refill, face_u, split, rows, blockIdx.x);
}

// The split 16-bit step in one launch: the per-cell blocks first (boundary
// work, bound by latency: it starts early), then the pair blocks over the
// plain cells.
// Block of per-cell slot v of a split step (SplitRows), or -1 for a slot past
// the end of its row or range: the boxes' slots first (their blocks carry the
// bodies, the longest work), then the face rows'.
__device__ inline int64_t percell_block(uint32_t v, const Grid& g, const SplitRows& r) {
    const int ny = g.ny, nz = g.nz;
    const uint32_t nbox = (uint32_t)r.base[r.n];
    int y, z;
    if (v >= nbox) {
        const bool zf = !(g.per & 4);
        const uint32_t bpr = (uint32_t)r.bpr;
        const uint32_t j = (v - nbox) % bpr, k0 = (v - nbox) / bpr;
        // z faces: ny rows at z = 0, ny at z = nz - 1; then y faces: y = 0 and
        // ny - 1 on every other z
        uint32_t k = k0;
        if (zf && k < 2u * (uint32_t)ny) {
            y = (int)(k % (uint32_t)ny);
            z = k < (uint32_t)ny ? 0 : nz - 1;
        } else {
            if (zf) k -= 2u * (uint32_t)ny;
            y = (k & 1u) ? ny - 1 : 0;
            z = (int)(k >> 1) + (zf ? 1 : 0);
        }
        const int64_t c0 = ((int64_t)y + (int64_t)ny * z) * g.nx;
        const int64_t first = c0 / kBlock, last = (c0 + g.nx - 1) / kBlock;
        return first + j <= last ? first + j : -1;
    }
    int b = 0;
    while (b + 1 < r.n && v >= (uint32_t)r.base[b + 1]) ++b;
    const uint32_t l = v - (uint32_t)r.base[b], bx = (uint32_t)r.bpx[b];
    const uint32_t jb = l % bx, row = l / bx;
    y = r.y0[b] + (int)(row % (uint32_t)r.ey[b]);
    z = r.z0[b] + (int)(row / (uint32_t)r.ey[b]);
    if (y >= ny) y -= ny;
    if (z >= nz) z -= nz;
    const int64_t c0 = ((int64_t)y + (int64_t)ny * z) * g.nx + r.x0[b];
    const int64_t first = c0 / kBlock, last = (c0 + r.ex[b] - 1) / kBlock;
    return first + jb <= last ? first + jb : -1;
}

template <class S, int COLL, bool FORCED, bool IBB>
__global__ __launch_bounds__(kBlock) void k27_split(const typename S::T* __restrict__ fin, typename S::T* __restrict__ fout,
                                                    Grid g, Phys<float> p, const uint8_t* __restrict__ fl_new,
                                                    const uint8_t* __restrict__ fl_old,
                                                    const DevBody* __restrict__ bnew, const DevBody* __restrict__ bold,
                                                    int nb, double* __restrict__ partials, int refill,
                                                    const double* __restrict__ face_u, SplitRows rows,
                                                    uint32_t cell_blocks) {
    const uint32_t i = blockIdx.x;
    const bool cellb = i < cell_blocks;
    const uint32_t vi = cellb ? i : i - cell_blocks;
    if (cellb) {
        const int64_t b = percell_block(vi, g, rows);
        if (b >= 0)
            step_block<S, COLL, FORCED, MODE_STEP, IBB>(fin, fout, nullptr, nullptr, nullptr, g, p, fl_new, fl_old,
                                                        bnew, bold, nb, partials, refill, face_u, 1, rows,
                                                        (uint32_t)b);
    } else {
        pair_block<S, COLL, FORCED>(fin, fout, g, p, face_u, rows, vi);
    }
}

// initialize_f (src/Initialization.jl:37-42)
template <class S>
__global__ __launch_bounds__(kBlock) void k27_equilibrium(typename S::T* __restrict__ f, const double* __restrict__ rho,
                                                          const double* __restrict__ u, double rho0, double u0x,
                                                          double u0y, double u0z, int64_t N, int64_t P, Shift sh) {
    using C = typename S::C;
    const int64_t cell = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (cell >= N) return;
    const C r = rho ? C(rho[cell]) : C(rho0);
    const C ux = u ? C(u[cell]) : C(u0x), uy = u ? C(u[N + cell]) : C(u0y), uz = u ? C(u[2 * N + cell]) : C(u0z);
#pragma unroll
    for (int d = 0; d < QQ; ++d) S::store(f + (int64_t)d * P + cell, d, feq27<C>(d, r, ux, uy, uz), sh);
}

template <class S>
__global__ __launch_bounds__(kBlock) void k27_from_double(typename S::T* __restrict__ f, const double* __restrict__ src,
                                                          int64_t N, int64_t P, Shift sh) {
    const int64_t cell = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (cell >= N) return;
#pragma unroll
    for (int d = 0; d < QQ; ++d) S::store(f + (int64_t)d * P + cell, d, (typename S::C)src[(int64_t)d * N + cell], sh);
}

template <class S, int MODE, bool IBB>
static void launch_mode(const StepArgs& a) {
    using T = typename S::T;
    using C = typename S::C;
    Phys<C> p;
    p.wp = (C)a.wp;
    p.wm = (C)a.wm;
    for (int k = 0; k < 3; ++k) p.g[k] = (C)a.grav[k];
    const dim3 grid((unsigned)nblocks_of(a.g.N)), block(kBlock);
    size_t lds = MODE == MODE_STEP ? (size_t)kWaves * a.nb * 6 * sizeof(double) : 0;
    if (IBB) lds = (size_t)kWaves * a.nb * 6 * sizeof(double) + (size_t)(QQ - 1) * kBlock * sizeof(float);
    if constexpr (S::kDev && MODE == MODE_STEP) {
        if (a.split) {
            const unsigned cb = (unsigned)a.rows.total, pb = (unsigned)nblocks_of(a.g.N / 2);
#define KS(COLL, FORCED)                                                                                          \
    hipLaunchKernelGGL((k27_split<S, COLL, FORCED, IBB>), dim3(cb + pb), block, lds, a.stream, (const T*)a.fin,    \
                       (T*)a.fout, a.g, p, a.fl_new, a.fl_old, a.bnew, a.bold, a.nb, a.partials, a.refill, a.face_u, \
                       a.rows, cb)
            if (a.trt) {
                if (a.forced) KS(LBM_TRT, true);
                else KS(LBM_TRT, false);
            } else {
                if (a.forced) KS(LBM_BGK, true);
                else KS(LBM_BGK, false);
            }
#undef KS
            return;
        }
    }
#define K27(COLL, FORCED)                                                                                           \
    hipLaunchKernelGGL((k27_step<S, COLL, FORCED, MODE, IBB>), grid, block, lds, a.stream, (const T*)a.fin, (T*)a.fout, \
                       a.gout, a.rho_out, a.u_out, a.g, p, a.fl_new, a.fl_old, a.bnew, a.bold, a.nb, a.partials,   \
                       a.refill, a.face_u, a.split, a.rows)
    if (a.trt) {
        if (a.forced) K27(LBM_TRT, true);
        else K27(LBM_TRT, false);
    } else {
        if (a.forced) K27(LBM_BGK, true);
        else K27(LBM_BGK, false);
    }
#undef K27
}

template <class S, bool IBB>
static void launch_storage(const StepArgs& a) {
    switch (a.mode) {
        case MODE_STEP: launch_mode<S, MODE_STEP, IBB>(a); break;
        case MODE_GET: launch_mode<S, MODE_GET, IBB>(a); break;
        case MODE_COLLIDE_ONLY: launch_mode<S, MODE_COLLIDE_ONLY, IBB>(a); break;
        default: launch_mode<S, MODE_GET_RAW, IBB>(a); break;
    }
}

static int last_error() {
    const hipError_t e = hipGetLastError();
    return e == hipSuccess ? 0 : (int)e;
}

}  // namespace q27
}  // namespace lbm
