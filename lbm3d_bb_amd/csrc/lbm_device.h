// lbm_device.h -- gfx950 kernels of the D3Q19 per-step path.
//
// State between steps is the post-collision population f* (the reference's
// f_new after collision!, src/Collision.jl:166-170).  One step is two launches:
//
//   k_bulk   every cell that is not on a non-periodic face: pull f*[d, x - c_d]
//            (periodic wrap by index arithmetic, src/Utilities.jl:757-767 and
//            :878-891), moments, BGK/TRT collision, store.  No flags, no lists:
//            exactly 19 loads and 19 stores per cell, fully coalesced SoA.
//   k_fixup  only the cells the bulk pass cannot finish: face cells (velocity /
//            anti-bounce-back / zero-gradient, src/OtherBC.jl:137-166) and
//            fluid cells with a solid neighbour (moving bounce-back,
//            src/SolidBC.jl:187-208, and momentum exchange, src/MEA.jl:98-108).
//            They are recomputed from scratch and overwrite the bulk result.
//
// The moments that the reference's boundary conditions read from the previous
// step (rho[i], u[i]: src/MainCore.jl:316,336) are the moments of the local
// f*: collision conserves mass, and momentum up to the known body-force term.
#pragma once

#include "lbm_lattice.h"
#include "lbm_types.h"

namespace lbm {



// ---------------------------------------------------------------------------
// Geometry predicates (FP32, no contraction, same op order as the oracle).
// ---------------------------------------------------------------------------
__host__ __device__ inline float min_image(float d, bool periodic, int L) {
#pragma clang fp contract(off)
    if (periodic) {
        float h = 0.5f * (float)L;
        if (d > h) d -= (float)L;
        else if (d < -h) d += (float)L;
    }
    return d;
}

// NACA 4-digit lower and upper surface heights at the chordwise position x
// from the leading edge (_naca_upper_lower_y, src/Geometry.jl:347-356): the
// half thickness (_naca_thickness_y, :325-330) and the camber line with its
// slope (_naca_camber_y_dydx, :332-345), in the reference's operation order.
// The camber line is in chord fractions, not scaled by the chord, as the
// reference evaluates it.  xi^4 is formed in FP64 and rounded once (Julia's
// ^(::Float32, ::Integer)); cos(atan(s)) is 1 / sqrt(1 + s^2), which uses
// only correctly rounded operations, so host and device agree bit for bit.
template <typename T>
__host__ __device__ inline void naca_lines(T x, T chord, T m, T p, T t, T x_int, T& yl, T& yu) {
#pragma clang fp contract(off)
    const T xbar = x / chord;
    T xi = x_int * xbar;
    xi = xi > T(0) ? xi : T(0);
    const T rt = sqrt(xi);
    const T x3 = xi * xi * xi;
    const double xd = (double)xi;
    const T x4 = (T)((xd * xd) * (xd * xd));
    T poly = T(0.2969) * rt - T(0.1260) * xi;
    poly = poly - T(0.3516) * xi * xi;
    poly = poly + T(0.2843) * x3;
    poly = poly - T(0.1015) * x4;
    const T yt = T(5) * t * chord * poly;
    T yc = T(0), dyb = T(0);
    if (m == T(0) || p == T(0)) {
    } else if (xbar < p) {
        yc = m / (p * p) * (T(2) * p * xbar - xbar * xbar);
        dyb = T(2) * m / (p * p) * (p - xbar);
    } else {
        const T den = T(1) - p;
        const T den2 = den * den;
        yc = m / den2 * ((T(1) - T(2) * p) + T(2) * p * xbar - xbar * xbar);
        dyb = T(2) * m / den2 * (p - xbar);
    }
    const T sl = dyb / chord;
    const T cs = T(1) / sqrt(T(1) + sl * sl);
    yu = yc + yt * cs;
    yl = yc - yt * cs;
}

// The airfoil's inside test (src/Geometry.jl:75-82) on the minimum-image
// offset from the pivot.  Kept out of line: the solidity tests are inlined at
// every neighbour of every fix-up cell, and the airfoil's divisions and square
// roots there would multiply the kernel's code size.
__device__ __noinline__ bool inside_naca(const DevBody& b, float dx, float dy) {
#pragma clang fp contract(off)
    const float xb = b.rot[0] * dx + b.rot[3] * dy;
    const float yb = b.rot[1] * dx + b.rot[4] * dy;
    const float xle = xb + b.naca[4];
    const float chord = b.semi[0];
    if (!(0.0f <= xle && xle <= chord)) return false;
    float yl, yu;
    naca_lines<float>(xle, chord, b.naca[0], b.naca[1], b.naca[2], b.naca[3], yl, yu);
    const float tol = 1e-6f * chord;
    return yl - tol <= yb && yb <= yu + tol;
}

// generate_mask_s_local!_gpu's tests (src/Geometry.jl:66-82): cylinder
// (extruded along z) and NACA airfoil; the sphere is the cylinder's test with
// the z term, the ellipsoid its body-frame analogue.
__device__ inline bool inside_body(const DevBody& b, float px, float py, float pz, const Grid& g) {
#pragma clang fp contract(off)
    float dx = min_image(px - b.c[0], g.per & 1, g.nx);
    float dy = min_image(py - b.c[1], g.per & 2, g.ny);
    float dz = min_image(pz - b.c[2], g.per & 4, g.nz);
    if (b.shape == LBM_SPHERE || b.shape == LBM_CYLINDER) {
        float s = dx * dx + dy * dy;
        if (b.shape == LBM_SPHERE) s = s + dz * dz;
        return s <= b.r2 + 1e-6f * b.r2;
    }
    if (__builtin_expect(b.shape == LBM_NACA4, 0)) return inside_naca(b, dx, dy);
    float xb = b.rot[0] * dx + b.rot[3] * dy;
    xb = xb + b.rot[6] * dz;
    float yb = b.rot[1] * dx + b.rot[4] * dy;
    yb = yb + b.rot[7] * dz;
    float zb = b.rot[2] * dx + b.rot[5] * dy;
    zb = zb + b.rot[8] * dz;
    float q = xb * xb * b.ia2[0] + yb * yb * b.ia2[1];
    q = q + zb * zb * b.ia2[2];
    return q <= 1.0f + 1e-6f;
}

// Level set of body b at the point p, FP32 at the half-step pose (as the
// reference evaluates it): negative inside, zero on the surface.  Cylinder:
// phi_cylinder_lat (src/Geometry.jl:368-372); NACA4: phi_airfoil_lat
// (src/Geometry.jl:375-396), the vertical distance to the nearer surface over
// the chord and the distance to the leading / trailing edge point beyond it;
// sphere |d| - R; ellipsoid sqrt(sum (x_i / a_i)^2) - 1.  The interpolated
// bounce-back uses only the ratio phi(x) / (phi(x) - phi(x')).  Only correctly
// rounded operations, so host and device agree bit for bit.
__device__ __forceinline__ float body_phi(const DevBody& b, float px, float py, float pz, const Grid& g) {
#pragma clang fp contract(off)
    const float dx = min_image(px - (float)b.ch[0], g.per & 1, g.nx);
    const float dy = min_image(py - (float)b.ch[1], g.per & 2, g.ny);
    const float dz = min_image(pz - (float)b.ch[2], g.per & 4, g.nz);
    if (b.shape == LBM_SPHERE || b.shape == LBM_CYLINDER) {
        float s = dx * dx + dy * dy;
        if (b.shape == LBM_SPHERE) s = s + dz * dz;
        return sqrtf(s) - b.semi[0];
    }
    if (b.shape == LBM_ELLIPSOID) {
        float e = 0.0f;
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            float r = b.rh[k] * dx + b.rh[3 + k] * dy;
            r = r + b.rh[6 + k] * dz;
            const float t = r / b.semi[k];
            e = e + t * t;
        }
        return sqrtf(e) - 1.0f;
    }
    const float xb = b.rh[0] * dx + b.rh[3] * dy;
    const float yb = b.rh[1] * dx + b.rh[4] * dy;
    const float xle = xb + b.naca[4];
    const float chord = b.semi[0];
    if (xle < 0.0f) return sqrtf(xle * xle + yb * yb);
    if (xle > chord) return sqrtf((xle - chord) * (xle - chord) + yb * yb);
    float yl, yu;
    naca_lines<float>(xle, chord, b.naca[0], b.naca[1], b.naca[2], b.naca[3], yl, yu);
    if (yb >= yl && yb <= yu) return -fminf(yu - yb, yb - yl);
    return yb > yu ? yb - yu : yl - yb;
}

// Flag of a cell: 0 fluid, k+1 inside body k (the first one wins).
__device__ inline int flag_of(const DevBody* bodies, int nb, int x, int y, int z, const Grid& g) {
    float px = (float)x + 0.5f, py = (float)y + 0.5f, pz = (float)z + 0.5f;
    for (int b = 0; b < nb; ++b)
        if (inside_body(bodies[b], px, py, pz, g)) return b + 1;
    return 0;
}

__device__ inline void cell_xyz(int64_t cell, const Grid& g, int& x, int& y, int& z) {
    int64_t t = cell / g.nx;
    x = (int)(cell - t * g.nx);
    z = (int)(t / g.ny);
    y = (int)(t - (int64_t)z * g.ny);
}

__device__ inline bool on_open_face(int x, int y, int z, const Grid& g) {
    return (!(g.per & 1) && (x == 0 || x == g.nx - 1)) || (!(g.per & 2) && (y == 0 || y == g.ny - 1)) ||
           (!(g.per & 4) && (z == 0 || z == g.nz - 1));
}

// Neighbour along +c (sx, sy, sz already offset); returns false outside a
// non-periodic face.
__device__ inline bool wrap_coord(int& s, int L, bool periodic) {
    if (s < 0 || s >= L) {
        if (!periodic) return false;
        s = s < 0 ? s + L : s - L;
    }
    return true;
}

// ---------------------------------------------------------------------------
// Collision: BGK (src/Collision.jl:232-240) and TRT (src/Collision.jl:136-170)
// on one cell, in place, from the pulled populations.  With a body force the
// velocity is u = (sum f c + F/2)/rho and the Guo source
// S_i = w_i [3 (c_i - u).F + 9 (c_i.u)(c_i.F)] is added with (1 - omega/2),
// split into even/odd parts for TRT.
// ---------------------------------------------------------------------------
template <typename C>
__device__ inline void moments(const C f[Q], C& rho, C& jx, C& jy, C& jz) {
    rho = C(0);
    jx = jy = jz = C(0);
#pragma unroll
    for (int d = 0; d < Q; ++d) {
        rho += f[d];
        if (kCx[d] > 0) jx += f[d];
        if (kCx[d] < 0) jx -= f[d];
        if (kCy[d] > 0) jy += f[d];
        if (kCy[d] < 0) jy -= f[d];
        if (kCz[d] > 0) jz += f[d];
        if (kCz[d] < 0) jz -= f[d];
    }
}

template <typename C, int COLL, bool FORCED>
__device__ inline void collide(C f[Q], const Phys<C>& p, C& rho_out, C u_out[3]) {
    C rho, jx, jy, jz;
    moments<C>(f, rho, jx, jy, jz);
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
    rho_out = rho;
    u_out[0] = ux;
    u_out[1] = uy;
    u_out[2] = uz;
    const C u2 = ux * ux + uy * uy + uz * uz;
    const C wp = p.wp;
    if (COLL == LBM_TRT) {
        const C wm = p.wm;
        const C kp = C(1) - C(0.5) * wp, km = C(1) - C(0.5) * wm;
#pragma unroll
        for (int d = 0; d < 18; d += 2) {
            const C w = weight<C>(d);
            const C cu = C(kCx[d]) * ux + C(kCy[d]) * uy + C(kCz[d]) * uz;
            const C feq_p = w * rho * (C(1) + C(4.5) * cu * cu - C(1.5) * u2);
            const C feq_m = w * rho * (C(3) * cu);
            const C fp = C(0.5) * (f[d] + f[d + 1]);
            const C fm = C(0.5) * (f[d] - f[d + 1]);
            C fpn = fp - wp * (fp - feq_p);
            C fmn = fm - wm * (fm - feq_m);
            if (FORCED) {
                const C cF = C(kCx[d]) * Fx + C(kCy[d]) * Fy + C(kCz[d]) * Fz;
                fpn += kp * (w * (C(9) * cu * cF - C(3) * uF));
                fmn += km * (w * (C(3) * cF));
            }
            f[d] = fpn + fmn;
            f[d + 1] = fpn - fmn;
        }
        const C w0 = weight<C>(18);
        const C feq0 = w0 * rho * (C(1) - C(1.5) * u2);
        C r = f[18] - wp * (f[18] - feq0);
        if (FORCED) r += kp * (w0 * (-C(3) * uF));
        f[18] = r;
    } else {
        const C kp = C(1) - C(0.5) * wp;
#pragma unroll
        for (int d = 0; d < Q; ++d) {
            const C w = weight<C>(d);
            const C cu = C(kCx[d]) * ux + C(kCy[d]) * uy + C(kCz[d]) * uz;
            const C feq = w * rho * (C(1) + C(3) * cu + C(4.5) * cu * cu - C(1.5) * u2);
            C r = f[d] - wp * (f[d] - feq);
            if (FORCED) {
                const C cF = C(kCx[d]) * Fx + C(kCy[d]) * Fy + C(kCz[d]) * Fz;
                r += kp * (w * (C(3) * cF + C(9) * cu * cF - C(3) * uF));
            }
            f[d] = r;
        }
    }
}

// The same collision on deviations delta_d = f_d - w_d (the shifted 16-bit
// formats), in FP32.  The populations stay at the size of their deviation
// through the arithmetic, so its rounding is relative to that, not to w_d; and
// the equilibrium is normalised with the sums of the FP32 weights
// (rho_w = sum w_d, s2 = sum w_d c_dx^2) so that it carries exactly the cell's
// mass and momentum: with exact weights every coefficient reduces to the
// reference's (1, 3, 4.5, 9), while the FP32 weights, which do not sum to 1,
// would otherwise add (rho_w - 1) omega of mass every step.
//   f_eq - w = w [rm + rh (a2 (c.u)^2 - 1.5 u^2)] + w c3 c.(rho u),
//   rm = sum delta / rho_w, rh = 1 + rm, c3 = 1 / s2, a2 = 1.5 rho_w / s2;
// Guo source w [kp (a9 (c.u)(c.F) - 3 u.F) + km c3 c.F], a9 = 3 rho_w / s2.
template <int COLL, bool FORCED>
__device__ inline void collide_dev(float f[Q], const Phys<float>& p) {
    constexpr float rw = (float)kRhoW19, irw = (float)(1.0 / kRhoW19);
    constexpr float c3 = (float)(1.0 / kS2_19), a2 = (float)(1.5 * kRhoW19 / kS2_19),
                    a9 = (float)(3.0 * kRhoW19 / kS2_19);
    float sd, jx, jy, jz;
    moments<float>(f, sd, jx, jy, jz);
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
#pragma unroll
        for (int d = 0; d < 18; d += 2) {
            const float w = weight<float>(d);
            const float cu = float(kCx[d]) * ux + float(kCy[d]) * uy + float(kCz[d]) * uz;
            const float cm = float(kCx[d]) * mx + float(kCy[d]) * my + float(kCz[d]) * mz;
            const float dp = w * (rm + rh * (a2 * cu * cu - 1.5f * u2));
            const float dm = w * (c3 * cm);
            const float fp = 0.5f * (f[d] + f[d + 1]);
            const float fm = 0.5f * (f[d] - f[d + 1]);
            float fpn = fp - wp * (fp - dp);
            float fmn = fm - wm * (fm - dm);
            if (FORCED) {
                const float cF = float(kCx[d]) * Fx + float(kCy[d]) * Fy + float(kCz[d]) * Fz;
                fpn += kp * (w * (a9 * cu * cF - 3.0f * uF));
                fmn += km * (w * (c3 * cF));
            }
            f[d] = fpn + fmn;
            f[d + 1] = fpn - fmn;
        }
        const float w0 = weight<float>(18);
        const float d0 = w0 * (rm - rh * (1.5f * u2));
        float r = f[18] - wp * (f[18] - d0);
        if (FORCED) r += kp * (w0 * (-3.0f * uF));
        f[18] = r;
    } else {
#pragma unroll
        for (int d = 0; d < Q; ++d) {
            const float w = weight<float>(d);
            const float cu = float(kCx[d]) * ux + float(kCy[d]) * uy + float(kCz[d]) * uz;
            const float cm = float(kCx[d]) * mx + float(kCy[d]) * my + float(kCz[d]) * mz;
            const float de = w * (rm + rh * (a2 * cu * cu - 1.5f * u2) + c3 * cm);
            float r = f[d] - wp * (f[d] - de);
            if (FORCED) {
                const float cF = float(kCx[d]) * Fx + float(kCy[d]) * Fy + float(kCz[d]) * Fz;
                r += kp * (w * (c3 * cF + a9 * cu * cF - 3.0f * uF));
            }
            f[d] = r;
        }
    }
}

// collision in the storage's register form
template <class S, int COLL, bool FORCED>
__device__ inline void collide_s(typename S::C f[Q], const Phys<typename S::C>& p) {
    if constexpr (S::kDev) {
        collide_dev<COLL, FORCED>(f, p);
    } else {
        typename S::C rho, u[3];
        collide<typename S::C, COLL, FORCED>(f, p, rho, u);
    }
}

// rest level of direction d in the storage's register form: w_d for the
// deviations of the shifted formats, 0 for absolute populations
template <class S>
__device__ inline double rest_level(int d) {
    if constexpr (S::kDev) return (double)weight<float>(d);
    return 0.0;
}

// Moments of a pulled (pre-collision) state for the getters.
template <typename C, bool FORCED>
__device__ inline void macros_of(const C f[Q], const Phys<C>& p, C& rho, C u[3]) {
    C jx, jy, jz;
    moments<C>(f, rho, jx, jy, jz);
    u[0] = jx / rho;
    u[1] = jy / rho;
    u[2] = jz / rho;
    if (FORCED) {
        u[0] += C(0.5) * p.g[0];
        u[1] += C(0.5) * p.g[1];
        u[2] += C(0.5) * p.g[2];
    }
}

// Writes the finished cell: f* in storage format (STEP, COLLIDE_ONLY) or the
// pulled f_old plus moments in FP64 (GET, GET_RAW).
template <class S, int COLL, bool FORCED, int MODE>
__device__ inline void finish_cell(typename S::C f[Q], const Phys<typename S::C>& p, int64_t cell, int64_t N, int64_t P,
                                   typename S::T* __restrict__ fout, double* __restrict__ gout,
                                   double* __restrict__ rho_out, double* __restrict__ u_out, const Shift& sh) {
    using C = typename S::C;
    if (MODE == MODE_STEP || MODE == MODE_COLLIDE_ONLY) {
        collide_s<S, COLL, FORCED>(f, p);
        const uint32_t h = sr_cell(cell, sh.seed);
#pragma unroll
        for (int d = 0; d < Q; ++d) S::store_sr(fout + (int64_t)d * P + cell, d, f[d], sh, sr_dir(h, d));
    } else {
        if (gout) {
#pragma unroll
            for (int d = 0; d < Q; ++d) gout[(int64_t)d * N + cell] = (double)f[d] + rest_level<S>(d);
        }
        if (rho_out) {
            C rho, u[3];
            if constexpr (S::kDev) {
                // rho = rho_w + sum delta, u = j / rho (+ g/2): the weights carry no momentum
                C sd, jx, jy, jz;
                moments<C>(f, sd, jx, jy, jz);
                rho = C(kRhoW19) + sd;
                u[0] = jx / rho;
                u[1] = jy / rho;
                u[2] = jz / rho;
                if (FORCED)
                    for (int a = 0; a < 3; ++a) u[a] += C(0.5) * p.g[a];
            } else {
                macros_of<C, FORCED>(f, p, rho, u);
            }
            rho_out[cell] = (double)rho;
            u_out[cell] = (double)u[0];
            u_out[N + cell] = (double)u[1];
            u_out[2 * N + cell] = (double)u[2];
        }
    }
}

// ---------------------------------------------------------------------------
// Geometry views shared by the kernels below.
// ---------------------------------------------------------------------------
struct SolidView {
    // mode 0: flags; mode 1: analytic (new and old poses)
    int mode;
    int nb;
    const uint8_t* flags;
    const DevBody* bnew;
    const DevBody* bold;
};

__device__ inline int solid_at(const SolidView& sv, int x, int y, int z, const Grid& g) {
    if (sv.mode == 0) return sv.flags[(int64_t)x + (int64_t)g.nx * y + g.nxy * z] & kFlagBody;
    return flag_of(sv.bnew, sv.nb, x, y, z, g);
}

__device__ inline bool fresh_at(const SolidView& sv, int x, int y, int z, const Grid& g) {
    if (sv.mode == 0) return false;
    return flag_of(sv.bnew, sv.nb, x, y, z, g) == 0 && flag_of(sv.bold, sv.nb, x, y, z, g) != 0;
}

__device__ inline bool in_box(const Box& b, int x, int y, int z, const Grid& g) {
    const int q[3] = {x, y, z};
    const int L[3] = {g.nx, g.ny, g.nz};
    for (int a = 0; a < 3; ++a) {
        int r = q[a] - b.lo[a];
        if ((g.per >> a) & 1) {
            r %= L[a];
            if (r < 0) r += L[a];
        }
        if (r < 0 || r >= b.ext[a]) return false;
    }
    return true;
}

// in_box for a box whose lo lies in [0, L) on periodic axes (as the host
// builds them): one conditional wrap instead of a remainder
__device__ inline bool in_box_normalised(const Box& b, int x, int y, int z, const Grid& g) {
    const int q[3] = {x, y, z};
    const int L[3] = {g.nx, g.ny, g.nz};
    bool in = true;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        int r = q[a] - b.lo[a];
        if (r < 0 && ((g.per >> a) & 1)) r += L[a];
        in = in && r >= 0 && r < b.ext[a];
    }
    return in;
}

// Body regions (box kind 1) cover every cell that can be a boundary node, a
// fresh node or a face cell next to a body, in both geometry modes.
__device__ inline bool in_body_box(const StepGeom& sg, int x, int y, int z, const Grid& g) {
    for (int m = 0; m < sg.nbox; ++m)
        if (sg.box[m].kind == 1 && in_box(sg.box[m], x, y, z, g)) return true;
    return false;
}

// Body-region tests of the bulk rows from the record's BulkGeom (one
// conditional wrap: lo is normalised into [0, L) on periodic axes).
__device__ inline bool in_bulk_box(const BulkGeom& bg, int m, int x, int y, int z, const Grid& g) {
    const int q[3] = {x, y, z};
    const int L[3] = {g.nx, g.ny, g.nz};
    bool in = true;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        int r = q[a] - bg.lo[a][m];
        if (r < 0 && ((g.per >> a) & 1)) r += L[a];
        in = in && r >= 0 && r < bg.ext[a][m];
    }
    return in;
}

__device__ inline bool row_in_bulk_box(const BulkGeom& bg, int m, int y, int z, const Grid& g) {
    int ry = y - bg.lo[1][m], rz = z - bg.lo[2][m];
    if (ry < 0 && (g.per & 2)) ry += g.ny;
    if (rz < 0 && (g.per & 4)) rz += g.nz;
    return ry >= 0 && ry < bg.ext[1][m] && rz >= 0 && rz < bg.ext[2][m];
}

__device__ inline bool cell_in_body_region(const BulkGeom& bg, const StepGeom* sgp, int x, int y, int z,
                                           const Grid& g) {
    if (bg.nbox < 0) return in_body_box(*sgp, x, y, z, g);
    bool hit = false;
#pragma unroll
    for (int m = 0; m < kBulkBoxes; ++m)
        if (m < bg.nbox) hit = hit || in_bulk_box(bg, m, x, y, z, g);
    return hit;
}

__device__ inline bool row_in_body_region(const BulkGeom& bg, const StepGeom* sgp, int y, int z, const Grid& g) {
    if (bg.nbox < 0) {
        for (int m = sgp->first_fix_box; m < sgp->nbox; ++m) {
            const Box& b = sgp->box[m];
            if (b.kind != 1) continue;
            int ry = y - b.lo[1], rz = z - b.lo[2];
            if (g.per & 2) ry = ((ry % g.ny) + g.ny) % g.ny;
            if (g.per & 4) rz = ((rz % g.nz) + g.nz) % g.nz;
            if (ry >= 0 && ry < b.ext[1] && rz >= 0 && rz < b.ext[2]) return true;
        }
        return false;
    }
    bool hit = false;
#pragma unroll
    for (int m = 0; m < kBulkBoxes; ++m)
        if (m < bg.nbox) hit = hit || row_in_bulk_box(bg, m, y, z, g);
    return hit;
}

// number of non-periodic faces the cell lies on
__device__ inline int open_face_count(int x, int y, int z, const Grid& g) {
    return (!(g.per & 1) && (x == 0 || x == g.nx - 1)) + (!(g.per & 2) && (y == 0 || y == g.ny - 1)) +
           (!(g.per & 4) && (z == 0 || z == g.nz - 1));
}

// Runtime-indexed reads of the face table through selects: indexing the
// by-value kernel argument directly would copy it to private memory.
__device__ inline int face_kind_at(const Grid& g, int i) {
    int k = g.face_kind[0];
    k = i == 1 ? g.face_kind[1] : k;
    k = i == 2 ? g.face_kind[2] : k;
    k = i == 3 ? g.face_kind[3] : k;
    k = i == 4 ? g.face_kind[4] : k;
    k = i == 5 ? g.face_kind[5] : k;
    return k;
}

__device__ inline double face_u_at(const Grid& g, int i, int c) { return g.face_u[3 * i + c]; }

// Face rules for direction d whose source lies outside face fidx:
// VELOCITY src/OtherBC.jl:143-148, PRESSURE_ABB src/OtherBC.jl:156-165.
// uf: the face's wall / inlet velocity (read by the caller ahead of its loads)
// For the shifted formats (S::kDev) own_opp is a deviation and rho_prev holds
// rho - 1 (own_moments, the face slots): the rule in deviation form,
// delta_in = -delta_opp + 2 w [(rho - 1) + rho (4.5 (c.u)^2 - 1.5 u^2)] for
// anti-bounce-back, the same expression as f for the velocity rule.
template <class S>
__device__ inline typename S::C face_value(int kind, const typename S::C uf[3], int d, typename S::C own_opp,
                                           typename S::C rho_prev, const typename S::C up[3], typename S::C uu2) {
    using C = typename S::C;
    const C w = weight<C>(d);
    const C rho = S::kDev ? C(1) + rho_prev : rho_prev;
    if (kind == LBM_FACE_VELOCITY) {
        const C cu = C(kCx[d]) * (-uf[0]) + C(kCy[d]) * (-uf[1]) + C(kCz[d]) * (-uf[2]);
        return own_opp - C(2) * w * rho * cu * C(3);
    }
    const C cu = C(kCx[d]) * up[0] + C(kCy[d]) * up[1] + C(kCz[d]) * up[2];
    if constexpr (S::kDev)
        return -own_opp + C(2) * w * (rho_prev + rho * (C(0.5) * cu * cu * C(9) - C(0.5) * uu2 * C(3)));
    return -own_opp + C(2) * w * rho * (C(1) + C(0.5) * cu * cu * C(9) - C(0.5) * uu2 * C(3));
}

template <class S, bool FORCED>
__device__ inline void own_moments(const typename S::C own[Q], const Phys<typename S::C>& p, typename S::C& rho,
                                   typename S::C up[3], typename S::C& uu2) {
    using C = typename S::C;
    C jx, jy, jz;
    moments<C>(own, rho, jx, jy, jz);
    // shifted formats: rho holds sum delta; the density is rho_w + sum delta,
    // and rho - 1 is what the rules take (face_value)
    const C dens = S::kDev ? C(kRhoW19) + rho : rho;
    if (S::kDev) rho = C(kRhoW19 - 1.0) + rho;
    up[0] = jx / dens;
    up[1] = jy / dens;
    up[2] = jz / dens;
    if (FORCED) {
        up[0] -= C(0.5) * p.g[0];
        up[1] -= C(0.5) * p.g[1];
        up[2] -= C(0.5) * p.g[2];
    }
    uu2 = up[0] * up[0] + up[1] * up[1] + up[2] * up[2];
}

// ---------------------------------------------------------------------------
// Face moments.  A velocity or anti-bounce-back face rule needs the moments of
// the cell's own f* (src/MainCore.jl:316,336: the previous step's rho, u).
// The step that produces a face cell's f* also stores them in its face slot,
// so the next step reads one slot instead of the cell's 19 populations.
// ---------------------------------------------------------------------------
__device__ inline int face_slot(const Grid& g, int f, int x, int y, int z) {
    const int a = f >> 1;
    const int i = a == 0 ? y + g.ny * z : (a == 1 ? x + g.nx * z : x + g.nx * y);
    return g.face_base[f] + i;
}

__device__ inline bool rule_face(const Grid& g, int f) { return g.face_base[f] >= 0; }

template <typename C>
__device__ inline void side_read(const Grid& g, int slot, C& rho, C up[3], C& uu2) {
    const C* s = reinterpret_cast<const C*>(g.side_in) + 4 * (int64_t)slot;
    rho = s[0];
    up[0] = s[1];
    up[1] = s[2];
    up[2] = s[3];
    uu2 = up[0] * up[0] + up[1] * up[1] + up[2] * up[2];
}

// moments of a finished (post-collision) f* into face slot `slot`
template <class S, bool FORCED>
__device__ inline void side_write(const Grid& g, const typename S::C f[Q], const Phys<typename S::C>& p, int slot) {
    using C = typename S::C;
    C rho, up[3], uu2;
    own_moments<S, FORCED>(f, p, rho, up, uu2);
    C* s = reinterpret_cast<C*>(g.side_out) + 4 * (int64_t)slot;
    s[0] = rho;
    s[1] = up[0];
    s[2] = up[1];
    s[3] = up[2];
}

// every face slot of cell (x, y, z): cells on several rule faces (edges,
// corners) fill each
template <class S, bool FORCED>
__device__ inline void side_write_all(const Grid& g, const typename S::C f[Q], const Phys<typename S::C>& p, int x,
                                      int y, int z) {
    const int q[3] = {x, y, z};
    const int L[3] = {g.nx, g.ny, g.nz};
    for (int a = 0; a < 3; ++a) {
        if (q[a] == 0 && rule_face(g, 2 * a)) side_write<S, FORCED>(g, f, p, face_slot(g, 2 * a, x, y, z));
        if (q[a] == L[a] - 1 && rule_face(g, 2 * a + 1))
            side_write<S, FORCED>(g, f, p, face_slot(g, 2 * a + 1, x, y, z));
    }
}

// finish_cell plus the face slots of the cell (slot >= 0: that one; -1: every
// rule face the cell lies on)
template <class S, int COLL, bool FORCED, int MODE>
__device__ inline void finish_face_cell(typename S::C f[Q], const Phys<typename S::C>& p, int64_t cell, const Grid& g,
                                        typename S::T* __restrict__ fout, double* __restrict__ gout,
                                        double* __restrict__ rho_out, double* __restrict__ u_out, int x, int y, int z,
                                        int slot) {
    finish_cell<S, COLL, FORCED, MODE>(f, p, cell, g.N, g.P, fout, gout, rho_out, u_out, g.sh);
    if (MODE == MODE_STEP || MODE == MODE_COLLIDE_ONLY) {
        if (slot >= 0) side_write<S, FORCED>(g, f, p, slot);
        else side_write_all<S, FORCED>(g, f, p, x, y, z);
    }
}

// ---------------------------------------------------------------------------
// Face cell on exactly one non-periodic y or z face (AXIS, SIDE): directions
// whose source lies outside the face take the face rule (or zero gradient);
// all others are plain pulls with periodic wrap on the remaining axes.
// ---------------------------------------------------------------------------
// Pulled populations of a face cell (the face rule applied); returns its
// face slot (-1 on a zero-gradient face).
template <class S, bool FORCED, int AXIS, int SIDE>
__device__ __forceinline__ int face_pull(const typename S::T* __restrict__ fin, const Grid& g,
                                         const Phys<typename S::C>& p, int x, int y, int z, typename S::C (&f)[Q]) {
    using C = typename S::C;
    const int64_t N = g.N;
    const int nx = g.nx, nxy = (int)g.nxy;
    const int cell = x + nx * y + nxy * z;
    const int fidx = 2 * AXIS + SIDE;
    const int kind = g.face_kind[fidx];
    const int wxm = ((g.per & 1) && x == 0) ? nx : 0, wxp = ((g.per & 1) && x == nx - 1) ? -nx : 0;
    const int wym = (AXIS != 1 && y == 0) ? nxy : 0, wyp = (AXIS != 1 && y == g.ny - 1) ? -nxy : 0;
    const int wzm = (AXIS != 2 && z == 0) ? (int)N : 0, wzp = (AXIS != 2 && z == g.nz - 1) ? -(int)N : 0;
    const int axis_stride = AXIS == 1 ? nx : nxy;
    const bool zg = kind == LBM_FACE_ZERO_GRADIENT;
    const C uf[3] = {C(face_u_at(g, fidx, 0)), C(face_u_at(g, fidx, 1)), C(face_u_at(g, fidx, 2))};
    // previous-step moments of the cell, from its face slot (face rules only)
    const int slot = zg ? -1 : face_slot(g, fidx, x, y, z);
    C rho_prev, up[3], uu2;
    side_read<C>(g, zg ? 0 : slot, rho_prev, up, uu2);  // unconditional: slot 0 always exists
#pragma unroll
    for (int d = 0; d < Q; ++d) {
        const int ca = AXIS == 1 ? kCy[d] : kCz[d];
        const bool out = SIDE == 0 ? (ca > 0) : (ca < 0);
        int src = cell - (kCx[d] + kCy[d] * nx + kCz[d] * nxy);
        if (kCx[d] > 0) src += wxm;
        if (kCx[d] < 0) src += wxp;
        if (kCy[d] > 0) src += wym;
        if (kCy[d] < 0) src += wyp;
        if (kCz[d] > 0) src += wzm;
        if (kCz[d] < 0) src += wzp;
        if (out) {
            // zero gradient: the source clamped back onto the face plane; the
            // other rules start from the cell's own f*[opp(d)].  One load
            // either way (an address select, no branch around the load).
            const typename S::T* a = fin + (int64_t)d * g.P + (src + ca * axis_stride);
            const typename S::T* b = fin + (int64_t)kOpp[d] * g.P + cell;
            const C v = S::load_pick(zg ? a : b, kOpp[d], d, zg, g.sh);
            const C fv = face_value<S>(kind, uf, d, v, rho_prev, up, uu2);
            f[d] = zg ? v : fv;
        } else {
            f[d] = S::load(fin + (int64_t)d * g.P + src, d, g.sh);
        }
    }
    (void)p;
    return slot;
}

// ---------------------------------------------------------------------------
// Bulk pass.  A thread owns V = S::kVec consecutive x cells of one row; a
// block is blockDim.x such groups along x in blockDim.y rows, so every wave
// streams 64 V consecutive cells of one row in each direction plane and the
// +-1 shifted pulls stay inside the block's own cache lines.  Every direction
// is one unconditional load (periodic wraps and zero-gradient clamps are
// address or value selects): a load under a divergent branch would make the
// compiler wait for the loads already in flight.  It finishes every cell
// except
//   * domain edges (cells on two or more non-periodic faces) and x-face cells
//     of a velocity or anti-bounce-back face,
//   * cells the fix-up blocks own: static boundary nodes (flag bit), face
//     cells inside a body region, and every cell of a moving body's region,
// which the fix-up blocks do.  A row on a y or z face takes the face path as
// a whole wave.
// ---------------------------------------------------------------------------

// collide every active cell of a group and store it (one vector store per
// direction when the whole group is active)
template <class S, int COLL, bool FORCED, int MODE, int V>
__device__ inline void finish_group(typename S::C (&f)[V][Q], const bool (&act)[V], const Phys<typename S::C>& p,
                                    int cell0, const Grid& g, typename S::T* __restrict__ fout,
                                    double* __restrict__ gout, double* __restrict__ rho_out,
                                    double* __restrict__ u_out) {
    using C = typename S::C;
    if (MODE == MODE_STEP || MODE == MODE_COLLIDE_ONLY) {
#pragma unroll
        for (int v = 0; v < V; ++v) collide_s<S, COLL, FORCED>(f[v], p);
        bool all = true;
#pragma unroll
        for (int v = 0; v < V; ++v) all = all && act[v];
        if (V == 2 && all) {
            const uint32_t h = sr_cell(cell0, g.sh.seed);
#pragma unroll
            for (int d = 0; d < Q; ++d)
                S::store2_sr(fout + (int64_t)d * g.P + cell0, d, f[0][d], f[V - 1][d], g.sh, sr_dir(h, d));
        } else {
#pragma unroll
            for (int v = 0; v < V; ++v)
                if (act[v]) {
                    const uint32_t h = sr_cell(cell0 + v, g.sh.seed);
#pragma unroll
                    for (int d = 0; d < Q; ++d)
                        S::store_sr(fout + (int64_t)d * g.P + cell0 + v, d, f[v][d], g.sh, sr_dir(h, d));
                }
        }
    } else {
#pragma unroll
        for (int v = 0; v < V; ++v)
            if (act[v])
                finish_cell<S, COLL, FORCED, MODE>(f[v], p, cell0 + v, g.N, g.P, fout, gout, rho_out, u_out, g.sh);
    }
}

// One group on a row of a single y or z face.
template <class S, int COLL, bool FORCED, int MODE, int AXIS, int SIDE>
__device__ __forceinline__ void face_row(const typename S::T* __restrict__ fin, typename S::T* __restrict__ fout,
                                         double* __restrict__ gout, double* __restrict__ rho_out,
                                         double* __restrict__ u_out, const Grid& g, const Phys<typename S::C>& p,
                                         const bool (&act_in)[S::kVec], const uint8_t* fixup_flags, int x0, int y,
                                         int z) {
    using C = typename S::C;
    constexpr int V = S::kVec;
    uint32_t fl = 0;
    if (fixup_flags) fl = V == 2 ? *reinterpret_cast<const uint16_t*>(fixup_flags) : *fixup_flags;
    C f[V][Q];
    int slot[V];
#pragma unroll
    for (int v = 0; v < V; ++v) slot[v] = face_pull<S, FORCED, AXIS, SIDE>(fin, g, p, x0 + v, y, z, f[v]);
    bool act[V];
#pragma unroll
    for (int v = 0; v < V; ++v) act[v] = act_in[v] && !((fl >> (8 * v)) & kFlagFixup);
    finish_group<S, COLL, FORCED, MODE, V>(f, act, p, g.nx * y + (int)g.nxy * z + x0, g, fout, gout, rho_out, u_out);
    if (MODE == MODE_STEP || MODE == MODE_COLLIDE_ONLY) {
#pragma unroll
        for (int v = 0; v < V; ++v)
            if (act[v] && slot[v] >= 0) side_write<S, FORCED>(g, f[v], p, slot[v]);
    }
}

// One pair of cells on a row of a single y or z velocity / anti-bounce-back
// face (16-bit storage): the face_pull of both cells with one pair load per
// direction, as interior_row loads them.  Directions leaving the face take
// the rule from the cells' own f*[opp(d)] (an aligned pair) and their face
// slots; the others pull, with the x shift of interior_row (periodic x: the
// element across the wrap is one extra load, selected by value).  The x-face
// cells of such a row are domain edges and inactive, so no x rule applies.
template <class S, int COLL, bool FORCED, int MODE, int AXIS, int SIDE>
__device__ __forceinline__ void face_row_pair(const typename S::T* __restrict__ fin, typename S::T* __restrict__ fout,
                                              double* __restrict__ gout, double* __restrict__ rho_out,
                                              double* __restrict__ u_out, const Grid& g, const Phys<typename S::C>& p,
                                              const bool (&act_in)[2], const uint8_t* fixup_flags, int x0, int y,
                                              int z) {
    using C = typename S::C;
    uint32_t fl = 0;
    if (fixup_flags) fl = *reinterpret_cast<const uint16_t*>(fixup_flags);
    const int nx = g.nx, nxy = (int)g.nxy;
    const int row = nx * y + nxy * z;
    const int fidx = 2 * AXIS + SIDE;
    const int kind = g.face_kind[fidx];
    const bool px = g.per & 1;
    const bool left = x0 == 0, right = x0 + 2 == nx;
    // periodic wraps of the axis along the face (zero when not periodic)
    const int wym = (AXIS != 1 && y == 0) ? nxy : 0, wyp = (AXIS != 1 && y == g.ny - 1) ? -nxy : 0;
    const int wzm = (AXIS != 2 && z == 0) ? (int)g.N : 0, wzp = (AXIS != 2 && z == g.nz - 1) ? -(int)g.N : 0;
    const int wxm = left ? (px ? nx : 1) : 0, wxp = right ? (px ? -nx : -1) : 0;
    const C uf[3] = {C(face_u_at(g, fidx, 0)), C(face_u_at(g, fidx, 1)), C(face_u_at(g, fidx, 2))};
    const int slot0 = face_slot(g, fidx, x0, y, z);
    C rho_prev[2], up[2][3], uu2[2];
    side_read<C>(g, slot0, rho_prev[0], up[0], uu2[0]);
    side_read<C>(g, slot0 + 1, rho_prev[1], up[1], uu2[1]);
    C f[2][Q];
#pragma unroll
    for (int d = 0; d < Q; ++d) {
        const int ca = AXIS == 1 ? kCy[d] : kCz[d];
        const bool out = SIDE == 0 ? (ca > 0) : (ca < 0);
        if (out) {
            C a, b;
            S::load2(fin + (int64_t)kOpp[d] * g.P + row + x0, kOpp[d], a, b, g.sh);
            f[0][d] = a;
            f[1][d] = b;
            continue;
        }
        int srow = row - (kCy[d] * nx + kCz[d] * nxy);
        if (kCy[d] > 0) srow += wym;
        if (kCy[d] < 0) srow += wyp;
        if (kCz[d] > 0) srow += wzm;
        if (kCz[d] < 0) srow += wzp;
        const typename S::T* plane = fin + (int64_t)d * g.P + srow;
        if (kCx[d] == 0) {
            S::load2(plane + x0, d, f[0][d], f[1][d], g.sh);
        } else if (kCx[d] > 0) {
            C a, b;
            S::load2(left ? plane + x0 : plane + x0 - 1, d, a, b, g.sh);
            f[1][d] = left ? a : b;
            const C w = S::load(plane + x0 - 1 + wxm, d, g.sh);
            f[0][d] = px ? w : a;
        } else {
            C a, b;
            S::load2(right ? plane + x0 : plane + x0 + 1, d, a, b, g.sh);
            f[0][d] = right ? b : a;
            const C w = S::load(plane + x0 + 2 + wxp, d, g.sh);
            f[1][d] = px ? w : b;
        }
    }
    // the rules (ALU only: every load above is issued)
#pragma unroll
    for (int d = 0; d < Q; ++d) {
        const int ca = AXIS == 1 ? kCy[d] : kCz[d];
        const bool out = SIDE == 0 ? (ca > 0) : (ca < 0);
        if (!out) continue;
#pragma unroll
        for (int v = 0; v < 2; ++v) f[v][d] = face_value<S>(kind, uf, d, f[v][d], rho_prev[v], up[v], uu2[v]);
    }
    bool act[2];
#pragma unroll
    for (int v = 0; v < 2; ++v) act[v] = act_in[v] && !((fl >> (8 * v)) & kFlagFixup);
    finish_group<S, COLL, FORCED, MODE, 2>(f, act, p, row + x0, g, fout, gout, rho_out, u_out);
    if (MODE == MODE_STEP || MODE == MODE_COLLIDE_ONLY) {
#pragma unroll
        for (int v = 0; v < 2; ++v)
            if (act[v]) side_write<S, FORCED>(g, f[v], p, slot0 + v);
    }
}

// x treatment of an interior row, a wave-uniform choice between branch-free
// copies of the pull: periodic; zero-gradient faces only (clamps are value
// selects); at least one velocity / anti-bounce-back face (its rule starts
// from the face cell's own f*[opp(d)], one extra load per direction pair
// that every other lane points at an element it already reads).
enum { XM_PERIODIC = 0, XM_CLAMP = 1, XM_RULE = 2 };

// Pull + collide of one group on a row off the y and z faces.
template <class S, int COLL, bool FORCED, int MODE, int XM>
__device__ __forceinline__ void interior_row(const typename S::T* __restrict__ fin, typename S::T* __restrict__ fout,
                                             double* __restrict__ gout, double* __restrict__ rho_out,
                                             double* __restrict__ u_out, const Grid& g,
                                             const Phys<typename S::C>& p, const bool (&act_in)[S::kVec],
                                             const uint8_t* fixup_flags, int x0, int y, int z) {
    using C = typename S::C;
    constexpr int V = S::kVec;
    constexpr bool PX = XM == XM_PERIODIC;
    // fix-up bits of static boundary nodes, loaded together with the populations
    uint32_t fl = 0;
    if (fixup_flags) fl = V == 2 ? *reinterpret_cast<const uint16_t*>(fixup_flags) : *fixup_flags;
    const int nx = g.nx, nxy = (int)g.nxy;
    const int row = nx * y + nxy * z;
    const bool left = x0 == 0, right = x0 + V == nx;
    // lanes applying a face rule at x = 0 (rl) or x = nx - 1 (rr)
    const bool rl = XM == XM_RULE && left && rule_face(g, 0);
    const bool rr = XM == XM_RULE && right && rule_face(g, 1);
    // periodic y/z wraps by index arithmetic (zero on non-periodic axes)
    const int wym = (y == 0) ? nxy : 0, wyp = (y == g.ny - 1) ? -nxy : 0;
    const int wzm = (z == 0) ? (int)g.N : 0, wzp = (z == g.nz - 1) ? -(int)g.N : 0;
    // x: the pull source of x = 0 (c_x > 0) or x = nx - 1 (c_x < 0) wraps
    // (periodic) or is clamped onto the face plane (zero gradient; a rule
    // face replaces the value)
    const int wxm = left ? (PX ? nx : 1) : 0, wxp = right ? (PX ? -nx : -1) : 0;
    C rho_prev = C(0), up[3] = {C(0), C(0), C(0)}, uu2 = C(0);
    C uf_lo[3] = {C(0), C(0), C(0)}, uf_hi[3] = {C(0), C(0), C(0)};
    int slot = -1;
    if (XM == XM_RULE) {
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            uf_lo[c] = C(face_u_at(g, 0, c));
            uf_hi[c] = C(face_u_at(g, 1, c));
        }
        // every lane reads a valid slot; only rl / rr lanes use it
        slot = rule_face(g, 0) ? face_slot(g, 0, 0, y, z) : face_slot(g, 1, nx - 1, y, z);
        if (rr && rule_face(g, 0)) slot = face_slot(g, 1, nx - 1, y, z);
        side_read<C>(g, slot, rho_prev, up, uu2);
    }
    C f[V][Q];
    C ex[5];  // own f*[opp(d)] of a rule lane, per c_x > 0 direction pair
#pragma unroll
    for (int d = 0; d < Q; ++d) {
        int srow = row - (kCy[d] * nx + kCz[d] * nxy);
        if (kCy[d] > 0) srow += wym;
        if (kCy[d] < 0) srow += wyp;
        if (kCz[d] > 0) srow += wzm;
        if (kCz[d] < 0) srow += wzp;
        const typename S::T* plane = fin + (int64_t)d * g.P + srow;
        if (V == 1) {
            int sx = x0 - kCx[d];
            if (kCx[d] > 0) sx += wxm;
            if (kCx[d] < 0) sx += wxp;
            f[0][d] = S::load(plane + sx, d, g.sh);
        } else if (kCx[d] == 0) {
            S::load2(plane + x0, d, f[0][d], f[V - 1][d], g.sh);
        } else if (kCx[d] > 0) {
            // pair (x0 - 1, x0); the left group loads (0, 1) instead, so no
            // load leaves the row: x = -1 is a wrap, a clamp onto x = 0 or a
            // face rule
            C a, b;
            S::load2(left ? plane + x0 : plane + x0 - 1, d, a, b, g.sh);
            f[V - 1][d] = left ? a : b;
            f[0][d] = PX ? S::load(plane + x0 - 1 + wxm, d, g.sh) : a;  // the same element as a off the left group
        } else {
            // pair (x0 + 1, x0 + 2); the right group loads (nx - 2, nx - 1)
            C a, b;
            S::load2(right ? plane + x0 : plane + x0 + 1, d, a, b, g.sh);
            f[0][d] = right ? b : a;
            f[V - 1][d] = PX ? S::load(plane + x0 + 2 + wxp, d, g.sh) : b;
        }
        if (XM == XM_RULE && kCx[d] > 0) {
            // d leaves x = 0 inward (rule at the left face reads f*[opp d] at x = 0);
            // opp(d) leaves x = nx - 1 (rule at the right face reads f*[d] there);
            // other lanes read an element of their own pair
            const typename S::T* own_l = fin + (int64_t)kOpp[d] * g.P + row;
            const typename S::T* own_r = fin + (int64_t)d * g.P + row + nx - 1;
            const typename S::T* dflt = plane + x0;
            const int k = (d == 0) ? 0 : (d == 6 ? 1 : (d == 8 ? 2 : (d == 12 ? 3 : 4)));
            ex[k] = S::load_pick(rl ? own_l : (rr ? own_r : dflt), d, kOpp[d], rl, g.sh);
        }
    }
    if (XM == XM_RULE) {
        // ALU-only branches: the loads above are already issued
        if (rl) {
            const int kind = g.face_kind[0];
#pragma unroll
            for (int d = 0; d < Q; ++d) {
                if (kCx[d] <= 0) continue;
                const int k = (d == 0) ? 0 : (d == 6 ? 1 : (d == 8 ? 2 : (d == 12 ? 3 : 4)));
                f[0][d] = face_value<S>(kind, uf_lo, d, ex[k], rho_prev, up, uu2);
            }
        }
        if (rr) {
            const int kind = g.face_kind[1];
#pragma unroll
            for (int d = 0; d < Q; ++d) {
                if (kCx[d] >= 0) continue;
                const int k = (d == 1) ? 0 : (d == 7 ? 1 : (d == 9 ? 2 : (d == 13 ? 3 : 4)));
                f[V - 1][d] = face_value<S>(kind, uf_hi, d, ex[k], rho_prev, up, uu2);
            }
        }
    }
    bool act[V];
#pragma unroll
    for (int v = 0; v < V; ++v) act[v] = act_in[v] && !((fl >> (8 * v)) & kFlagFixup);
    finish_group<S, COLL, FORCED, MODE, V>(f, act, p, row + x0, g, fout, gout, rho_out, u_out);
    if (XM == XM_RULE && (MODE == MODE_STEP || MODE == MODE_COLLIDE_ONLY)) {
        if (rl && act[0]) side_write<S, FORCED>(g, f[0], p, face_slot(g, 0, 0, y, z));
        if (rr && act[V - 1]) side_write<S, FORCED>(g, f[V - 1], p, face_slot(g, 1, nx - 1, y, z));
    }
}

template <class S, int COLL, bool FORCED, int MODE>
__device__ __forceinline__ void bulk_group(const typename S::T* __restrict__ fin, typename S::T* __restrict__ fout,
                                           double* __restrict__ gout, double* __restrict__ rho_out,
                                           double* __restrict__ u_out, const Grid& g, const Phys<typename S::C>& p,
                                           const StepGeom* __restrict__ sgp, const uint8_t* __restrict__ flags,
                                           int x0, int y, int z) {
    using C = typename S::C;
    constexpr int V = S::kVec;  // V == 2 requires an even nx (checked at context creation)
    if (x0 >= g.nx || y >= g.ny) return;
    const int nx = g.nx, nxy = (int)g.nxy;
    const int row = nx * y + nxy * z;
    const int cell0 = row + x0;
    bool act[V];
#pragma unroll
    for (int v = 0; v < V; ++v) act[v] = true;
    if (MODE == MODE_COLLIDE_ONLY || MODE == MODE_GET_RAW) {
        C f[V][Q];
#pragma unroll
        for (int d = 0; d < Q; ++d) {
            if (V == 2) S::load2(fin + (int64_t)d * g.P + cell0, d, f[0][d], f[V - 1][d], g.sh);
            else f[0][d] = S::load(fin + (int64_t)d * g.P + cell0, d, g.sh);
        }
        finish_group<S, COLL, FORCED, MODE, V>(f, act, p, cell0, g, fout, gout, rho_out, u_out);
        if (MODE == MODE_COLLIDE_ONLY) {
#pragma unroll
            for (int v = 0; v < V; ++v) side_write_all<S, FORCED>(g, f[v], p, x0 + v, y, z);
        }
        return;
    }
    const bool px = g.per & 1;
    const bool ylo = !(g.per & 2) && y == 0, yhi = !(g.per & 2) && y == g.ny - 1;
    const bool zlo = !(g.per & 4) && z == 0, zhi = !(g.per & 4) && z == g.nz - 1;
    const int nyz = (ylo || yhi) + (zlo || zhi);
    if (nyz >= 2) return;  // domain edge row
    const bool left = x0 == 0, right = x0 + V == nx;  // the group holds x = 0 / x = nx - 1
    const uint8_t* fixup_flags = nullptr;
    // x-face cells of a face row are domain edges
    if (!px && nyz == 1) {
        if (left) act[0] = false;
        if (right) act[V - 1] = false;
    }
    const int bmode = sgp ? sgp->bg.mode : 0, bnb = sgp ? sgp->bg.nbodies : 0;
    if (bmode == 1) {
        const BulkGeom& bg = sgp->bg;
        // moving bodies: the fix-up blocks own their whole regions
        if (row_in_body_region(bg, sgp, y, z, g)) {
#pragma unroll
            for (int v = 0; v < V; ++v)
                if (cell_in_body_region(bg, sgp, x0 + v, y, z, g)) act[v] = false;
        }
    } else if (bnb > 0) {
        const BulkGeom& bg = sgp->bg;
        // static bodies: boundary nodes carry the fix-up bit (read with the
        // populations, below), face cells of a body region belong to the fix-up
        fixup_flags = flags + cell0;
#pragma unroll
        for (int v = 0; v < V; ++v) {
            const bool xf = !px && ((v == 0 && left) || (v == V - 1 && right));
            if ((nyz == 1 || xf) && cell_in_body_region(bg, sgp, x0 + v, y, z, g)) act[v] = false;
        }
    }
    bool any = false;
#pragma unroll
    for (int v = 0; v < V; ++v) any = any || act[v];
    if (!any) return;
    if (nyz == 1) {
        // a row on one y or z face: the whole wave takes the face path (16-bit
        // pairs on a velocity / anti-bounce-back face: the pair-load copy)
        if constexpr (V == 2) {
            const int fidx = ylo ? 2 : (yhi ? 3 : (zlo ? 4 : 5));
            if (rule_face(g, fidx)) {
                if (ylo)
                    face_row_pair<S, COLL, FORCED, MODE, 1, 0>(fin, fout, gout, rho_out, u_out, g, p, act, fixup_flags,
                                                               x0, y, z);
                else if (yhi)
                    face_row_pair<S, COLL, FORCED, MODE, 1, 1>(fin, fout, gout, rho_out, u_out, g, p, act, fixup_flags,
                                                               x0, y, z);
                else if (zlo)
                    face_row_pair<S, COLL, FORCED, MODE, 2, 0>(fin, fout, gout, rho_out, u_out, g, p, act, fixup_flags,
                                                               x0, y, z);
                else
                    face_row_pair<S, COLL, FORCED, MODE, 2, 1>(fin, fout, gout, rho_out, u_out, g, p, act, fixup_flags,
                                                               x0, y, z);
                return;
            }
        }
        if (ylo) face_row<S, COLL, FORCED, MODE, 1, 0>(fin, fout, gout, rho_out, u_out, g, p, act, fixup_flags, x0, y, z);
        else if (yhi)
            face_row<S, COLL, FORCED, MODE, 1, 1>(fin, fout, gout, rho_out, u_out, g, p, act, fixup_flags, x0, y, z);
        else if (zlo)
            face_row<S, COLL, FORCED, MODE, 2, 0>(fin, fout, gout, rho_out, u_out, g, p, act, fixup_flags, x0, y, z);
        else face_row<S, COLL, FORCED, MODE, 2, 1>(fin, fout, gout, rho_out, u_out, g, p, act, fixup_flags, x0, y, z);
        return;
    }
    // interior row; the x treatment is a wave-uniform choice between
    // branch-free copies of the pull.  A wave lies in one row (blockDim.x is a
    // multiple of 64), so its first lane tells whether it holds x = 0 or
    // x = nx - 1: only those waves pay for the face-rule loads.
    const int wx0 = __builtin_amdgcn_readfirstlane(x0);
    const bool wave_left = wx0 == 0, wave_right = wx0 + 64 * V >= nx;
    if (px)
        interior_row<S, COLL, FORCED, MODE, XM_PERIODIC>(fin, fout, gout, rho_out, u_out, g, p, act, fixup_flags, x0,
                                                         y, z);
    else if ((wave_left && rule_face(g, 0)) || (wave_right && rule_face(g, 1)))
        interior_row<S, COLL, FORCED, MODE, XM_RULE>(fin, fout, gout, rho_out, u_out, g, p, act, fixup_flags, x0, y, z);
    else
        interior_row<S, COLL, FORCED, MODE, XM_CLAMP>(fin, fout, gout, rho_out, u_out, g, p, act, fixup_flags, x0, y,
                                                      z);
}

template <class S, int COLL, bool FORCED, int MODE>
__global__ __launch_bounds__(kBlock) void k_bulk(const typename S::T* __restrict__ fin,
                                                 typename S::T* __restrict__ fout, double* __restrict__ gout,
                                                 double* __restrict__ rho_out, double* __restrict__ u_out, Grid g,
                                                 Phys<typename S::C> p, const StepGeom* __restrict__ sgp,
                                                 const uint8_t* __restrict__ flags) {
    bulk_group<S, COLL, FORCED, MODE>(fin, fout, gout, rho_out, u_out, g, p, sgp, flags,
                                      (blockIdx.x * blockDim.x + threadIdx.x) * S::kVec,
                                      blockIdx.y * blockDim.y + threadIdx.y, blockIdx.z);
}

// ---------------------------------------------------------------------------
// Force accumulation: per thread one body in registers; a link of a fluid node
// to a second body goes into its wave's own LDS slots (wave_slot_add); per
// wave a deterministic segmented shuffle reduction into the same slots; per
// block a fixed-order sum over the waves.  No two waves share a slot, one
// wave's LDS operations complete in program order, and within one atomic
// instruction the LDS applies the lanes that hit one address in a fixed
// order, so the force of a step is bitwise reproducible.
// ---------------------------------------------------------------------------
struct ForceAcc {
    int b0;
    double a0[6];
};

__device__ inline double wave_sum(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// v[6] of a link to body `key` into this wave's slots (see above)
__device__ __forceinline__ void wave_slot_add(int key, const double v[6], double* lds_wave) {
#pragma unroll
    for (int c = 0; c < 6; ++c) atomicAdd(&lds_wave[key * 6 + c], v[c]);
}

__device__ inline void wave_reduce_to_lds(int key, const double vals[6], double* lds_wave /*[nb][6]*/) {
    const int lane = threadIdx.x & 63;
    uint64_t pending = __ballot(key >= 0);
    while (pending) {
        const int leader = __ffsll((unsigned long long)pending) - 1;
        const int k = __shfl(key, leader, 64);
        const bool mine = (key == k);
        const uint64_t sel = __ballot(mine);
#pragma unroll
        for (int c = 0; c < 6; ++c) {
            double s = wave_sum(mine ? vals[c] : 0.0);
            if (lane == 0) lds_wave[k * 6 + c] += s;
        }
        pending &= ~sel;
    }
}

// ---------------------------------------------------------------------------
// Plain cell off every non-periodic face: the k_bulk pull, for cells of a
// moving-body region that are neither boundary nor fresh nodes.
// ---------------------------------------------------------------------------
template <class S, int COLL, bool FORCED, int MODE>
__device__ __forceinline__ void plain_cell(const typename S::T* __restrict__ fin, typename S::T* __restrict__ fout,
                                           double* __restrict__ gout, double* __restrict__ rho_out,
                                           double* __restrict__ u_out, const Grid& g, const Phys<typename S::C>& p,
                                           int x, int y, int z) {
    using C = typename S::C;
    const int nx = g.nx, nxy = (int)g.nxy;
    const int64_t N = g.N;
    const int cell = x + nx * y + nxy * z;
    const int wxm = (x == 0) ? nx : 0, wxp = (x == nx - 1) ? -nx : 0;
    const int wym = (y == 0) ? nxy : 0, wyp = (y == g.ny - 1) ? -nxy : 0;
    const int wzm = (z == 0) ? (int)N : 0, wzp = (z == g.nz - 1) ? -(int)N : 0;
    C f[Q];
#pragma unroll
    for (int d = 0; d < Q; ++d) {
        int src = cell - (kCx[d] + kCy[d] * nx + kCz[d] * nxy);
        if (kCx[d] > 0) src += wxm;
        if (kCx[d] < 0) src += wxp;
        if (kCy[d] > 0) src += wym;
        if (kCy[d] < 0) src += wyp;
        if (kCz[d] > 0) src += wzm;
        if (kCz[d] < 0) src += wzp;
        f[d] = S::load(fin + (int64_t)d * g.P + src, d, g.sh);
    }
    finish_cell<S, COLL, FORCED, MODE>(f, p, cell, N, g.P, fout, gout, rho_out, u_out, g.sh);
}

// ---------------------------------------------------------------------------
// The general cell: faces (any number, with precedence), bounce-back links,
// refill, momentum exchange.  The cell's previous-step moments come first
// from its own f*; then each direction is classified and its value loaded or
// formed in one unrolled pass, so only the 19 pulled values stay live.
// ---------------------------------------------------------------------------
enum { DIR_PULL = 0, DIR_FACE = 1, DIR_LINK = 2 };

// Link codes of the 18 pull sources, 10 bits per direction (kind, body).
struct LinkCodes {
    uint32_t w[(Q + 2) / 3];
};

// Solidity of a cell and of its 18 neighbours against the new pose of
// analytic bodies, with inside_body's FP32 operations at every neighbour:
// per body, the three offsets of each axis (neighbour coordinate + 1/2 minus
// the centre, minimum image) are formed once and combined per neighbour, and a
// body out of reach of every neighbour is skipped.  Sets the link codes of the
// in-domain pull sources x - c_d (the first body containing a source wins, as
// in flag_of) and returns the cell's own body + 1 (0: fluid).
__device__ inline int link_codes_moving(const SolidView& sv, int x, int y, int z, const Grid& g, LinkCodes& lc) {
#pragma clang fp contract(off)
    const int q[3] = {x, y, z};
    const int L[3] = {g.nx, g.ny, g.nz};
    int qq[3][3];
    uint32_t valid = 0u;  // bit 3a + k: neighbour offset k - 1 along axis a is in the domain
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            qq[a][k] = q[a] + k - 1;
            if (wrap_coord(qq[a][k], L[a], (g.per >> a) & 1)) valid |= 1u << (3 * a + k);
        }
    uint32_t taken = 0u;  // bit e: neighbour x + c_e already inside an earlier body
    int self = 0;
    for (int b = 0; b < sv.nb; ++b) {
        const DevBody& B = sv.bnew[b];
        float dv[3][3];
#pragma unroll
        for (int a = 0; a < 3; ++a)
#pragma unroll
            for (int k = 0; k < 3; ++k) dv[a][k] = min_image((float)qq[a][k] + 0.5f - B.c[a], (g.per >> a) & 1, L[a]);
        if (fabsf(dv[0][1]) > B.ext[0] || fabsf(dv[1][1]) > B.ext[1] || fabsf(dv[2][1]) > B.ext[2]) continue;
        uint32_t hit = 0u;
        if (B.shape == LBM_SPHERE || B.shape == LBM_CYLINDER) {
            float s2[3][3];
#pragma unroll
            for (int a = 0; a < 3; ++a)
#pragma unroll
                for (int k = 0; k < 3; ++k) s2[a][k] = dv[a][k] * dv[a][k];
            const float thr = B.r2 + 1e-6f * B.r2;
            const bool sph = B.shape == LBM_SPHERE;
#pragma unroll 1
            for (int e = 0; e < 19; ++e) {
                float t = s2[0][c_at(0, e) + 1] + s2[1][c_at(1, e) + 1];
                if (sph) t = t + s2[2][c_at(2, e) + 1];
                hit |= (t <= thr ? 1u : 0u) << e;
                __builtin_amdgcn_sched_barrier(0);
            }
        } else if (B.shape == LBM_ELLIPSOID) {
#pragma unroll 1
            for (int e = 0; e < 19; ++e) {
                const float dx = dv[0][c_at(0, e) + 1], dy = dv[1][c_at(1, e) + 1], dz = dv[2][c_at(2, e) + 1];
                float xb = B.rot[0] * dx + B.rot[3] * dy;
                xb = xb + B.rot[6] * dz;
                float yb = B.rot[1] * dx + B.rot[4] * dy;
                yb = yb + B.rot[7] * dz;
                float zb = B.rot[2] * dx + B.rot[5] * dy;
                zb = zb + B.rot[8] * dz;
                float qe = xb * xb * B.ia2[0] + yb * yb * B.ia2[1];
                qe = qe + zb * zb * B.ia2[2];
                hit |= (qe <= 1.0f + 1e-6f ? 1u : 0u) << e;
                __builtin_amdgcn_sched_barrier(0);
            }
        } else {
#pragma unroll 1
            for (int e = 0; e < 19; ++e)
                if (inside_naca(B, dv[0][c_at(0, e) + 1], dv[1][c_at(1, e) + 1])) hit |= 1u << e;
        }
        hit &= ~taken;
        taken |= hit;
        if ((hit >> 18) & 1u) self = b + 1;
        // pull direction d reads the neighbour e = opp(d)
#pragma unroll
        for (int d = 0; d < 18; ++d) {
            const int e = kOpp[d];
            if ((hit >> e) & 1u) lc.w[d / 3] |= (uint32_t)(DIR_LINK | (b << 2)) << (10 * (d % 3));
        }
    }
    // drop the sources outside a non-periodic face
#pragma unroll
    for (int d = 0; d < 18; ++d) {
        const int e = kOpp[d];
        const bool ok = ((valid >> (kCx[e] + 1)) & (valid >> (3 + kCy[e] + 1)) & (valid >> (6 + kCz[e] + 1)) & 1u) != 0u;
        if (!ok) lc.w[d / 3] &= ~(0x3ffu << (10 * (d % 3)));
    }
    return self;
}

// The same from the flags (static bodies): the link codes of the in-domain
// pull sources and the cell's own body + 1.
__device__ inline int link_codes_flags(const SolidView& sv, int x, int y, int z, const Grid& g, LinkCodes& lc) {
#pragma unroll
    for (int d = 0; d < 18; ++d) {
        int sx = x - kCx[d], sy = y - kCy[d], sz = z - kCz[d];
        const bool okx = wrap_coord(sx, g.nx, g.per & 1), oky = wrap_coord(sy, g.ny, g.per & 2),
                   okz = wrap_coord(sz, g.nz, g.per & 4);
        const bool ok = okx && oky && okz;
        // outside a non-periodic face: read the cell itself and drop it
        const int sbr = solid_at(sv, ok ? sx : x, ok ? sy : y, ok ? sz : z, g);
        const int sb = ok ? sbr : 0;
        if (sb != 0) lc.w[d / 3] |= (uint32_t)(DIR_LINK | ((sb - 1) << 2)) << (10 * (d % 3));
    }
    return solid_at(sv, x, y, z, g);
}

template <class S, int COLL, bool FORCED, int MODE, bool IBB = false>
__device__ __forceinline__ void general_cell(const typename S::T* __restrict__ fin, typename S::T* __restrict__ fout,
                             double* __restrict__ gout, double* __restrict__ rho_out, double* __restrict__ u_out,
                             const Grid& g, const Phys<typename S::C>& p, const SolidView& sv, int x, int y, int z,
                             ForceAcc& acc, double* lds_wave, const double* lds_face_u, int pre_self,
                             const LinkCodes pre, float* lds_q = nullptr) {
    using C = typename S::C;
    const int64_t cell = (int64_t)x + (int64_t)g.nx * y + g.nxy * z;
    const int L[3] = {g.nx, g.ny, g.nz};
    const int q0[3] = {x, y, z};
    // link codes of the pull sources: from the caller (moving bodies: the
    // analytic solidity it computed), or below from the flags (static bodies)
    uint32_t codes[(Q + 2) / 3];
#pragma unroll
    for (int i = 0; i < (Q + 2) / 3; ++i) codes[i] = pre.w[i];
    // solid cells are streamed like fluid ones (no links), as in the reference
    const bool self_solid = sv.nb > 0 && pre_self != 0;
    const bool fresh = sv.nb > 0 && !self_solid && g.refill && sv.mode == 1 &&
                       flag_of(sv.bold, sv.nb, x, y, z, g) != 0;
    C f[Q];
    if (fresh) {
        // equilibrium of the mean previous-step rho, u of fluid neighbours
        C sr = C(0), su = C(0), sv_ = C(0), sw = C(0);
        int cnt = 0;
#pragma unroll 3
        for (int d = 0; d < 18; ++d) {
            int q[3] = {x + c_at(0, d), y + c_at(1, d), z + c_at(2, d)};
            bool ok = true;
#pragma unroll
            for (int a = 0; a < 3; ++a) ok = ok && wrap_coord(q[a], L[a], (g.per >> a) & 1);
            if (!ok || solid_at(sv, q[0], q[1], q[2], g) != 0 || fresh_at(sv, q[0], q[1], q[2], g)) continue;
            const int64_t nc = (int64_t)q[0] + (int64_t)g.nx * q[1] + g.nxy * q[2];
            C own[Q];
#pragma unroll
            for (int e = 0; e < Q; ++e) own[e] = S::load(fin + (int64_t)e * g.P + nc, e, g.sh);
            C r, mx, my, mz;
            moments<C>(own, r, mx, my, mz);
            // shifted formats: r is sum delta; average rho - 1 there (precise), rho elsewhere
            const C dens = S::kDev ? C(kRhoW19) + r : r;
            if (S::kDev) r = C(kRhoW19 - 1.0) + r;
            C ux = mx / dens, uy = my / dens, uz = mz / dens;
            if (FORCED) {
                ux -= C(0.5) * p.g[0];
                uy -= C(0.5) * p.g[1];
                uz -= C(0.5) * p.g[2];
            }
            sr += r;
            su += ux;
            sv_ += uy;
            sw += uz;
            ++cnt;
        }
        C r = S::kDev ? C(g.rho0 - 1.0) : C(g.rho0), ux = C(0), uy = C(0), uz = C(0);
        if (cnt > 0) {
            const C n = C(cnt);
            r = sr / n;
            ux = su / n;
            uy = sv_ / n;
            uz = sw / n;
        }
        const C u2 = ux * ux + uy * uy + uz * uz;
#pragma unroll
        for (int d = 0; d < Q; ++d) {
            const C cu = C(kCx[d]) * ux + C(kCy[d]) * uy + C(kCz[d]) * uz;
            if constexpr (S::kDev)  // f_eq - w with r = rho - 1
                f[d] = weight<C>(d) * (r + (C(1) + r) * (cu * C(3) + C(0.5) * cu * cu * C(9) - C(0.5) * u2 * C(3)));
            else
                f[d] = weight<C>(d) * r * (C(1) + cu * C(3) + C(0.5) * cu * cu * C(9) - C(0.5) * u2 * C(3));
        }
        finish_face_cell<S, COLL, FORCED, MODE>(f, p, cell, g, fout, gout, rho_out, u_out, x, y, z, -1);
        return;
    }
    // previous-step moments of the cell (src/MainCore.jl:316,336) and the
    // solidity of every in-domain pull source x - c_d, in one memory round
    // trip; the body index of a link source is packed 10 bits per direction
    C rho_prev, up[3], uu2;
    if (self_solid) {
#pragma unroll
        for (int i = 0; i < (Q + 2) / 3; ++i) codes[i] = 0u;
    }
    // a cell on a velocity or anti-bounce-back face reads the moments its face
    // slot received when its f* was produced (one load, which then overlaps the
    // pulls below); any other cell forms them from its own 19 populations
    int slot = -1;
#pragma unroll
    for (int a = 2; a >= 0; --a) {
        if ((g.per >> a) & 1) continue;
        if (q0[a] == 0 && rule_face(g, 2 * a)) slot = face_slot(g, 2 * a, x, y, z);
        if (q0[a] == L[a] - 1 && rule_face(g, 2 * a + 1)) slot = face_slot(g, 2 * a + 1, x, y, z);
    }
    if (slot >= 0) {
        side_read<C>(g, slot, rho_prev, up, uu2);
    } else {
        C own[Q];
#pragma unroll
        for (int d = 0; d < Q; ++d) own[d] = S::load(fin + (int64_t)d * g.P + cell, d, g.sh);
        own_moments<S, FORCED>(own, p, rho_prev, up, uu2);
    }
    __builtin_amdgcn_sched_barrier(0);
    // interpolated bounce-back: wall fraction q and rule of every link from the
    // body's level set, in a loop that is not unrolled (few values are live
    // here); rule 1: q < 1/2 with a fluid second node x - c_dl; rule 2: q >= 1/2;
    // rule 0: no sign change, or no second node (the simple bounce-back)
    // q is kept in LDS (lds_q[d * kBlock + thread]) rather than registers
    float* qv = lds_q + threadIdx.x;
    uint64_t rules = 0;
    if (IBB) {
#pragma unroll 1
        for (int d = 0; d < 18; ++d) {
            if (((codes[d / 3] >> (10 * (d % 3))) & 3u) != DIR_LINK) continue;
            const int dl = d ^ 1;  // opposite of a moving direction
            const DevBody& B = sv.bnew[(codes[d / 3] >> (10 * (d % 3) + 2)) & 0xffu];
            const int cx = c_at(0, dl), cy = c_at(1, dl), cz = c_at(2, dl);
            const float pbx = (float)x + 0.5f, pby = (float)y + 0.5f, pbz = (float)z + 0.5f;
            const float phb = body_phi(B, pbx, pby, pbz, g);
            const float phs = body_phi(B, pbx + (float)cx, pby + (float)cy, pbz + (float)cz, g);
            const float den = phb - phs;
            if (!(phb > 0.0f && phs < 0.0f && fabsf(den) >= 1e-8f)) continue;
            const float q = fminf(fmaxf(phb / den, kEpsQ), 1.0f - kEpsQ);
            int rule = 2;
            if (q < 0.5f) {
                int s2[3] = {x - cx, y - cy, z - cz};
                bool ok = true;
                for (int a = 0; a < 3; ++a) ok = ok && wrap_coord(s2[a], L[a], (g.per >> a) & 1);
                rule = ok && solid_at(sv, s2[0], s2[1], s2[2], g) == 0 ? 1 : 0;
            }
            if (rule != 0) {
                qv[d * kBlock] = q;
                rules |= (uint64_t)rule << (2 * d);
            }
        }
    }
    __builtin_amdgcn_sched_barrier(0);
    // 1) faces (with precedence) and the plain pulls
#pragma unroll
    for (int d = 0; d < Q; ++d) {
        const int cd[3] = {kCx[d], kCy[d], kCz[d]};
        int s[3];
        int fas[3] = {-1, -1, -1};
        bool out = false;
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            s[a] = q0[a] - cd[a];
            if (s[a] < 0 || s[a] >= L[a]) {
                if ((g.per >> a) & 1) {
                    s[a] = s[a] < 0 ? s[a] + L[a] : s[a] - L[a];
                } else {
                    fas[a] = 2 * a + (s[a] >= L[a] ? 1 : 0);
                    out = true;
                }
            }
        }
        bool pull = ((codes[d / 3] >> (10 * (d % 3))) & 3u) == 0u;
        if (out) {
            // edge precedence: VELOCITY > PRESSURE_ABB > ZERO_GRADIENT, lower axis first
            int chosen = -1, kind = LBM_FACE_ZERO_GRADIENT;
            const int order[2] = {LBM_FACE_VELOCITY, LBM_FACE_PRESSURE_ABB};
            for (int k = 0; k < 2 && chosen < 0; ++k)
                for (int a = 0; a < 3 && chosen < 0; ++a)
                    if (fas[a] >= 0 && face_kind_at(g, fas[a]) == order[k]) {
                        chosen = fas[a];
                        kind = order[k];
                    }
            if (kind == LBM_FACE_ZERO_GRADIENT) {
                for (int a = 0; a < 3; ++a)
                    if (fas[a] >= 0) s[a] = q0[a];
            } else {
                codes[d / 3] |= (uint32_t)(DIR_FACE | (chosen << 2)) << (10 * (d % 3));
                pull = false;
            }
        }
        // a face or link direction loads the cell's own f*[opp(d)], which its
        // rule below starts from, so the rules need no further memory access
        const typename S::T* a = fin + (int64_t)d * g.P + (s[0] + g.nx * s[1] + (int)g.nxy * s[2]);
        const typename S::T* b = fin + (int64_t)kOpp[d] * g.P + cell;
        f[d] = S::load_pick(pull ? a : b, kOpp[d], d, pull, g.sh);
    }
    __builtin_amdgcn_sched_barrier(0);
    // 2) face and link values; wall point of the moving-wall velocity:
    //    physical x + c_dl/2, or the reference's x - c_dl/2 (src/SolidBC.jl:195-196)
    const double hs = g.wall_point == LBM_WALL_REFERENCE ? -0.5 : 0.5;
#pragma unroll
    for (int d = 0; d < Q; ++d) {
        const uint32_t code = (codes[d / 3] >> (10 * (d % 3))) & 0x3ffu;
        const int kind_d = (int)(code & 3u), aux = (int)(code >> 2);
        if (kind_d == DIR_FACE) {
            // face velocities from LDS: a global load here would wait once per face direction
            const C uf[3] = {C(lds_face_u[3 * aux]), C(lds_face_u[3 * aux + 1]), C(lds_face_u[3 * aux + 2])};
            f[d] = face_value<S>(face_kind_at(g, aux), uf, d, f[d], rho_prev, up, uu2);
        } else if (kind_d == DIR_LINK) {
            // boundary link (x, dl), dl = opp(d): the solid node is x + c_dl
            const int dl = kOpp[d];
            const DevBody& B = sv.bnew[aux];
            const int rule = IBB ? (int)((rules >> (2 * d)) & 3u) : 0;
            const double qw = IBB && rule != 0 ? (double)qv[d * kBlock] : 0.5;
            // wall point x + h c_dl of the wall velocity: h = q (IBB) or 1/2 (SBB),
            // mirrored to -h for the reference's sampling point
            const double hw = rule == 0 ? hs : (g.wall_point == LBM_WALL_REFERENCE ? -qw : qw);
            double r[3];
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                const int cdl = a == 0 ? kCx[dl] : (a == 1 ? kCy[dl] : kCz[dl]);
                r[a] = (double)q0[a] + 0.5 + hw * (double)cdl - B.ch[a];
                if ((g.per >> a) & 1) {
                    const double h = 0.5 * L[a];
                    if (r[a] > h) r[a] -= L[a];
                    else if (r[a] < -h) r[a] += L[a];
                }
            }
            const C uwx = C(B.v[0] + (B.om[1] * r[2] - B.om[2] * r[1]));
            const C uwy = C(B.v[1] + (B.om[2] * r[0] - B.om[0] * r[2]));
            const C uwz = C(B.v[2] + (B.om[0] * r[1] - B.om[1] * r[0]));
            const C cuw = C(kCx[dl]) * uwx + C(kCy[dl]) * uwy + C(kCz[dl]) * uwz;
            const C fin_dl = f[d];
            const C corr = C(2) * weight<C>(dl) * (S::kDev ? C(1) + rho_prev : rho_prev) * cuw * C(3);
            C val = fin_dl - corr;
            if (IBB && rule != 0) {
                // rule 1: f*_dl at the second fluid node x - c_dl; rule 2: the
                // cell's own f*_d, which leaves towards that node.  The offsets are
                // formed here (an opaque copy of the cell index), not hoisted ahead
                // of the loop for every direction.
                int xs = x, ys = y, zs = z;
                asm volatile("" : "+v"(xs), "+v"(ys), "+v"(zs));
                int s2[3] = {xs - kCx[dl], ys - kCy[dl], zs - kCz[dl]};
#pragma unroll
                for (int a = 0; a < 3; ++a) wrap_coord(s2[a], L[a], (g.per >> a) & 1);
                const int64_t cell2 = (int64_t)s2[0] + (int64_t)g.nx * s2[1] + g.nxy * s2[2];
                const int64_t own = (int64_t)xs + (int64_t)g.nx * ys + g.nxy * zs;
                const bool near = rule == 1;
                const int64_t off = near ? (int64_t)dl * g.P + cell2 : (int64_t)d * g.P + own;
                const C fx = S::load_pick(fin + off, dl, d, !near, g.sh);
                const C q2 = C(2) * C(qw);
                val = near ? q2 * fin_dl + (C(1) - q2) * fx - corr : (fin_dl + (q2 - C(1)) * fx - corr) / q2;
            }
            f[d] = val;
            if (MODE == MODE_STEP) {
                // momentum exchange of this link (src/MEA.jl:102-107), FP64, arm to the
                // wall point x + c_dl/2 (SBB) or x + q c_dl (IBB)
                const double t = S::kDev ? (double)fin_dl + (double)val + 2.0 * rest_level<S>(dl)
                                         : (double)(fin_dl + val);
                const double P[3] = {t * kCx[dl], t * kCy[dl], t * kCz[dl]};
                const double ha = rule == 0 ? 0.5 : qw;
                double ra[3];
#pragma unroll
                for (int a = 0; a < 3; ++a) {
                    const int cdl = a == 0 ? kCx[dl] : (a == 1 ? kCy[dl] : kCz[dl]);
                    ra[a] = (double)q0[a] + 0.5 + ha * (double)cdl - B.ch[a];
                    if ((g.per >> a) & 1) {
                        const double h = 0.5 * L[a];
                        if (ra[a] > h) ra[a] -= L[a];
                        else if (ra[a] < -h) ra[a] += L[a];
                    }
                }
                const double v6[6] = {P[0], P[1], P[2], ra[1] * P[2] - ra[2] * P[1], ra[2] * P[0] - ra[0] * P[2],
                                      ra[0] * P[1] - ra[1] * P[0]};
                if (acc.b0 < 0 || acc.b0 == aux) {
                    acc.b0 = aux;
                    for (int c = 0; c < 6; ++c) acc.a0[c] += v6[c];
                } else {
                    wave_slot_add(aux, v6, lds_wave);
                }
            }
        }
        // the interpolated rule's extra loads and FP64 sums: one direction at a time
        if (IBB) __builtin_amdgcn_sched_barrier(0);
    }
    finish_face_cell<S, COLL, FORCED, MODE>(f, p, cell, g, fout, gout, rho_out, u_out, x, y, z, -1);
}

// ---------------------------------------------------------------------------
// k_fixup: the general path.  Blocks [0, sg.nfix_blocks) walk the boxes of
// kind 2 (domain edges: every cell) and kind 1 (body regions); blocks
// [sg.nfix_blocks, + sg.list_blocks) walk the static boundary-node list.
//  * body region, moving bodies (mode 1): every cell (k_bulk skips the
//    region, so the two kernels write disjoint cells and run concurrently);
//    the thread also rewrites its cell's flag;
//  * body region, static bodies (mode 0): face cells only;
//  * list: boundary nodes not on a face.
// Cells in several regions go to the first.  Every block writes nbodies*6
// partial force/torque sums (fixed order) to partials[blockIdx].
// ---------------------------------------------------------------------------
template <class S, int COLL, bool FORCED, int MODE, bool IBB = false>
__device__ __forceinline__ void fixup_block(const int blk, const typename S::T* __restrict__ fin,
                                            typename S::T* __restrict__ fout, double* __restrict__ gout,
                                            double* __restrict__ rho_out, double* __restrict__ u_out, const Grid& g,
                                            const Phys<typename S::C>& p, const StepGeom* __restrict__ sgp,
                                            const DevBody* __restrict__ bnew, const DevBody* __restrict__ bold,
                                            uint8_t* __restrict__ flags, const int32_t* __restrict__ list,
                                            const int32_t* __restrict__ list_count, double* __restrict__ partials,
                                            int write_flags, double* lds) {
    // The fix-up waves share their CUs with bulk waves; their latency is the
    // tail of the step, so they take issue priority.
    __builtin_amdgcn_s_setprio(3);
    const StepGeom& sg = *sgp;
    const int nb = sg.nbodies;
    const int n6 = nb * 6;
    const int tid = threadIdx.x;
    // LDS: per wave nbodies x 6 partial sums
    double* s_wave = lds + (tid >> 6) * n6;
    for (int i = tid; i < kWaves * n6; i += kBlock) lds[i] = 0.0;
    // then the bodies (new poses, then old), when few enough
    const bool staged = nb <= kLdsBodies;
    DevBody* s_bodies = reinterpret_cast<DevBody*>(lds + kWaves * n6);
    // then the faces' wall / inlet velocities (6 x 3), then, with the interpolated
    // bounce-back, the wall fractions (18 per thread)
    double* s_face_u = reinterpret_cast<double*>(reinterpret_cast<char*>(s_bodies) +
                                                 (staged ? 2 * (size_t)nb * sizeof(DevBody) : 0));
    float* s_q = reinterpret_cast<float*>(s_face_u + 18);
    if (tid < 18) s_face_u[tid] = g.face_u[tid];
    if (staged) {
        const int nw = 2 * nb * (int)(sizeof(DevBody) / 4);
        const uint32_t* src_new = reinterpret_cast<const uint32_t*>(bnew);
        const uint32_t* src_old = reinterpret_cast<const uint32_t*>(bold);
        uint32_t* dst = reinterpret_cast<uint32_t*>(s_bodies);
        const int half = nw / 2;
        for (int i = tid; i < nw; i += kBlock) dst[i] = i < half ? src_new[i] : src_old[i - half];
    }
    __syncthreads();

    SolidView sv;
    sv.mode = sg.mode;
    sv.nb = nb;
    sv.flags = flags;
    sv.bnew = staged ? s_bodies : bnew;
    sv.bold = staged ? s_bodies + nb : bold;

    ForceAcc acc;
    acc.b0 = -1;
    for (int c = 0; c < 6; ++c) acc.a0[c] = 0.0;

    bool work = false, plain = false;
    int x = 0, y = 0, z = 0;
    // moving bodies: the cell's own body and its link codes (general_cell's format)
    int pre_self = 0;
    LinkCodes pre = {{0u, 0u, 0u, 0u, 0u, 0u, 0u}};
    if (blk < sg.nfix_blocks) {
        // the box of this block: the number of boxes starting at or before
        // it, from independent loads of the first kScanBoxes start blocks
        int bi;
        if (sg.nbox <= kScanBoxes) {
            int cnt = 0;
#pragma unroll
            for (int k = 0; k < kScanBoxes; ++k) {
                const int fb = sg.box[k].first_block;
                cnt += (k >= sg.first_fix_box && k < sg.nbox && fb <= blk) ? 1 : 0;
            }
            bi = sg.first_fix_box + cnt - 1;
        } else {
            bi = sg.first_fix_box;
            while (bi + 1 < sg.nbox && sg.box[bi + 1].first_block <= blk) ++bi;
        }
        const Box b = sg.box[bi];
        const int t = (blk - b.first_block) * kBlock + tid;  // box volumes stay far below 2^31
        const int vol = b.ext[0] * b.ext[1] * b.ext[2];
        if (t < vol) {
            const int i = t % b.ext[0];
            const int r = t / b.ext[0];
            const int j = r % b.ext[1];
            const int k = r / b.ext[1];
            x = b.lo[0] + i;
            y = b.lo[1] + j;
            z = b.lo[2] + k;
            if (x < 0) x += g.nx;
            if (x >= g.nx) x -= g.nx;
            if (y < 0) y += g.ny;
            if (y >= g.ny) y -= g.ny;
            if (z < 0) z += g.nz;
            if (z >= g.nz) z -= g.nz;
            bool dup = false;
            if (b.ovl) {
                if (bi <= kScanBoxes) {
#pragma unroll
                    for (int m = 0; m < kScanBoxes; ++m) {
                        const bool in = in_box_normalised(sg.box[m], x, y, z, g);
                        dup = dup || (m >= sg.first_fix_box && m < bi && in);
                    }
                } else {
                    for (int m = sg.first_fix_box; m < bi && !dup; ++m) dup = in_box(sg.box[m], x, y, z, g);
                }
            }
            if (!dup && sv.mode == 1) {
                // moving bodies: solidity of the cell and its neighbours, once
                pre_self = link_codes_moving(sv, x, y, z, g, pre);
                if (write_flags) {
                    // the region's flags follow the new pose (the static fix-up
                    // bits are rebuilt when the bodies stop)
                    const int64_t cell = (int64_t)x + (int64_t)g.nx * y + g.nxy * z;
                    flags[cell] = (uint8_t)pre_self;
                }
            }
            if (b.kind == 2) {
                work = !dup;  // domain edge: every cell
            } else if (!dup) {
                // moving bodies: the bulk pass skipped the whole region, so every
                // cell is done here (boundary, fresh and face cells on the general
                // path, the rest on the plain one); static bodies: only the
                // region's face cells (the list has the rest)
                const bool face = open_face_count(x, y, z, g) == 1;
                if (sv.mode == 1) {
                    work = true;
                    if (!face) {
                        // a fluid cell with a solid neighbour, or a fresh one, takes
                        // the general path
                        bool gen = false;
                        if (pre_self == 0) {
                            gen = g.refill && flag_of(sv.bold, nb, x, y, z, g) != 0;
#pragma unroll
                            for (int i = 0; i < 7; ++i) gen = gen || pre.w[i] != 0u;
                        }
                        plain = !gen;
                    }
                } else {
                    work = face;
                }
            }
        }
    } else {
        const int64_t t = (int64_t)(blk - sg.nfix_blocks) * kBlock + tid;
        if (t < (int64_t)(*list_count)) {
            const int64_t cell = list[t];
            int64_t r = cell / g.nx;
            x = (int)(cell - r * g.nx);
            z = (int)(r / g.ny);
            y = (int)(r - (int64_t)z * g.ny);
            work = open_face_count(x, y, z, g) == 0;  // face cells belong to the regions
        }
    }
    if (work && !plain && sv.mode == 0 && nb > 0) pre_self = link_codes_flags(sv, x, y, z, g, pre);
    if (work && (MODE == MODE_STEP || MODE == MODE_GET)) {
        if (plain)
            plain_cell<S, COLL, FORCED, MODE>(fin, fout, gout, rho_out, u_out, g, p, x, y, z);
        else
            general_cell<S, COLL, FORCED, MODE, IBB>(fin, fout, gout, rho_out, u_out, g, p, sv, x, y, z, acc, s_wave,
                                                     s_face_u, pre_self, pre, s_q);
    }
    if (MODE == MODE_STEP && nb > 0) {
        wave_reduce_to_lds(acc.b0, acc.a0, s_wave);
        __syncthreads();
        for (int i = tid; i < n6; i += kBlock) {
            double s = lds[i];
            for (int w = 1; w < kWaves; ++w) s += lds[w * n6 + i];
            partials[(int64_t)blk * n6 + i] = s;
        }
    }
}

// ---------------------------------------------------------------------------
// k_step: one launch per time step.  Blocks [0, nfix) are the fix-up blocks
// (edges, body regions, static boundary nodes) and are dispatched first; the
// remaining blocks are the bulk pass over rows of blockDim-sized x runs.  The
// two sets write disjoint cells, so they run side by side with no barrier.
// ---------------------------------------------------------------------------
struct BulkShape {
    int bx, by;   // cell groups per block along x and rows per block
    int gx, gy;   // blocks along x and y
};

template <class S>
inline BulkShape bulk_shape(const Grid& g) {
    const int groups = (g.nx + S::kVec - 1) / S::kVec;
    BulkShape bs;
    bs.bx = groups >= 256 ? 256 : (groups >= 128 ? 128 : 64);
    bs.by = kBlock / bs.bx;
    bs.gx = (groups + bs.bx - 1) / bs.bx;
    bs.gy = (g.ny + bs.by - 1) / bs.by;
    return bs;
}

// dynamic LDS of a step launch: the fix-up blocks' force partials, staged
// bodies, face velocities and (interpolated bounce-back) wall fractions
inline size_t step_lds_bytes(int nbodies, bool ibb = false) {
    size_t b = (size_t)kWaves * (size_t)nbodies * 6 * sizeof(double);
    if (nbodies <= kLdsBodies) b += 2 * (size_t)nbodies * sizeof(DevBody);
    b += 18 * sizeof(double);
    if (ibb) b += 18 * (size_t)kBlock * sizeof(float);
    return b;
}

// Minimum waves per SIMD the step kernel is compiled for.  The bulk rows need
// few registers; the rare boundary path needs many, and a higher floor makes
// it spill to scratch instead of lowering the occupancy of every bulk wave.
#ifndef LBM_STEP_MIN_WAVES
#define LBM_STEP_MIN_WAVES 1
#endif
template <class S, int COLL, bool FORCED, int MODE, bool IBB = false>
__global__ __launch_bounds__(kBlock, LBM_STEP_MIN_WAVES) void k_step(const typename S::T* __restrict__ fin,
                                                 typename S::T* __restrict__ fout, double* __restrict__ gout,
                                                 double* __restrict__ rho_out, double* __restrict__ u_out, Grid g,
                                                 Phys<typename S::C> p, const StepGeom* __restrict__ sgp,
                                                 const DevBody* __restrict__ bnew, const DevBody* __restrict__ bold,
                                                 uint8_t* __restrict__ flags, const int32_t* __restrict__ list,
                                                 const int32_t* __restrict__ list_count,
                                                 double* __restrict__ partials, int write_flags, BulkShape bs,
                                                 int nfix) {
    extern __shared__ double lds_dyn[];
    const int blk = blockIdx.x;
    if (blk < nfix) {
        fixup_block<S, COLL, FORCED, MODE, IBB>(blk, fin, fout, gout, rho_out, u_out, g, p, sgp, bnew, bold, flags, list,
                                           list_count, partials, write_flags, lds_dyn);
        return;
    }
    const int b = blk - nfix;
    const int bxi = b % bs.gx;
    const int r = b / bs.gx;
    const int byi = r % bs.gy;
    const int bz = r / bs.gy;
    const int tx = threadIdx.x % bs.bx, ty = threadIdx.x / bs.bx;
    bulk_group<S, COLL, FORCED, MODE>(fin, fout, gout, rho_out, u_out, g, p, sgp, flags,
                                      (bxi * bs.bx + tx) * S::kVec,
                                      byi * bs.by + ty, bz);
}

// ---------------------------------------------------------------------------
// Start state: f_old = f_eq(rho, u) in storage format (initialize_f_gpu,
// src/Initialization.jl:27-44, without the random perturbation), or f_old from
// FP64 host data.
// ---------------------------------------------------------------------------
template <class S>
__global__ __launch_bounds__(kBlock) void k_equilibrium(typename S::T* __restrict__ f, const double* __restrict__ rho,
                                                        const double* __restrict__ u, double rho0, double u0x,
                                                        double u0y, double u0z, int64_t N, int64_t P, Shift sh) {
    using C = typename S::C;
    const int64_t cell = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (cell >= N) return;
    const C r = rho ? C(rho[cell]) : C(rho0);
    const C ux = u ? C(u[cell]) : C(u0x);
    const C uy = u ? C(u[N + cell]) : C(u0y);
    const C uz = u ? C(u[2 * N + cell]) : C(u0z);
    const C u2 = ux * ux + uy * uy + uz * uz;
#pragma unroll
    for (int d = 0; d < Q; ++d) {
        const C cu = C(kCx[d]) * ux + C(kCy[d]) * uy + C(kCz[d]) * uz;
        const C feq = weight<C>(d) * r * (C(1) + cu * C(3) + C(0.5) * cu * cu * C(9) - C(0.5) * u2 * C(3));
        S::store(f + (int64_t)d * P + cell, d, feq, sh);
    }
}

template <class S>
__global__ __launch_bounds__(kBlock) void k_from_double(typename S::T* __restrict__ f, const double* __restrict__ src,
                                                        int64_t N, int64_t P, Shift sh) {
    const int64_t cell = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (cell >= N) return;
#pragma unroll
    for (int d = 0; d < Q; ++d) S::store(f + (int64_t)d * P + cell, d, (typename S::C)src[(int64_t)d * N + cell], sh);
}

}  // namespace lbm
