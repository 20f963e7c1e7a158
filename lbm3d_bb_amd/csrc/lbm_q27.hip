// lbm_q27.hip -- the D3Q27 lattice (src/Utilities.jl:962-976): the same time
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
// (the address is selected, the load is unconditional).  FP64 and FP32.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "lbm_device.h"
#include "lbm_q27.h"

namespace lbm {
namespace q27 {

constexpr int QQ = kQ;
// ex, ey, ez of src/Utilities.jl:962-964: six faces, twelve edges, eight
// corners, rest last; moving directions pair as (2k, 2k + 1)
constexpr int kX[QQ] = {1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 0, 0, 1, -1, 1, -1, 0, 0, 1, -1, 1, -1, 1, -1, -1, 1, 0};
constexpr int kY[QQ] = {0, 0, 1, -1, 0, 0, 1, -1, 0, 0, 1, -1, -1, 1, 0, 0, 1, -1, 1, -1, 1, -1, -1, 1, 1, -1, 0};
constexpr int kZ[QQ] = {0, 0, 0, 0, 1, -1, 0, 0, 1, -1, 1, -1, 0, 0, -1, 1, -1, 1, 1, -1, -1, 1, 1, -1, 1, -1, 0};

// bbOpp (src/Utilities.jl:976)
__host__ __device__ constexpr int opp(int d) { return d == QQ - 1 ? d : (d ^ 1); }

// w (src/Utilities.jl:967-973): 8/27 rest, 2/27 face, 1/54 edge, 1/216 corner
template <typename C>
__device__ inline C w27(int d) {
    const int n = (kX[d] != 0) + (kY[d] != 0) + (kZ[d] != 0);
    return n == 0 ? C(8) / C(27) : (n == 1 ? C(2) / C(27) : (n == 2 ? C(1) / C(54) : C(1) / C(216)));
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
// collision keeps rho and adds F to the momentum, so u = j*/rho - g/2.
template <typename C, bool FORCED>
__device__ inline void prev_moments(C rho, C jx, C jy, C jz, const Phys<C>& p, C& rho_out, C u[3], C& uu2) {
    rho_out = rho;
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

// Whether the rows (y, z) of this block's cells come within reach of a body,
// from block-uniform values: the y and z intervals of the rows, against each
// body's extent (min image on periodic axes; an interval of half the axis or
// more counts as near).
__device__ inline bool interval_near(int lo, int hi, float c, float ext, bool per, int L) {
    const float half = 0.5f * (float)(hi - lo + 1);
    if (per && 2 * (hi - lo + 1) >= L) return true;
    const float m = 0.5f * (float)(lo + hi + 1);
    return fabsf(min_image(m - c, per, L)) <= ext + half;
}

__device__ inline bool block_near(uint32_t block, const DevBody* __restrict__ bnew, const DevBody* __restrict__ bold,
                                  int nb, const Grid& g) {
    if (nb <= 0) return false;
    // 32-bit arithmetic: cell counts stay below 2^31
    const uint32_t c0 = block * (uint32_t)kBlock;
    const uint32_t c1 = (c0 + kBlock < (uint32_t)g.N ? c0 + kBlock : (uint32_t)g.N) - 1u;
    const uint32_t nx = (uint32_t)g.nx, ny = (uint32_t)g.ny;
    const uint32_t r0 = c0 / nx, r1 = c1 / nx;  // rows y + ny z
    const int z0 = (int)(r0 / ny), z1 = (int)(r1 / ny);
    const int y0 = z0 == z1 ? (int)(r0 - (uint32_t)z0 * ny) : 0;
    const int y1 = z0 == z1 ? (int)(r1 - (uint32_t)z1 * ny) : g.ny - 1;
    bool hit = false;
    for (int b = 0; b < nb && !hit; ++b) {
        for (int k = 0; k < 2 && !hit; ++k) {
            if (k == 1 && bold == nullptr) break;
            const DevBody& B = k == 0 ? bnew[b] : bold[b];
            hit = B.reach < 0.0f || (interval_near(y0, y1, B.c[1], B.ext[1], g.per & 2, g.ny) &&
                                     interval_near(z0, z1, B.c[2], B.ext[2], g.per & 4, g.nz));
        }
    }
    return hit;
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
template <typename C>
__device__ inline C velocity_rule(int d, C own_opp, const C uf[3], C rho_prev) {
    const C cu = C(kX[d]) * (-uf[0]) + C(kY[d]) * (-uf[1]) + C(kZ[d]) * (-uf[2]);
    return own_opp - C(2) * w27<C>(d) * rho_prev * cu * C(3);
}

template <typename C>
__device__ inline C abb_rule(int d, C own_opp, const C up[3], C uu2, C rho_prev) {
    const C cu = C(kX[d]) * up[0] + C(kY[d]) * up[1] + C(kZ[d]) * up[2];
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
template <typename C, int FC>
__device__ inline void one_face_rule(C f[QQ], bool velocity, const C uf[3], const C up[3], C uu2, C rho_prev) {
#pragma unroll
    for (int d = 0; d < QQ; ++d) {
        if (!((face_dirs(FC) >> d) & 1u)) continue;
        f[d] = velocity ? velocity_rule<C>(d, f[d], uf, rho_prev) : abb_rule<C>(d, f[d], up, uu2, rho_prev);
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

template <class S, int COLL, bool FORCED, int MODE>
__global__ __launch_bounds__(kBlock) void k27_step(const typename S::T* __restrict__ fin, typename S::T* __restrict__ fout,
                                                   double* __restrict__ gout, double* __restrict__ rho_out,
                                                   double* __restrict__ u_out, Grid g, Phys<typename S::C> p,
                                                   const uint8_t* __restrict__ fl_new, const uint8_t* __restrict__ fl_old,
                                                   const DevBody* __restrict__ bnew, const DevBody* __restrict__ bold,
                                                   int nb, double* __restrict__ partials, int refill,
                                                   const double* __restrict__ face_u) {
    using C = typename S::C;
    extern __shared__ double lds_dyn[];
    const int n6 = nb * 6;
    double* s_over = lds_dyn + kWaves * n6;  // links of a second body in one thread
    // Only blocks within reach of a body can hold links: they alone reduce
    // forces and write their partials (the sum visits the same blocks).
    const bool bnear = MODE != MODE_COLLIDE_ONLY && MODE != MODE_GET_RAW && block_near(blockIdx.x, bnew, bold, nb, g);
    if (MODE == MODE_STEP && bnear) {
        for (int i = threadIdx.x; i < (kWaves + 1) * n6; i += kBlock) lds_dyn[i] = 0.0;
        __syncthreads();
    }
    ForceAcc acc;
    acc.b0 = -1;
#pragma unroll
    for (int c = 0; c < 6; ++c) acc.a0[c] = 0.0;
    const int64_t N = g.N;
    const int64_t cell = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (cell < N) {
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
            const bool fresh = near && refill && self == 0 && (fl_old[cell] & kFlagBody) != 0;
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
            if (fresh) {
                // equilibrium of the mean previous-step rho, u of the neighbours
                // x + c_d that are fluid and not fresh (refill, extension)
                // (a rare path: one neighbour at a time keeps the registers of
                // the step's other paths low)
                C sr = C(0), su = C(0), sv = C(0), sw = C(0);
                int cnt = 0;
                for (int d = 0; d < QQ - 1; ++d) {
                    const int64_t nc = in_domain_source(q0, opp(d), g);
                    if (nc < 0 || (fl_new[nc] & kFlagBody) != 0 || (fl_old[nc] & kFlagBody) != 0) continue;
                    C r = C(0), jx = C(0), jy = C(0), jz = C(0);
#pragma unroll
                    for (int e = 0; e < QQ; ++e) {
                        const C v = S::load(fin + (int64_t)e * g.P + nc, e, g.sh);
                        r += v;
                        if (kX[e] > 0) jx += v;
                        if (kX[e] < 0) jx -= v;
                        if (kY[e] > 0) jy += v;
                        if (kY[e] < 0) jy -= v;
                        if (kZ[e] > 0) jz += v;
                        if (kZ[e] < 0) jz -= v;
                    }
                    C rr, uu[3], u2;
                    prev_moments<C, FORCED>(r, jx, jy, jz, p, rr, uu, u2);
                    sr += rr;
                    su += uu[0];
                    sv += uu[1];
                    sw += uu[2];
                    ++cnt;
                }
                C r = C(g.rho0), ux = C(0), uy = C(0), uz = C(0);
                if (cnt > 0) {
                    const C n = C(cnt);
                    r = sr / n;
                    ux = su / n;
                    uy = sv / n;
                    uz = sw / n;
                }
#pragma unroll
                for (int d = 0; d < QQ; ++d) f[d] = feq27<C>(d, r, ux, uy, uz);
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
                        prev_moments<C, FORCED>(r, jx, jy, jz, p, rho_prev, up, uu2);
                    } else {
#pragma unroll
                        for (int d = 0; d < QQ; ++d) r += S::load(fin + (int64_t)d * g.P + cell, d, g.sh);
                        rho_prev = r;
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
                        case 0: one_face_rule<C, 0>(f, vel, uf, up, uu2, rho_prev); break;
                        case 1: one_face_rule<C, 1>(f, vel, uf, up, uu2, rho_prev); break;
                        case 2: one_face_rule<C, 2>(f, vel, uf, up, uu2, rho_prev); break;
                        case 3: one_face_rule<C, 3>(f, vel, uf, up, uu2, rho_prev); break;
                        case 4: one_face_rule<C, 4>(f, vel, uf, up, uu2, rho_prev); break;
                        default: one_face_rule<C, 5>(f, vel, uf, up, uu2, rho_prev); break;
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
                            if (vel) f[d] = velocity_rule<C>(d, own_opp, uf, rho_prev);
                            else f[d] = abb_rule<C>(d, own_opp, up, uu2, rho_prev);
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
                            double r[3];
#pragma unroll
                            for (int a = 0; a < 3; ++a)
                                r[a] = wrap_arm((double)q0[a] + 0.5 + hs * (double)cdl[a] - ch[a], a, g);
                            const C uwx = C(v[0] + (om[1] * r[2] - om[2] * r[1]));
                            const C uwy = C(v[1] + (om[2] * r[0] - om[0] * r[2]));
                            const C uwz = C(v[2] + (om[0] * r[1] - om[1] * r[0]));
                            const C cuw = C(cdl[0]) * uwx + C(cdl[1]) * uwy + C(cdl[2]) * uwz;
                            const C val = own_opp - C(2) * w * rho_prev * cuw * C(3);
                            f[d] = val;
                            if (MODE == MODE_STEP) {
                                // momentum exchange (src/MEA.jl:102-107) in FP64, torque arm to x + c_dl/2
                                const double t = (double)(own_opp + val);
                                const double P[3] = {t * cdl[0], t * cdl[1], t * cdl[2]};
                                double ra[3];
#pragma unroll
                                for (int a = 0; a < 3; ++a)
                                    ra[a] = wrap_arm((double)q0[a] + 0.5 + 0.5 * (double)cdl[a] - ch[a], a, g);
                                const double v6[6] = {P[0], P[1], P[2], ra[1] * P[2] - ra[2] * P[1],
                                                      ra[2] * P[0] - ra[0] * P[2], ra[0] * P[1] - ra[1] * P[0]};
                                if (acc.b0 < 0 || acc.b0 == bi) {
                                    acc.b0 = bi;
#pragma unroll
                                    for (int c = 0; c < 6; ++c) acc.a0[c] += v6[c];
                                } else {
#pragma unroll
                                    for (int c = 0; c < 6; ++c) atomicAdd(&s_over[bi * 6 + c], v6[c]);
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
            C ux = jx / rho, uy = jy / rho, uz = jz / rho;
            if (FORCED) {
                ux += C(0.5) * p.g[0];
                uy += C(0.5) * p.g[1];
                uz += C(0.5) * p.g[2];
            }
            if (gout) {
#pragma unroll
                for (int d = 0; d < QQ; ++d) gout[(int64_t)d * N + cell] = (double)f[d];
            }
            if (rho_out) rho_out[cell] = (double)rho;
            if (u_out) {
                u_out[cell] = (double)ux;
                u_out[N + cell] = (double)uy;
                u_out[2 * N + cell] = (double)uz;
            }
        } else {
            collide27<C, COLL, FORCED>(f, p);
#pragma unroll
            for (int d = 0; d < QQ; ++d) S::store(fout + (int64_t)d * g.P + cell, d, f[d], g.sh);
        }
    }
    if (MODE == MODE_STEP && bnear) {
        const int wave = threadIdx.x >> 6;
        wave_reduce_to_lds(acc.b0, acc.a0, lds_dyn + wave * n6);
        __syncthreads();
        for (int i = threadIdx.x; i < n6; i += kBlock) {
            double s = s_over[i];
            for (int w = 0; w < kWaves; ++w) s += lds_dyn[w * n6 + i];
            partials[(int64_t)blockIdx.x * n6 + i] = s;
        }
    }
}

// Fixed-order sum of the block partials of step blockIdx.y over the blocks
// within reach of a body (the step record holds the step's bodies, new pose
// then old pose): one block per (body, component); thread t takes blocks t,
// t + kBlock, ..., then an LDS tree.
__global__ __launch_bounds__(kBlock) void k27_force_sum(const double* __restrict__ partials, int64_t nblocks,
                                                         int64_t stride, int n6, const char* __restrict__ recs,
                                                         size_t rec_bytes, Grid g, double* __restrict__ out) {
    __shared__ double s[kBlock];
    const int k = blockIdx.x;
    const int nb = n6 / 6;
    const double* ps = partials + (int64_t)blockIdx.y * stride;
    const DevBody* bn = reinterpret_cast<const DevBody*>(recs + (size_t)blockIdx.y * rec_bytes);
    double a = 0.0;
    for (int64_t b = threadIdx.x; b < nblocks; b += kBlock)
        if (block_near((uint32_t)b, bn, bn + nb, nb, g)) a += ps[b * n6 + k];
    s[threadIdx.x] = a;
    __syncthreads();
    for (int w = kBlock / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[(int64_t)blockIdx.y * n6 + k] = s[0];
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

// Solid flags (flag_of, the mask of k_flags_full) of the cells of one box per
// body (blockIdx.y): where a moved body's flags can change.
__global__ __launch_bounds__(kBlock) void k27_flags_boxes(uint8_t* __restrict__ flags, const DevBody* __restrict__ bodies,
                                                          int nb, Grid g, const int32_t* __restrict__ boxes) {
    const int32_t* bx = boxes + 6 * blockIdx.y;
    const uint32_t ex = (uint32_t)bx[3], ey = (uint32_t)bx[4], ez = (uint32_t)bx[5];
    const uint32_t vol = ex * ey * ez;
    for (uint32_t i = blockIdx.x * kBlock + threadIdx.x; i < vol; i += gridDim.x * kBlock) {
        const uint32_t t = i / ex, k = t / ey;
        int x = bx[0] + (int)(i - t * ex), y = bx[1] + (int)(t - k * ey), z = bx[2] + (int)k;
        x = x >= g.nx ? x - g.nx : x;  // lo lies in [0, L) and ext <= L
        y = y >= g.ny ? y - g.ny : y;
        z = z >= g.nz ? z - g.nz : z;
        flags[(int64_t)x + (int64_t)g.nx * y + g.nxy * z] = (uint8_t)flag_of(bodies, nb, x, y, z, g);
    }
}

// Links (fluid cell x, direction d with a solid node x + c_d), ascending cell
// then direction (kernel_fill_bb_csr!, src/Geometry.jl:249-257).
__device__ inline int links_of(const uint8_t* __restrict__ flags, int64_t cell, const Grid& g, uint32_t* mask) {
    *mask = 0u;
    if (cell >= g.N || (flags[cell] & kFlagBody) != 0) return 0;
    int x, y, z;
    xyz_of(cell, g, x, y, z);
    const int q0[3] = {x, y, z};
    for (int d = 0; d < QQ - 1; ++d) {
        const int64_t nc = in_domain_source(q0, opp(d), g);  // x + c_d
        if (nc >= 0 && (flags[nc] & kFlagBody) != 0) *mask |= 1u << d;
    }
    return __popc(*mask);
}

__global__ __launch_bounds__(kBlock) void k27_link_count(const uint8_t* __restrict__ flags, Grid g,
                                                         int32_t* __restrict__ counts) {
    __shared__ int s_l[kWaves];
    const int64_t cell = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    uint32_t m;
    int n = links_of(flags, cell, g, &m);
    for (int off = 32; off > 0; off >>= 1) n += __shfl_xor(n, off, 64);
    if ((threadIdx.x & 63) == 0) s_l[threadIdx.x >> 6] = n;
    __syncthreads();
    if (threadIdx.x == 0) {
        int t = 0;
        for (int w = 0; w < kWaves; ++w) t += s_l[w];
        counts[blockIdx.x] = t;
    }
}

__global__ __launch_bounds__(kBlock) void k27_link_fill(const uint8_t* __restrict__ flags, Grid g,
                                                        const int32_t* __restrict__ block_off,
                                                        int32_t* __restrict__ cells, int8_t* __restrict__ dirs) {
    __shared__ int s_l[kWaves];
    const int64_t cell = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    uint32_t m;
    const int n = links_of(flags, cell, g, &m);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    int incl = n;
    for (int off = 1; off < 64; off <<= 1) {
        const int t = __shfl_up(incl, off, 64);
        if (lane >= off) incl += t;
    }
    if (lane == 63) s_l[wave] = incl;
    __syncthreads();
    int base = block_off[blockIdx.x];
    for (int w = 0; w < wave; ++w) base += s_l[w];
    int pos = base + incl - n;
    for (int d = 0; d < QQ - 1; ++d)
        if ((m >> d) & 1u) {
            cells[pos] = (int32_t)cell;
            dirs[pos] = (int8_t)d;
            ++pos;
        }
}

template <class S, int MODE>
static void launch_mode(const StepArgs& a) {
    using T = typename S::T;
    using C = typename S::C;
    Phys<C> p;
    p.wp = (C)a.wp;
    p.wm = (C)a.wm;
    for (int k = 0; k < 3; ++k) p.g[k] = (C)a.grav[k];
    const dim3 grid((unsigned)nblocks_of(a.g.N)), block(kBlock);
    const size_t lds = MODE == MODE_STEP ? (size_t)(kWaves + 1) * a.nb * 6 * sizeof(double) : 0;
#define K27(COLL, FORCED)                                                                                           \
    hipLaunchKernelGGL((k27_step<S, COLL, FORCED, MODE>), grid, block, lds, a.stream, (const T*)a.fin, (T*)a.fout, \
                       a.gout, a.rho_out, a.u_out, a.g, p, a.fl_new, a.fl_old, a.bnew, a.bold, a.nb, a.partials,   \
                       a.refill, a.face_u)
    if (a.trt) {
        if (a.forced) K27(LBM_TRT, true);
        else K27(LBM_TRT, false);
    } else {
        if (a.forced) K27(LBM_BGK, true);
        else K27(LBM_BGK, false);
    }
#undef K27
}

template <class S>
static void launch_storage(const StepArgs& a) {
    switch (a.mode) {
        case MODE_STEP: launch_mode<S, MODE_STEP>(a); break;
        case MODE_GET: launch_mode<S, MODE_GET>(a); break;
        case MODE_COLLIDE_ONLY: launch_mode<S, MODE_COLLIDE_ONLY>(a); break;
        default: launch_mode<S, MODE_GET_RAW>(a); break;
    }
}

static int last_error() {
    const hipError_t e = hipGetLastError();
    return e == hipSuccess ? 0 : (int)e;
}

}  // namespace q27

int q27_launch_step(const q27::StepArgs& a) {
    if (a.storage == LBM_FP64) q27::launch_storage<StoreF64>(a);
    else q27::launch_storage<StoreF32>(a);
    return q27::last_error();
}

int q27_launch_force_sum(const double* partials, int64_t nblocks, int64_t stride, int n6, int nsteps,
                         const char* recs, size_t rec_bytes, const Grid& g, double* out, hipStream_t st) {
    hipLaunchKernelGGL(q27::k27_force_sum, dim3((unsigned)n6, (unsigned)nsteps), dim3(kBlock), 0, st, partials,
                       nblocks, stride, n6, recs, rec_bytes, g, out);
    return q27::last_error();
}

int q27_launch_equilibrium(int storage, const Grid& g, hipStream_t st, void* f, const double* rho, const double* u,
                           double rho0, const double u0[3]) {
    const dim3 grid((unsigned)nblocks_of(g.N)), block(kBlock);
    if (storage == LBM_FP64)
        hipLaunchKernelGGL(q27::k27_equilibrium<StoreF64>, grid, block, 0, st, (double*)f, rho, u, rho0, u0[0], u0[1],
                           u0[2], g.N, g.P, g.sh);
    else
        hipLaunchKernelGGL(q27::k27_equilibrium<StoreF32>, grid, block, 0, st, (float*)f, rho, u, rho0, u0[0], u0[1],
                           u0[2], g.N, g.P, g.sh);
    return q27::last_error();
}

int q27_launch_from_double(int storage, const Grid& g, hipStream_t st, void* f, const double* src) {
    const dim3 grid((unsigned)nblocks_of(g.N)), block(kBlock);
    if (storage == LBM_FP64)
        hipLaunchKernelGGL(q27::k27_from_double<StoreF64>, grid, block, 0, st, (double*)f, src, g.N, g.P, g.sh);
    else hipLaunchKernelGGL(q27::k27_from_double<StoreF32>, grid, block, 0, st, (float*)f, src, g.N, g.P, g.sh);
    return q27::last_error();
}

int q27_launch_flags_boxes(uint8_t* flags, const DevBody* bodies, int nb, const Grid& g, const int32_t* boxes,
                           int max_volume, hipStream_t st) {
    if (nb <= 0 || max_volume <= 0) return 0;
    const int64_t blocks = std::min<int64_t>(nblocks_of(max_volume), 2048);
    hipLaunchKernelGGL(q27::k27_flags_boxes, dim3((unsigned)blocks, (unsigned)nb), dim3(kBlock), 0, st, flags, bodies,
                       nb, g, boxes);
    return q27::last_error();
}

int q27_launch_link_count(const uint8_t* flags, const Grid& g, int32_t* counts, hipStream_t st) {
    hipLaunchKernelGGL(q27::k27_link_count, dim3((unsigned)nblocks_of(g.N)), dim3(kBlock), 0, st, flags, g, counts);
    return q27::last_error();
}

int q27_launch_link_fill(const uint8_t* flags, const Grid& g, const int32_t* block_off, int32_t* cells, int8_t* dirs,
                         hipStream_t st) {
    hipLaunchKernelGGL(q27::k27_link_fill, dim3((unsigned)nblocks_of(g.N)), dim3(kBlock), 0, st, flags, g, block_off,
                       cells, dirs);
    return q27::last_error();
}

}  // namespace lbm
