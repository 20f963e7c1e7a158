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

// generate_mask_s_local!_gpu's test (src/Geometry.jl:66-73) for a sphere, and
// its body-frame analogue for an ellipsoid.
__device__ inline bool inside_body(const DevBody& b, float px, float py, float pz, const Grid& g) {
#pragma clang fp contract(off)
    float dx = min_image(px - b.c[0], g.per & 1, g.nx);
    float dy = min_image(py - b.c[1], g.per & 2, g.ny);
    float dz = min_image(pz - b.c[2], g.per & 4, g.nz);
    if (b.shape == LBM_SPHERE) {
        float s = dx * dx + dy * dy;
        s = s + dz * dz;
        return s <= b.r2 + 1e-6f * b.r2;
    }
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
                                   double* __restrict__ rho_out, double* __restrict__ u_out) {
    using C = typename S::C;
    if (MODE == MODE_STEP || MODE == MODE_COLLIDE_ONLY) {
        C rho, u[3];
        collide<C, COLL, FORCED>(f, p, rho, u);
#pragma unroll
        for (int d = 0; d < Q; ++d) S::store(fout + (int64_t)d * P + cell, d, f[d]);
    } else {
        if (gout) {
#pragma unroll
            for (int d = 0; d < Q; ++d) gout[(int64_t)d * N + cell] = (double)f[d];
        }
        if (rho_out) {
            C rho, u[3];
            macros_of<C, FORCED>(f, p, rho, u);
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

// Body regions (box kind 1) cover every cell that can be a boundary node, a
// fresh node or a face cell next to a body, in both geometry modes.
__device__ inline bool in_body_box(const StepGeom& sg, int x, int y, int z, const Grid& g) {
    for (int m = 0; m < sg.nbox; ++m)
        if (sg.box[m].kind == 1 && in_box(sg.box[m], x, y, z, g)) return true;
    return false;
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

__device__ inline double face_u_at(const Grid& g, int i, int c) {
    double v = g.face_u[0][c];
    v = i == 1 ? g.face_u[1][c] : v;
    v = i == 2 ? g.face_u[2][c] : v;
    v = i == 3 ? g.face_u[3][c] : v;
    v = i == 4 ? g.face_u[4][c] : v;
    v = i == 5 ? g.face_u[5][c] : v;
    return v;
}

// Face rules for direction d whose source lies outside face fidx:
// VELOCITY src/OtherBC.jl:143-148, PRESSURE_ABB src/OtherBC.jl:156-165.
template <typename C>
__device__ inline C face_value(int kind, int fidx, int d, const Grid& g, C own_opp, C rho_prev, const C up[3],
                               C uu2) {
    const C w = weight<C>(d);
    if (kind == LBM_FACE_VELOCITY) {
        const C cu = C(kCx[d]) * (-C(face_u_at(g, fidx, 0))) + C(kCy[d]) * (-C(face_u_at(g, fidx, 1))) +
                     C(kCz[d]) * (-C(face_u_at(g, fidx, 2)));
        return own_opp - C(2) * w * rho_prev * cu * C(3);
    }
    const C cu = C(kCx[d]) * up[0] + C(kCy[d]) * up[1] + C(kCz[d]) * up[2];
    return -own_opp + C(2) * w * rho_prev * (C(1) + C(0.5) * cu * cu * C(9) - C(0.5) * uu2 * C(3));
}

template <class S, bool FORCED>
__device__ inline void own_moments(const typename S::C own[Q], const Phys<typename S::C>& p, typename S::C& rho,
                                   typename S::C up[3], typename S::C& uu2) {
    using C = typename S::C;
    C jx, jy, jz;
    moments<C>(own, rho, jx, jy, jz);
    up[0] = jx / rho;
    up[1] = jy / rho;
    up[2] = jz / rho;
    if (FORCED) {
        up[0] -= C(0.5) * p.g[0];
        up[1] -= C(0.5) * p.g[1];
        up[2] -= C(0.5) * p.g[2];
    }
    uu2 = up[0] * up[0] + up[1] * up[1] + up[2] * up[2];
}

// ---------------------------------------------------------------------------
// Face cell on exactly one non-periodic y or z face (AXIS, SIDE): directions
// whose source lies outside the face take the face rule (or zero gradient);
// all others are plain pulls with periodic wrap on the remaining axes.
// ---------------------------------------------------------------------------
template <class S, int COLL, bool FORCED, int MODE, int AXIS, int SIDE>
__device__ __forceinline__ void face_cell(const typename S::T* __restrict__ fin, typename S::T* __restrict__ fout,
                          double* __restrict__ gout, double* __restrict__ rho_out, double* __restrict__ u_out,
                          const Grid& g, const Phys<typename S::C>& p, int x, int y, int z) {
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
    C own[Q], f[Q];
#pragma unroll
    for (int d = 0; d < Q; ++d) own[d] = S::load(fin + (int64_t)d * g.P + cell, d);
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
        // zero gradient: the source clamped back onto the face plane
        if (out) src = kind == LBM_FACE_ZERO_GRADIENT ? src + ca * axis_stride : cell;
        f[d] = S::load(fin + (int64_t)d * g.P + src, d);
    }
    if (kind != LBM_FACE_ZERO_GRADIENT) {
        C rho_prev, up[3], uu2;
        own_moments<S, FORCED>(own, p, rho_prev, up, uu2);
#pragma unroll
        for (int d = 0; d < Q; ++d) {
            const int ca = AXIS == 1 ? kCy[d] : kCz[d];
            const bool out = SIDE == 0 ? (ca > 0) : (ca < 0);
            if (out) f[d] = face_value<C>(kind, fidx, d, g, own[kOpp[d]], rho_prev, up, uu2);
        }
    }
    finish_cell<S, COLL, FORCED, MODE>(f, p, cell, N, g.P, fout, gout, rho_out, u_out);
}

// ---------------------------------------------------------------------------
// k_bulk: a block is a run of blockDim.x consecutive x cells (up to 256, a
// whole row when Nx <= 256) in blockDim.y rows, so every wave streams 64
// consecutive cells of one row in each direction plane, the +-1 shifted loads
// of the x neighbours stay inside the block's own cache lines, and no thread
// divides.  It finishes every cell except
//   * domain edges (cells on two or more non-periodic faces), and
//   * face cells inside a body region,
// which k_fixup does.  A row on a y or z face takes the face path as a whole
// wave; a cell on an x face takes the face rule for its five outgoing
// directions inside the row's own coalesced loads.
// ---------------------------------------------------------------------------
template <class S, int COLL, bool FORCED, int MODE>
__device__ __forceinline__ void bulk_cell(const typename S::T* __restrict__ fin, typename S::T* __restrict__ fout,
                                          double* __restrict__ gout, double* __restrict__ rho_out,
                                          double* __restrict__ u_out, const Grid& g, const Phys<typename S::C>& p,
                                          const StepGeom* __restrict__ sgp, const uint8_t* __restrict__ flags, int x,
                                          int y, int z) {
    using C = typename S::C;
    if (x >= g.nx || y >= g.ny) return;
    const int nx = g.nx, nxy = (int)g.nxy;
    const int64_t N = g.N;
    const int cell = x + nx * y + nxy * z;
    C f[Q];
    if (MODE == MODE_COLLIDE_ONLY || MODE == MODE_GET_RAW) {
#pragma unroll
        for (int d = 0; d < Q; ++d) f[d] = S::load(fin + (int64_t)d * g.P + cell, d);
        finish_cell<S, COLL, FORCED, MODE>(f, p, cell, N, g.P, fout, gout, rho_out, u_out);
        return;
    }
    const bool px = g.per & 1;
    const bool xlo = !px && x == 0, xhi = !px && x == nx - 1;
    const bool ylo = !(g.per & 2) && y == 0, yhi = !(g.per & 2) && y == g.ny - 1;
    const bool zlo = !(g.per & 4) && z == 0, zhi = !(g.per & 4) && z == g.nz - 1;
    const int nyz = (ylo || yhi) + (zlo || zhi);
    const bool xf = xlo || xhi;
    if (nyz + xf >= 2) return;  // domain edge
    if (sgp && sgp->mode == 1) {
        // moving bodies: the fix-up blocks own their whole regions
        bool row = false;
        for (int m = sgp->first_fix_box; m < sgp->nbox && !row; ++m) {
            const Box& b = sgp->box[m];
            if (b.kind != 1) continue;
            int ry = y - b.lo[1], rz = z - b.lo[2];
            if (g.per & 2) ry = ((ry % g.ny) + g.ny) % g.ny;
            if (g.per & 4) rz = ((rz % g.nz) + g.nz) % g.nz;
            row = ry >= 0 && ry < b.ext[1] && rz >= 0 && rz < b.ext[2];
        }
        if (row && in_body_box(*sgp, x, y, z, g)) return;
    } else if (sgp && sgp->nbodies > 0) {
        // static bodies: boundary nodes carry the fix-up bit, face cells of a
        // body region belong to the fix-up too
        if (flags[cell] & kFlagFixup) return;
        if ((nyz + xf) == 1 && in_body_box(*sgp, x, y, z, g)) return;
    }
    if (nyz == 1) {
        if (ylo) face_cell<S, COLL, FORCED, MODE, 1, 0>(fin, fout, gout, rho_out, u_out, g, p, x, y, z);
        else if (yhi) face_cell<S, COLL, FORCED, MODE, 1, 1>(fin, fout, gout, rho_out, u_out, g, p, x, y, z);
        else if (zlo) face_cell<S, COLL, FORCED, MODE, 2, 0>(fin, fout, gout, rho_out, u_out, g, p, x, y, z);
        else face_cell<S, COLL, FORCED, MODE, 2, 1>(fin, fout, gout, rho_out, u_out, g, p, x, y, z);
        return;
    }
    // periodic wraps (zero on non-periodic axes)
    const int wxm = (px && x == 0) ? nx : 0, wxp = (px && x == nx - 1) ? -nx : 0;
    const int wym = (y == 0) ? nxy : 0, wyp = (y == g.ny - 1) ? -nxy : 0;
    const int wzm = (z == 0) ? (int)N : 0, wzp = (z == g.nz - 1) ? -(int)N : 0;
    const int xkind = xlo ? g.face_kind[0] : (xhi ? g.face_kind[1] : -1);
#pragma unroll
    for (int d = 0; d < Q; ++d) {
        int src = cell - (kCx[d] + kCy[d] * nx + kCz[d] * nxy);
        if (kCx[d] > 0) src += wxm;
        if (kCx[d] < 0) src += wxp;
        if (kCy[d] > 0) src += wym;
        if (kCy[d] < 0) src += wyp;
        if (kCz[d] > 0) src += wzm;
        if (kCz[d] < 0) src += wzp;
        if (kCx[d] != 0) {
            // x-face lane, outgoing direction: zero gradient clamps onto the face plane,
            // the other rules read the cell itself
            const bool out = kCx[d] > 0 ? xlo : xhi;
            if (out) src = xkind == LBM_FACE_ZERO_GRADIENT ? src + kCx[d] : cell;
        }
        f[d] = S::load(fin + (int64_t)d * g.P + src, d);
    }
    if (xf && xkind != LBM_FACE_ZERO_GRADIENT) {
        C own[Q];
#pragma unroll
        for (int d = 0; d < Q; ++d) own[d] = S::load(fin + (int64_t)d * g.P + cell, d);
        C rho_prev, up[3], uu2;
        own_moments<S, FORCED>(own, p, rho_prev, up, uu2);
        const int fidx = xlo ? 0 : 1;
#pragma unroll
        for (int d = 0; d < Q; ++d) {
            if (kCx[d] == 0) continue;
            const bool out = kCx[d] > 0 ? xlo : xhi;
            if (out) f[d] = face_value<C>(xkind, fidx, d, g, own[kOpp[d]], rho_prev, up, uu2);
        }
    }
    finish_cell<S, COLL, FORCED, MODE>(f, p, cell, N, g.P, fout, gout, rho_out, u_out);
}

template <class S, int COLL, bool FORCED, int MODE>
__global__ __launch_bounds__(kBlock) void k_bulk(const typename S::T* __restrict__ fin,
                                                 typename S::T* __restrict__ fout, double* __restrict__ gout,
                                                 double* __restrict__ rho_out, double* __restrict__ u_out, Grid g,
                                                 Phys<typename S::C> p, const StepGeom* __restrict__ sgp,
                                                 const uint8_t* __restrict__ flags) {
    bulk_cell<S, COLL, FORCED, MODE>(fin, fout, gout, rho_out, u_out, g, p, sgp, flags,
                                     blockIdx.x * blockDim.x + threadIdx.x, blockIdx.y * blockDim.y + threadIdx.y,
                                     blockIdx.z);
}

// ---------------------------------------------------------------------------
// Force accumulation: per thread one body in registers (a fluid node next to a
// second body accumulates that one through LDS atomics), per wave a
// deterministic segmented shuffle reduction, per block a fixed-order sum.
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
        f[d] = S::load(fin + (int64_t)d * g.P + src, d);
    }
    finish_cell<S, COLL, FORCED, MODE>(f, p, cell, N, g.P, fout, gout, rho_out, u_out);
}

// ---------------------------------------------------------------------------
// The general cell: faces (any number, with precedence), bounce-back links,
// refill, momentum exchange.  Every direction is classified first, then all
// loads are issued together, so the cell costs one memory round trip.
// ---------------------------------------------------------------------------
enum { DIR_PULL = 0, DIR_FACE = 1, DIR_LINK = 2 };

template <class S, int COLL, bool FORCED, int MODE>
__device__ __forceinline__ void general_cell(const typename S::T* __restrict__ fin, typename S::T* __restrict__ fout,
                             double* __restrict__ gout, double* __restrict__ rho_out, double* __restrict__ u_out,
                             const Grid& g, const Phys<typename S::C>& p, const SolidView& sv, int x, int y, int z,
                             ForceAcc& acc, double* lds_overflow) {
    using C = typename S::C;
    const int64_t N = g.N;
    const int64_t cell = (int64_t)x + (int64_t)g.nx * y + g.nxy * z;
    const int L[3] = {g.nx, g.ny, g.nz};
    const int q0[3] = {x, y, z};
    // solid cells are streamed like fluid ones (no links), as in the reference
    const bool self_solid = sv.nb > 0 && solid_at(sv, x, y, z, g) != 0;
    const bool fresh = sv.nb > 0 && !self_solid && g.refill && fresh_at(sv, x, y, z, g);
    C own[Q];
#pragma unroll
    for (int d = 0; d < Q; ++d) own[d] = S::load(fin + (int64_t)d * g.P + cell, d);
    C f[Q];
    if (fresh) {
        // equilibrium of the mean previous-step rho, u of fluid neighbours; all
        // neighbour loads are unrolled (unused ones read the cell itself)
        C sr = C(0), su = C(0), sv_ = C(0), sw = C(0);
        int cnt = 0;
#pragma unroll
        for (int d = 0; d < 18; ++d) {
            int q[3] = {x + kCx[d], y + kCy[d], z + kCz[d]};
            bool ok = true;
#pragma unroll
            for (int a = 0; a < 3; ++a) ok = ok && wrap_coord(q[a], L[a], (g.per >> a) & 1);
            const bool use = ok && solid_at(sv, q[0], q[1], q[2], g) == 0 && !fresh_at(sv, q[0], q[1], q[2], g);
            const int64_t nc = use ? (int64_t)q[0] + (int64_t)g.nx * q[1] + g.nxy * q[2] : cell;
            C r = C(0), mx = C(0), my = C(0), mz = C(0);
#pragma unroll
            for (int e = 0; e < Q; ++e) {
                const C v = S::load(fin + (int64_t)e * g.P + nc, e);
                r += v;
                if (kCx[e] > 0) mx += v;
                if (kCx[e] < 0) mx -= v;
                if (kCy[e] > 0) my += v;
                if (kCy[e] < 0) my -= v;
                if (kCz[e] > 0) mz += v;
                if (kCz[e] < 0) mz -= v;
            }
            C ux = mx / r, uy = my / r, uz = mz / r;
            if (FORCED) {
                ux -= C(0.5) * p.g[0];
                uy -= C(0.5) * p.g[1];
                uz -= C(0.5) * p.g[2];
            }
            if (use) {
                sr += r;
                su += ux;
                sv_ += uy;
                sw += uz;
                ++cnt;
            }
        }
        C r = C(g.rho0), ux = C(0), uy = C(0), uz = C(0);
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
            f[d] = weight<C>(d) * r * (C(1) + cu * C(3) + C(0.5) * cu * cu * C(9) - C(0.5) * u2 * C(3));
        }
        finish_cell<S, COLL, FORCED, MODE>(f, p, cell, N, g.P, fout, gout, rho_out, u_out);
        return;
    }
    // 1) classify every direction and issue its load right away (the loaded
    //    values are only consumed after the loop, so the loads overlap)
    int code[Q];  // DIR_* | (face index or body index) << 2
#pragma unroll
    for (int d = 0; d < Q; ++d) {
        const int cd[3] = {kCx[d], kCy[d], kCz[d]};
        int s[3];
        int fas[3] = {-1, -1, -1};
        bool out = false;
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
        code[d] = DIR_PULL;
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
                code[d] = DIR_FACE | (chosen << 2);
                s[0] = x;
                s[1] = y;
                s[2] = z;
            }
        } else if (!self_solid && sv.nb > 0) {
            const int sb = solid_at(sv, s[0], s[1], s[2], g);
            if (sb != 0) {
                code[d] = DIR_LINK | ((sb - 1) << 2);
                s[0] = x;
                s[1] = y;
                s[2] = z;
            }
        }
        f[d] = S::load(fin + (int64_t)d * g.P + (s[0] + g.nx * s[1] + (int)g.nxy * s[2]), d);
    }
    // 2) boundary values
    C rho_prev, up[3], uu2;
    own_moments<S, FORCED>(own, p, rho_prev, up, uu2);
#pragma unroll
    for (int d = 0; d < Q; ++d) {
        const int kd = code[d] & 3, aux = code[d] >> 2;
        if (kd == DIR_FACE) {
            f[d] = face_value<C>(face_kind_at(g, aux), aux, d, g, own[kOpp[d]], rho_prev, up, uu2);
        } else if (kd == DIR_LINK) {
            // boundary link (x, dl), dl = opp(d): the solid node is x + c_dl
            const int dl = kOpp[d];
            const DevBody& B = sv.bnew[aux];
            // wall point: physical x + c_dl/2, or the reference's x - c_dl/2
            const double hs = g.wall_point == LBM_WALL_REFERENCE ? -0.5 : 0.5;
            double r[3];
            for (int a = 0; a < 3; ++a) {
                const int cdl = a == 0 ? kCx[dl] : (a == 1 ? kCy[dl] : kCz[dl]);
                r[a] = (double)q0[a] + 0.5 + hs * (double)cdl - B.ch[a];
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
            const C fin_dl = own[dl];
            const C val = fin_dl - C(2) * weight<C>(dl) * rho_prev * cuw * C(3);
            f[d] = val;
            if (MODE == MODE_STEP) {
                // momentum exchange of this link (src/MEA.jl:102-107), FP64, arm to x + c_dl/2
                const double t = (double)(fin_dl + val);
                const double P[3] = {t * kCx[dl], t * kCy[dl], t * kCz[dl]};
                double ra[3];
                for (int a = 0; a < 3; ++a) {
                    const int cdl = a == 0 ? kCx[dl] : (a == 1 ? kCy[dl] : kCz[dl]);
                    ra[a] = (double)q0[a] + 0.5 + 0.5 * (double)cdl - B.ch[a];
                    if ((g.per >> a) & 1) {
                        const double h = 0.5 * L[a];
                        if (ra[a] > h) ra[a] -= L[a];
                        else if (ra[a] < -h) ra[a] += L[a];
                    }
                }
                const double v6[6] = {P[0], P[1], P[2], ra[1] * P[2] - ra[2] * P[1], ra[2] * P[0] - ra[0] * P[2],
                                      ra[0] * P[1] - ra[1] * P[0]};
                const int bi = aux;
                if (acc.b0 < 0 || acc.b0 == bi) {
                    acc.b0 = bi;
                    for (int c = 0; c < 6; ++c) acc.a0[c] += v6[c];
                } else {
                    for (int c = 0; c < 6; ++c) atomicAdd(&lds_overflow[bi * 6 + c], v6[c]);
                }
            }
        }
    }
    finish_cell<S, COLL, FORCED, MODE>(f, p, cell, N, g.P, fout, gout, rho_out, u_out);
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
template <class S, int COLL, bool FORCED, int MODE>
__device__ __forceinline__ void fixup_block(const int blk, const typename S::T* __restrict__ fin,
                                            typename S::T* __restrict__ fout, double* __restrict__ gout,
                                            double* __restrict__ rho_out, double* __restrict__ u_out, const Grid& g,
                                            const Phys<typename S::C>& p, const StepGeom* __restrict__ sgp,
                                            const DevBody* __restrict__ bnew, const DevBody* __restrict__ bold,
                                            uint8_t* __restrict__ flags, const int32_t* __restrict__ list,
                                            const int32_t* __restrict__ list_count, double* __restrict__ partials,
                                            int write_flags, double (*s_acc)[LBM_MAX_BODIES * 6], double* s_over) {
    const StepGeom& sg = *sgp;
    const int nb = sg.nbodies;
    const int tid = threadIdx.x;
    for (int i = tid; i < nb * 6; i += kBlock) {
        for (int w = 0; w < kWaves; ++w) s_acc[w][i] = 0.0;
        s_over[i] = 0.0;
    }
    __syncthreads();

    SolidView sv;
    sv.mode = sg.mode;
    sv.nb = nb;
    sv.flags = flags;
    sv.bnew = bnew;
    sv.bold = bold;

    ForceAcc acc;
    acc.b0 = -1;
    for (int c = 0; c < 6; ++c) acc.a0[c] = 0.0;

    bool work = false, plain = false;
    int x = 0, y = 0, z = 0;
    if (blk < sg.nfix_blocks) {
        int bi = sg.first_fix_box;
        while (bi + 1 < sg.nbox && sg.box[bi + 1].first_block <= blk) ++bi;
        const Box b = sg.box[bi];
        const int64_t t = (int64_t)(blk - b.first_block) * kBlock + tid;
        const int64_t vol = (int64_t)b.ext[0] * b.ext[1] * b.ext[2];
        if (t < vol) {
            const int i = (int)(t % b.ext[0]);
            const int64_t r = t / b.ext[0];
            const int j = (int)(r % b.ext[1]);
            const int k = (int)(r / b.ext[1]);
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
            for (int m = sg.first_fix_box; m < bi && !dup; ++m) dup = in_box(sg.box[m], x, y, z, g);
            if (!dup && sv.mode == 1 && write_flags) {
                const int fl = flag_of(bnew, nb, x, y, z, g);
                const int64_t cell = (int64_t)x + (int64_t)g.nx * y + g.nxy * z;
                if ((flags[cell] & kFlagBody) != fl) flags[cell] = (uint8_t)fl;
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
                        bool gen = false;
                        if (flag_of(bnew, nb, x, y, z, g) == 0) {
                            gen = g.refill && fresh_at(sv, x, y, z, g);
                            const int L[3] = {g.nx, g.ny, g.nz};
                            for (int d = 0; d < 18 && !gen; ++d) {
                                int q[3] = {x + kCx[d], y + kCy[d], z + kCz[d]};
                                bool ok = true;
                                for (int a = 0; a < 3; ++a) ok = ok && wrap_coord(q[a], L[a], (g.per >> a) & 1);
                                if (ok && solid_at(sv, q[0], q[1], q[2], g) != 0) gen = true;
                            }
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
    if (work && (MODE == MODE_STEP || MODE == MODE_GET)) {
        if (plain)
            plain_cell<S, COLL, FORCED, MODE>(fin, fout, gout, rho_out, u_out, g, p, x, y, z);
        else
            general_cell<S, COLL, FORCED, MODE>(fin, fout, gout, rho_out, u_out, g, p, sv, x, y, z, acc, s_over);
    }
    if (MODE == MODE_STEP && nb > 0) {
        const int wave = tid >> 6;
        wave_reduce_to_lds(acc.b0, acc.a0, s_acc[wave]);
        __syncthreads();
        for (int i = tid; i < nb * 6; i += kBlock) {
            double s = s_over[i];
            for (int w = 0; w < kWaves; ++w) s += s_acc[w][i];
            partials[(int64_t)blk * nb * 6 + i] = s;
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
    int bx, by;   // cells per block along x and rows per block
    int gx, gy;   // blocks along x and y
};

template <class S, int COLL, bool FORCED, int MODE>
__global__ __launch_bounds__(kBlock) void k_step(const typename S::T* __restrict__ fin,
                                                 typename S::T* __restrict__ fout, double* __restrict__ gout,
                                                 double* __restrict__ rho_out, double* __restrict__ u_out, Grid g,
                                                 Phys<typename S::C> p, const StepGeom* __restrict__ sgp,
                                                 const DevBody* __restrict__ bnew, const DevBody* __restrict__ bold,
                                                 uint8_t* __restrict__ flags, const int32_t* __restrict__ list,
                                                 const int32_t* __restrict__ list_count,
                                                 double* __restrict__ partials, int write_flags, BulkShape bs,
                                                 int nfix) {
    __shared__ double s_acc[kWaves][LBM_MAX_BODIES * 6];
    __shared__ double s_over[LBM_MAX_BODIES * 6];
    const int blk = blockIdx.x;
    if (blk < nfix) {
        fixup_block<S, COLL, FORCED, MODE>(blk, fin, fout, gout, rho_out, u_out, g, p, sgp, bnew, bold, flags, list,
                                           list_count, partials, write_flags, s_acc, s_over);
        return;
    }
    const int b = blk - nfix;
    const int bxi = b % bs.gx;
    const int r = b / bs.gx;
    const int byi = r % bs.gy;
    const int bz = r / bs.gy;
    const int tx = threadIdx.x % bs.bx, ty = threadIdx.x / bs.bx;
    bulk_cell<S, COLL, FORCED, MODE>(fin, fout, gout, rho_out, u_out, g, p, sgp, flags, bxi * bs.bx + tx,
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
                                                        double u0y, double u0z, int64_t N, int64_t P) {
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
        S::store(f + (int64_t)d * P + cell, d, feq);
    }
}

template <class S>
__global__ __launch_bounds__(kBlock) void k_from_double(typename S::T* __restrict__ f, const double* __restrict__ src,
                                                        int64_t N, int64_t P) {
    const int64_t cell = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (cell >= N) return;
#pragma unroll
    for (int d = 0; d < Q; ++d) S::store(f + (int64_t)d * P + cell, d, (typename S::C)src[(int64_t)d * N + cell]);
}

}  // namespace lbm
