/*
 * oracle_impl.h -- TEST INFRASTRUCTURE ONLY (the oracle).  Included twice by
 * lbm_oracle.c with REAL = double and REAL = float; SFX() names each copy.
 *
 * A literal CPU restatement of one reference time step, in the reference's order
 * and with its two population arrays (src/MainCore.jl:245-379):
 *   collision!      f_new = C(f_old; rho, u of the previous step)
 *                   TRT src/Collision.jl:136-170, BGK src/Collision.jl:212-242
 *   streaming!      f_old[d,x] = f_new[d, x - c_d]; an out-of-domain source on a
 *                   non-periodic axis copies the same node (src/Utilities.jl:757-776,
 *                   src/Streaming.jl:94-108)
 *   periodic_BC!    periodic axes wrap (src/OtherBC.jl:65-82, src/Utilities.jl:878-891)
 *   solid_BC! SBB   f_old[opp,x] = f_new[d,x] - 2 w_d rho(x) 3 (c_d . u_w)
 *                   over the link list (src/SolidBC.jl:187-208)
 *   velocity_BC!    moving wall / anti-bounce-back faces (src/OtherBC.jl:137-166)
 *   compute_aero!   F = sum c_d (f_new[d,x] + f_old[opp,x]) (src/MEA.jl:98-108)
 *   compute_macros! rho = sum f, u = sum f c / rho (src/ComputeMacro.jl:26-45)
 * Extensions that the reference lacks are stated where they occur: Guo body
 * forcing, periodic x/y, zero-gradient faces, several 3-D bodies, torque and
 * the fresh-node refill.
 */

typedef struct {
    lbm_config cfg;
    int nx, ny, nz;
    int64_t N;
    int per[3], L[3];
    int q;                           /* 19 or 27 */
    const int *cx, *cy, *cz, *opp;   /* lattice tables (src/Utilities.jl:945-993) */
    REAL w[27];
    REAL *f_old, *f_new;
    REAL *rho, *ux, *uy, *uz;
    uint8_t *flags, *fresh;
    /* CSR of boundary links, rows = cells (src/Geometry.jl:214-261) */
    int32_t* bb_ptr;
    int32_t* bb_cell;
    int8_t* bb_dir;
    int64_t nnz, cap;
    int nbodies;
    lbm_body bodies[LBM_MAX_BODIES];
    orc_body ob[LBM_MAX_BODIES];
    int nlaws;
    lbm_motion laws[LBM_MAX_BODIES];
    int geom_pending, geom_built;
    int64_t step;
    int initialised;
} SFX(orc);

/* D3Q19 weights 1/18, 1/36, 1/3 (src/Utilities.jl:954-958); D3Q27 weights
 * 2/27, 1/54, 1/216, 8/27 by the number of non-zero components (:967-973). */
static REAL SFX(orc_wgt)(int q, int d) {
    const REAL one = (REAL)1;
    if (q == 19) return OWCLS[d] == 0 ? one / (REAL)18 : (OWCLS[d] == 1 ? one / (REAL)36 : one / (REAL)3);
    int n = (OCX27[d] != 0) + (OCY27[d] != 0) + (OCZ27[d] != 0);
    return n == 0 ? (REAL)8 / (REAL)27 : (n == 1 ? (REAL)2 / (REAL)27 : (n == 2 ? one / (REAL)54 : one / (REAL)216));
}

static void* SFX(orc_create)(const lbm_config* cfg) {
    SFX(orc)* o = (SFX(orc)*)calloc(1, sizeof(SFX(orc)));
    if (!o) return NULL;
    o->cfg = *cfg;
    o->nx = cfg->nx;
    o->ny = cfg->ny;
    o->nz = cfg->nz;
    o->N = (int64_t)cfg->nx * cfg->ny * cfg->nz;
    o->L[0] = cfg->nx;
    o->L[1] = cfg->ny;
    o->L[2] = cfg->nz;
    for (int a = 0; a < 3; ++a) o->per[a] = cfg->faces[2 * a].kind == LBM_FACE_PERIODIC;
    o->q = cfg->q == 27 ? 27 : 19;
    o->cx = o->q == 27 ? OCX27 : OCX19;
    o->cy = o->q == 27 ? OCY27 : OCY19;
    o->cz = o->q == 27 ? OCZ27 : OCZ19;
    o->opp = o->q == 27 ? OOPP27 : OOPP19;
    for (int d = 0; d < o->q; ++d) o->w[d] = SFX(orc_wgt)(o->q, d);
    int64_t N = o->N;
    o->f_old = (REAL*)calloc((size_t)(o->q * N), sizeof(REAL));
    o->f_new = (REAL*)calloc((size_t)(o->q * N), sizeof(REAL));
    o->rho = (REAL*)calloc((size_t)N, sizeof(REAL));
    o->ux = (REAL*)calloc((size_t)N, sizeof(REAL));
    o->uy = (REAL*)calloc((size_t)N, sizeof(REAL));
    o->uz = (REAL*)calloc((size_t)N, sizeof(REAL));
    o->flags = (uint8_t*)calloc((size_t)N, 1);
    o->fresh = (uint8_t*)calloc((size_t)N, 1);
    o->bb_ptr = (int32_t*)calloc((size_t)(N + 1), sizeof(int32_t));
    if (!o->f_old || !o->f_new || !o->rho || !o->ux || !o->uy || !o->uz || !o->flags ||
        !o->fresh || !o->bb_ptr) {
        free(o);
        return NULL;
    }
    return o;
}

static void SFX(orc_destroy)(void* p) {
    SFX(orc)* o = (SFX(orc)*)p;
    if (!o) return;
    free(o->f_old);
    free(o->f_new);
    free(o->rho);
    free(o->ux);
    free(o->uy);
    free(o->uz);
    free(o->flags);
    free(o->fresh);
    free(o->bb_ptr);
    free(o->bb_cell);
    free(o->bb_dir);
    free(o);
}

/* Neighbour of cell (x,y,z) along c_d: wraps on periodic axes, -1 outside
 * (nbr_lli, src/Utilities.jl:556-572; the wrap is an extension). */
static int64_t SFX(orc_nbr)(const SFX(orc)* o, int x, int y, int z, int d) {
    int q[3] = {x + o->cx[d], y + o->cy[d], z + o->cz[d]};
    for (int a = 0; a < 3; ++a) {
        if (q[a] < 0 || q[a] >= o->L[a]) {
            if (!o->per[a]) return -1;
            q[a] = (q[a] + o->L[a]) % o->L[a];
        }
    }
    return (int64_t)q[0] + (int64_t)o->nx * ((int64_t)q[1] + (int64_t)o->ny * q[2]);
}

/* Solid mask, boundary-node mask and link CSR, brute force over all cells:
 * generate_mask_s_local! (src/Geometry.jl:49-87), generate_mask_bb_local!
 * (:122-148), kernel_count_bb_links! (:181-204), cumsum, ptr (:214-227) and the
 * ascending-direction fill (:233-261). */
static int SFX(orc_build_geometry)(SFX(orc)* o, int keep_flags) {
    int64_t N = o->N;
    int nx = o->nx, ny = o->ny;
    uint8_t* prev = NULL;
    if (o->cfg.refill && o->geom_built) {
        prev = (uint8_t*)malloc((size_t)N);
        if (!prev) return LBM_ERR_ARG;
        memcpy(prev, o->flags, (size_t)N);
    }
    if (!keep_flags) {
#pragma omp parallel for schedule(static)
        for (int64_t c = 0; c < N; ++c) {
            int x = (int)(c % nx), y = (int)((c / nx) % ny), z = (int)(c / ((int64_t)nx * ny));
            float px = (float)x + 0.5f, py = (float)y + 0.5f, pz = (float)z + 0.5f;
            uint8_t f = 0;
            for (int b = 0; b < o->nbodies; ++b) {
                if (orc_inside(&o->ob[b], px, py, pz, o->per, o->L)) {
                    f = (uint8_t)(b + 1);
                    break;
                }
            }
            o->flags[c] = f;
        }
    }
    memset(o->fresh, 0, (size_t)N);
    if (prev) {
        for (int64_t c = 0; c < N; ++c) o->fresh[c] = (prev[c] != 0 && o->flags[c] == 0);
        free(prev);
    }
    /* count (cnt) and inclusive scan into bb_ptr */
    o->bb_ptr[0] = 0;
    int64_t total = 0;
    for (int64_t c = 0; c < N; ++c) {
        int cnt = 0;
        if (o->flags[c] == 0) {
            int x = (int)(c % nx), y = (int)((c / nx) % ny), z = (int)(c / ((int64_t)nx * ny));
            for (int d = 0; d < o->q; ++d) {
                int64_t b = SFX(orc_nbr)(o, x, y, z, d);
                if (b >= 0 && o->flags[b] != 0) ++cnt;
            }
        }
        total += cnt;
        o->bb_ptr[c + 1] = (int32_t)total;
    }
    if (total > o->cap) {
        free(o->bb_cell);
        free(o->bb_dir);
        o->cap = total > 16 ? total : 16;
        o->bb_cell = (int32_t*)malloc((size_t)o->cap * sizeof(int32_t));
        o->bb_dir = (int8_t*)malloc((size_t)o->cap);
        if (!o->bb_cell || !o->bb_dir) return LBM_ERR_ARG;
    }
    o->nnz = total;
#pragma omp parallel for schedule(static)
    for (int64_t c = 0; c < N; ++c) {
        int64_t i = o->bb_ptr[c];
        if (i == o->bb_ptr[c + 1]) continue;
        int x = (int)(c % nx), y = (int)((c / nx) % ny), z = (int)(c / ((int64_t)nx * ny));
        for (int d = 0; d < o->q; ++d) {
            int64_t b = SFX(orc_nbr)(o, x, y, z, d);
            if (b >= 0 && o->flags[b] != 0) {
                o->bb_cell[i] = (int32_t)c;
                o->bb_dir[i] = (int8_t)d;
                ++i;
            }
        }
    }
    o->geom_built = 1;
    o->geom_pending = 0;
    return LBM_OK;
}

/* Equilibrium in the reference's BGK form w rho (1 + 3cu + 4.5cu^2 - 1.5u^2)
 * (src/Initialization.jl:37-42). */
static REAL SFX(orc_feq)(const SFX(orc)* o, REAL w, REAL rho, REAL ux, REAL uy, REAL uz, int d) {
    REAL cu = (REAL)o->cx[d] * ux + (REAL)o->cy[d] * uy + (REAL)o->cz[d] * uz;
    REAL u2 = ux * ux + uy * uy + uz * uz;
    REAL cu2 = cu * cu;
    return w * rho * ((REAL)1 + cu * (REAL)3 + (REAL)0.5 * cu2 * (REAL)9 - (REAL)0.5 * u2 * (REAL)3);
}

/* rho, u of f_old (src/ComputeMacro.jl:26-45); with a body force the velocity is
 * the Guo half-force-corrected u = sum f c / rho + g/2. */
static void SFX(orc_macros)(SFX(orc)* o) {
    int64_t N = o->N;
    REAL gx = (REAL)o->cfg.g[0], gy = (REAL)o->cfg.g[1], gz = (REAL)o->cfg.g[2];
    int forced = (gx != 0 || gy != 0 || gz != 0);
#pragma omp parallel for schedule(static)
    for (int64_t c = 0; c < N; ++c) {
        REAL r = 0, mx = 0, my = 0, mz = 0;
        for (int d = 0; d < o->q; ++d) {
            REAL f = o->f_old[(int64_t)d * N + c];
            r += f;
            mx += f * (REAL)o->cx[d];
            my += f * (REAL)o->cy[d];
            mz += f * (REAL)o->cz[d];
        }
        o->rho[c] = r;
        o->ux[c] = mx / r;
        o->uy[c] = my / r;
        o->uz[c] = mz / r;
        if (forced) {
            o->ux[c] += (REAL)0.5 * gx;
            o->uy[c] += (REAL)0.5 * gy;
            o->uz[c] += (REAL)0.5 * gz;
        }
    }
}

/* collision! over every cell, solid ones included as in the reference
 * (src/Collision.jl:95-105 has no mask).  TRT: src/Collision.jl:136-170;
 * BGK: src/Collision.jl:232-240.  Guo source (extension):
 * S_i = w_i [3 (c_i - u).F + 9 (c_i.u)(c_i.F)], F = rho g, split into even and
 * odd parts and scaled by (1 - omega_pm/2). */
static void SFX(orc_collide)(SFX(orc)* o) {
    int64_t N = o->N;
    REAL wp = (REAL)o->cfg.omega_plus, wm = (REAL)o->cfg.omega_minus;
    int trt = o->cfg.collision == LBM_TRT;
    REAL gx = (REAL)o->cfg.g[0], gy = (REAL)o->cfg.g[1], gz = (REAL)o->cfg.g[2];
    int forced = (gx != 0 || gy != 0 || gz != 0);
#pragma omp parallel for schedule(static)
    for (int64_t c = 0; c < N; ++c) {
        REAL rho = o->rho[c], ux = o->ux[c], uy = o->uy[c], uz = o->uz[c];
        REAL u2 = ux * ux + uy * uy + uz * uz;
        REAL Fx = rho * gx, Fy = rho * gy, Fz = rho * gz;
        REAL uF = ux * Fx + uy * Fy + uz * Fz;
        for (int d = 0; d < o->q; ++d) {
            REAL w = o->w[d];
            REAL fi = o->f_old[(int64_t)d * N + c];
            REAL cu = (REAL)o->cx[d] * ux + (REAL)o->cy[d] * uy + (REAL)o->cz[d] * uz;
            REAL cu2 = cu * cu;
            REAL out;
            REAL Sp = 0, Sm = 0;
            if (forced) {
                REAL cF = (REAL)o->cx[d] * Fx + (REAL)o->cy[d] * Fy + (REAL)o->cz[d] * Fz;
                Sp = w * ((REAL)9 * cu * cF - (REAL)3 * uF);
                Sm = w * ((REAL)3 * cF);
            }
            if (trt) {
                REAL fo = o->f_old[(int64_t)o->opp[d] * N + c];
                REAL feq_p = w * rho * ((REAL)1 + (REAL)0.5 * cu2 * (REAL)9 - (REAL)0.5 * u2 * (REAL)3);
                REAL feq_m = w * rho * (cu * (REAL)3);
                REAL fp = (REAL)0.5 * (fi + fo);
                REAL fm = (REAL)0.5 * (fi - fo);
                REAL fpn = fp - wp * (fp - feq_p);
                REAL fmn = fm - wm * (fm - feq_m);
                out = fpn + fmn;
                if (forced) out += ((REAL)1 - (REAL)0.5 * wp) * Sp + ((REAL)1 - (REAL)0.5 * wm) * Sm;
            } else {
                REAL feq = w * rho * ((REAL)1 + (REAL)3 * cu + (REAL)4.5 * cu2 - (REAL)1.5 * u2);
                out = fi - wp * (fi - feq);
                if (forced) out += ((REAL)1 - (REAL)0.5 * wp) * (Sp + Sm);
            }
            o->f_new[(int64_t)d * N + c] = out;
        }
    }
}

/* streaming! + periodic_BC!: pull from x - c_d, wrapping periodic axes; a source
 * outside a non-periodic face copies the same node, as build_stream_lists does
 * for X/Y (src/Utilities.jl:773-776); the face pass below overwrites those. */
static void SFX(orc_stream)(SFX(orc)* o) {
    int64_t N = o->N;
    int nx = o->nx, ny = o->ny, nz = o->nz;
#pragma omp parallel for schedule(static)
    for (int64_t c = 0; c < N; ++c) {
        int x = (int)(c % nx), y = (int)((c / nx) % ny), z = (int)(c / ((int64_t)nx * ny));
        for (int d = 0; d < o->q; ++d) {
            int s[3] = {x - o->cx[d], y - o->cy[d], z - o->cz[d]};
            int out = 0;
            for (int a = 0; a < 3; ++a) {
                if (s[a] < 0 || s[a] >= o->L[a]) {
                    if (o->per[a]) s[a] = (s[a] + o->L[a]) % o->L[a];
                    else out = 1;
                }
            }
            int64_t src = out ? c : (int64_t)s[0] + (int64_t)nx * ((int64_t)s[1] + (int64_t)ny * s[2]);
            o->f_old[(int64_t)d * N + c] = o->f_new[(int64_t)d * N + src];
        }
    }
    (void)nz;
}

/* Wall velocity u_w = U + Omega x (x_w - X) at the half step
 * (uw_at_wall_lat_planar, src/Geometry.jl:412-419, generalised to 3-D);
 * x_w = x + h c_d (physical) or x - h c_d (src/SolidBC.jl:195-196), h = 1/2 on a
 * simple bounce-back link and q on an interpolated one (src/SolidBC.jl:324). */
static void SFX(orc_wall_velocity)(const SFX(orc)* o, const orc_body* b, int x, int y, int z,
                                   int d, int mode, REAL h, REAL uw[3]) {
    REAL sgn = mode == LBM_WALL_REFERENCE ? -h : h;
    REAL p[3] = {(REAL)x + (REAL)0.5 + sgn * (REAL)o->cx[d], (REAL)y + (REAL)0.5 + sgn * (REAL)o->cy[d],
                 (REAL)z + (REAL)0.5 + sgn * (REAL)o->cz[d]};
    REAL r[3];
    for (int a = 0; a < 3; ++a) {
        r[a] = p[a] - (REAL)b->ch[a];
        if (o->per[a]) {
            REAL h = (REAL)0.5 * (REAL)o->L[a];
            if (r[a] > h) r[a] -= (REAL)o->L[a];
            else if (r[a] < -h) r[a] += (REAL)o->L[a];
        }
    }
    REAL om[3] = {(REAL)b->om[0], (REAL)b->om[1], (REAL)b->om[2]};
    if (orc_planted == ORC_PLANT_SPIN_SIGN)
        for (int a = 0; a < 3; ++a) om[a] = -om[a];
    uw[0] = (REAL)b->v[0] + (om[1] * r[2] - om[2] * r[1]);
    uw[1] = (REAL)b->v[1] + (om[2] * r[0] - om[0] * r[2]);
    uw[2] = (REAL)b->v[2] + (om[0] * r[1] - om[1] * r[0]);
}

/* Rule of the link (cell c, direction d into the solid) under the interpolated
 * bounce-back (IBB!, src/SolidBC.jl:253-350, with the solid-side point taken
 * at x + c_d and the second fluid node at x - c_d): 0 = simple bounce-back
 * (no sign change of the level set, or q < 1/2 without a fluid second node),
 * 1 = q < 1/2, 2 = q >= 1/2.  q = phi(x) / (phi(x) - phi(x + c_d)) in FP32,
 * clamped to [1e-3, 1 - 1e-3] (EPS_Q, undefined in the reference). */
static int SFX(orc_link_rule)(const SFX(orc)* o, int x, int y, int z, int d, const orc_body* b, float* q_out,
                              int64_t* second) {
    *q_out = 0.5f;
    *second = -1;
    if (o->cfg.solid_bc != LBM_SOLID_IBB) return 0;
    float pb[3] = {(float)x + 0.5f, (float)y + 0.5f, (float)z + 0.5f};
    float phb = orc_phi(b, pb[0], pb[1], pb[2], o->per, o->L);
    float phs = orc_phi(b, pb[0] + (float)o->cx[d], pb[1] + (float)o->cy[d], pb[2] + (float)o->cz[d], o->per, o->L);
    float den = phb - phs;
    if (!(phb > 0.0f && phs < 0.0f && fabsf(den) >= 1e-8f)) return 0;
    float q = phb / den;
    if (q < 1e-3f) q = 1e-3f;
    if (q > 1.0f - 1e-3f) q = 1.0f - 1e-3f;
    if (q >= 0.5f) {
        *q_out = q;
        return 2;
    }
    int64_t s2 = SFX(orc_nbr)(o, x, y, z, o->opp[d]);
    if (s2 < 0 || o->flags[s2] != 0) return 0;
    *q_out = q;
    *second = s2;
    return 1;
}

/* solid_BC! over the CSR: SBB! (src/SolidBC.jl:187-208) or IBB!
 * (src/SolidBC.jl:253-350) with the moving-wall correction of Lallemand & Luo
 * (2003), corr / (2q) on the q >= 1/2 branch. */
static void SFX(orc_sbb)(SFX(orc)* o) {
    int64_t N = o->N;
    int nx = o->nx, ny = o->ny;
#pragma omp parallel for schedule(static)
    for (int64_t k = 0; k < o->nnz; ++k) {
        int64_t c = o->bb_cell[k];
        int d = o->bb_dir[k];
        if (o->fresh[c]) continue;
        int x = (int)(c % nx), y = (int)((c / nx) % ny), z = (int)(c / ((int64_t)nx * ny));
        int64_t s = SFX(orc_nbr)(o, x, y, z, d);
        const orc_body* b = &o->ob[o->flags[s] - 1];
        float qf;
        int64_t c2;
        int rule = SFX(orc_link_rule)(o, x, y, z, d, b, &qf, &c2);
        REAL uw[3];
        SFX(orc_wall_velocity)(o, b, x, y, z, d, o->cfg.wall_point, rule == 0 ? (REAL)0.5 : (REAL)qf, uw);
        REAL cuw = (REAL)o->cx[d] * uw[0] + (REAL)o->cy[d] * uw[1] + (REAL)o->cz[d] * uw[2];
        REAL corr = (REAL)2 * o->w[d] * o->rho[c] * cuw * (REAL)3;
        if (orc_planted == ORC_PLANT_WALL_HALF) corr *= (REAL)0.5;
        REAL fd = o->f_new[(int64_t)d * N + c];
        REAL val;
        if (rule == 1) {
            REAL q2 = (REAL)2 * (REAL)qf;
            val = q2 * fd + ((REAL)1 - q2) * o->f_new[(int64_t)d * N + c2] - corr;
        } else if (rule == 2) {
            REAL q2 = (REAL)2 * (REAL)qf;
            val = (fd + (q2 - (REAL)1) * o->f_new[(int64_t)o->opp[d] * N + c] - corr) / q2;
        } else {
            val = fd - corr;
        }
        o->f_old[(int64_t)o->opp[d] * N + c] = val;
    }
}

/* velocity_BC! (src/OtherBC.jl:137-166) on every non-periodic face, with the
 * deterministic edge precedence VELOCITY > PRESSURE_ABB > ZERO_GRADIENT, lower
 * axis first; zero-gradient (extension) pulls from the source clamped onto the
 * face. */
static void SFX(orc_faces)(SFX(orc)* o) {
    int64_t N = o->N;
    int nx = o->nx, ny = o->ny;
    if (o->per[0] && o->per[1] && o->per[2]) return;
#pragma omp parallel for schedule(static)
    for (int64_t c = 0; c < N; ++c) {
        int x = (int)(c % nx), y = (int)((c / nx) % ny), z = (int)(c / ((int64_t)nx * ny));
        int q[3] = {x, y, z};
        int onface = 0;
        for (int a = 0; a < 3; ++a)
            if (!o->per[a] && (q[a] == 0 || q[a] == o->L[a] - 1)) onface = 1;
        if (!onface) continue;
        for (int d = 0; d < o->q; ++d) {
            int cd[3] = {o->cx[d], o->cy[d], o->cz[d]};
            int s[3];
            int face_of_axis[3] = {-1, -1, -1};
            int any = 0;
            for (int a = 0; a < 3; ++a) {
                s[a] = q[a] - cd[a];
                if (s[a] < 0 || s[a] >= o->L[a]) {
                    if (o->per[a]) {
                        s[a] = (s[a] + o->L[a]) % o->L[a];
                    } else {
                        face_of_axis[a] = 2 * a + (s[a] >= o->L[a] ? 1 : 0);
                        any = 1;
                    }
                }
            }
            if (!any) continue;
            int chosen = -1, kind = -1;
            const int order[3] = {LBM_FACE_VELOCITY, LBM_FACE_PRESSURE_ABB, LBM_FACE_ZERO_GRADIENT};
            for (int k = 0; k < 3 && chosen < 0; ++k)
                for (int a = 0; a < 3; ++a)
                    if (face_of_axis[a] >= 0 && o->cfg.faces[face_of_axis[a]].kind == order[k]) {
                        chosen = face_of_axis[a];
                        kind = order[k];
                        break;
                    }
            REAL w = o->w[d];
            REAL val;
            if (kind == LBM_FACE_VELOCITY) {
                const double* U = o->cfg.faces[chosen].u;
                REAL cu = (REAL)cd[0] * (-(REAL)U[0]) + (REAL)cd[1] * (-(REAL)U[1]) + (REAL)cd[2] * (-(REAL)U[2]);
                val = o->f_new[(int64_t)o->opp[d] * N + c] - (REAL)2 * w * o->rho[c] * cu * (REAL)3;
            } else if (kind == LBM_FACE_PRESSURE_ABB) {
                REAL uw = o->ux[c], vw = o->uy[c], ww = o->uz[c];
                REAL cu = (REAL)cd[0] * uw + (REAL)cd[1] * vw + (REAL)cd[2] * ww;
                REAL cu2 = cu * cu, uu2 = uw * uw + vw * vw + ww * ww;
                val = -o->f_new[(int64_t)o->opp[d] * N + c] +
                      (REAL)2 * w * o->rho[c] * ((REAL)1 + (REAL)0.5 * cu2 * (REAL)9 - (REAL)0.5 * uu2 * (REAL)3);
            } else {
                /* every out axis is zero-gradient: clamp onto the face */
                for (int a = 0; a < 3; ++a) {
                    if (face_of_axis[a] >= 0) s[a] = q[a];
                }
                int64_t src = (int64_t)s[0] + (int64_t)nx * ((int64_t)s[1] + (int64_t)ny * s[2]);
                val = o->f_new[(int64_t)d * N + src];
            }
            o->f_old[(int64_t)d * N + c] = val;
        }
    }
}

/* Fresh-node refill (extension): a cell that turned fluid this step takes the
 * equilibrium of the mean rho, u (previous step) of its fluid neighbours. */
static void SFX(orc_refill)(SFX(orc)* o) {
    if (!o->cfg.refill) return;
    int64_t N = o->N;
    int nx = o->nx, ny = o->ny;
    for (int64_t c = 0; c < N; ++c) {
        if (!o->fresh[c]) continue;
        int x = (int)(c % nx), y = (int)((c / nx) % ny), z = (int)(c / ((int64_t)nx * ny));
        REAL sr = 0, su = 0, sv = 0, sw = 0;
        int cnt = 0;
        for (int d = 0; d < o->q - 1; ++d) {
            int64_t b = SFX(orc_nbr)(o, x, y, z, d);
            if (b < 0 || o->flags[b] != 0 || o->fresh[b]) continue;
            sr += o->rho[b];
            su += o->ux[b];
            sv += o->uy[b];
            sw += o->uz[b];
            ++cnt;
        }
        REAL r, u, v, ww;
        if (cnt > 0) {
            REAL inv = (REAL)cnt;
            r = sr / inv;
            u = su / inv;
            v = sv / inv;
            ww = sw / inv;
        } else {
            r = (REAL)o->cfg.rho0;
            u = v = ww = 0;
        }
        for (int d = 0; d < o->q; ++d) o->f_old[(int64_t)d * N + c] = SFX(orc_feq)(o, o->w[d], r, u, v, ww, d);
    }
}

/* compute_aero! (src/MEA.jl:98-108), per body, FP64 accumulation, plus the
 * torque (x + c_d/2 - X_half) x dP (extension). out: nbodies * 6. */
static void SFX(orc_mea)(SFX(orc)* o, double* out) {
    int64_t N = o->N;
    int nx = o->nx, ny = o->ny;
    int nb = o->nbodies;
    for (int i = 0; i < 6 * nb; ++i) out[i] = 0.0;
    for (int64_t k = 0; k < o->nnz; ++k) {
        int64_t c = o->bb_cell[k];
        int d = o->bb_dir[k];
        if (o->fresh[c]) continue;
        int x = (int)(c % nx), y = (int)((c / nx) % ny), z = (int)(c / ((int64_t)nx * ny));
        int64_t s = SFX(orc_nbr)(o, x, y, z, d);
        int bi = o->flags[s] - 1;
        const orc_body* b = &o->ob[bi];
        REAL tmp = o->f_new[(int64_t)d * N + c] + o->f_old[(int64_t)o->opp[d] * N + c];
        double t = (double)tmp;
        double P[3] = {t * o->cx[d], t * o->cy[d], t * o->cz[d]};
        /* arm to the wall point x + h c_d: h = 1/2 (SBB) or q (IBB) */
        float qf;
        int64_t c2;
        double h = SFX(orc_link_rule)(o, x, y, z, d, b, &qf, &c2) == 0 ? 0.5 : (double)qf;
        if (orc_planted == ORC_PLANT_ARM_MIRROR) h = -h;
        double r[3] = {(double)x + 0.5 + h * o->cx[d] - b->ch[0], (double)y + 0.5 + h * o->cy[d] - b->ch[1],
                       (double)z + 0.5 + h * o->cz[d] - b->ch[2]};
        if (orc_planted == ORC_PLANT_ARM_SIGN)
            for (int a = 0; a < 3; ++a) r[a] = -r[a];
        for (int a = 0; a < 3; ++a) {
            if (o->per[a]) {
                double h = 0.5 * o->L[a];
                if (r[a] > h) r[a] -= o->L[a];
                else if (r[a] < -h) r[a] += o->L[a];
            }
        }
        double* ob = out + 6 * bi;
        ob[0] += P[0];
        ob[1] += P[1];
        ob[2] += P[2];
        ob[3] += r[1] * P[2] - r[2] * P[1];
        ob[4] += r[2] * P[0] - r[0] * P[2];
        ob[5] += r[0] * P[1] - r[1] * P[0];
    }
}

static int SFX(orc_set_bodies)(void* p, int n, const lbm_body* b) {
    SFX(orc)* o = (SFX(orc)*)p;
    if (n < 0 || n > LBM_MAX_BODIES) return LBM_ERR_ARG;
    o->nbodies = n;
    o->nlaws = 0;
    for (int i = 0; i < n; ++i) {
        o->bodies[i] = b[i];
        orc_prepare_body(&b[i], &o->ob[i]);
    }
    o->geom_pending = 1;
    return LBM_OK;
}

static int SFX(orc_set_motion)(void* p, int n, const lbm_motion* m) {
    SFX(orc)* o = (SFX(orc)*)p;
    if (n < 0 || n > LBM_MAX_BODIES) return LBM_ERR_ARG;
    o->nlaws = n;
    o->nbodies = n;
    for (int i = 0; i < n; ++i) o->laws[i] = m[i];
    o->geom_pending = 1;
    return LBM_OK;
}

static void SFX(orc_apply_laws)(SFX(orc)* o, int64_t n) {
    for (int i = 0; i < o->nlaws; ++i) {
        orc_body_from_law(&o->laws[i], n, &o->bodies[i]);
        orc_prepare_body(&o->bodies[i], &o->ob[i]);
    }
}

static int SFX(orc_set_flags)(void* p, const uint8_t* flags) {
    SFX(orc)* o = (SFX(orc)*)p;
    o->nbodies = 0;
    o->nlaws = 0;
    int maxf = 0;
    for (int64_t c = 0; c < o->N; ++c)
        if (flags[c] > maxf) maxf = flags[c];
    if (maxf > LBM_MAX_BODIES) return LBM_ERR_ARG;
    /* bodies without kinematics: at rest, centre at the origin */
    o->nbodies = maxf;
    for (int i = 0; i < maxf; ++i) {
        memset(&o->bodies[i], 0, sizeof(lbm_body));
        memset(&o->ob[i], 0, sizeof(orc_body));
    }
    memcpy(o->flags, flags, (size_t)o->N);
    return SFX(orc_build_geometry)(o, 1);
}

static int SFX(orc_update_geometry)(void* p) {
    SFX(orc)* o = (SFX(orc)*)p;
    if (o->nlaws) SFX(orc_apply_laws)(o, o->step + 1);
    return SFX(orc_build_geometry)(o, 0);
}

static int SFX(orc_init)(void* p, const double* rho, const double* u) {
    SFX(orc)* o = (SFX(orc)*)p;
    int64_t N = o->N;
#pragma omp parallel for schedule(static)
    for (int64_t c = 0; c < N; ++c) {
        REAL r = rho ? (REAL)rho[c] : (REAL)o->cfg.rho0;
        REAL ux = u ? (REAL)u[c] : (REAL)o->cfg.u0[0];
        REAL uy = u ? (REAL)u[N + c] : (REAL)o->cfg.u0[1];
        REAL uz = u ? (REAL)u[2 * N + c] : (REAL)o->cfg.u0[2];
        for (int d = 0; d < o->q; ++d) o->f_old[(int64_t)d * N + c] = SFX(orc_feq)(o, o->w[d], r, ux, uy, uz, d);
    }
    SFX(orc_macros)(o);
    o->step = 0;
    o->initialised = 1;
    return LBM_OK;
}

static int SFX(orc_set_populations)(void* p, const double* f) {
    SFX(orc)* o = (SFX(orc)*)p;
    for (int64_t i = 0; i < (int64_t)o->q * o->N; ++i) o->f_old[i] = (REAL)f[i];
    SFX(orc_macros)(o);
    o->step = 0;
    o->initialised = 1;
    return LBM_OK;
}

/* One reference step n = step + 1 (src/MainCore.jl:245-360). */
static int SFX(orc_step)(void* p, int nsteps, double* force_out) {
    SFX(orc)* o = (SFX(orc)*)p;
    if (!o->initialised) return LBM_ERR_STATE;
    for (int s = 0; s < nsteps; ++s) {
        int64_t n = o->step + 1;
        if (o->nlaws) {
            SFX(orc_apply_laws)(o, n);
            o->geom_pending = 1;
        }
        if (o->geom_pending) {
            int rc = SFX(orc_build_geometry)(o, 0);
            if (rc) return rc;
        } else if (o->cfg.refill) {
            memset(o->fresh, 0, (size_t)o->N);
        }
        SFX(orc_collide)(o);
        SFX(orc_stream)(o);
        SFX(orc_sbb)(o);
        SFX(orc_faces)(o);
        SFX(orc_refill)(o);
        if (force_out) SFX(orc_mea)(o, force_out + (size_t)s * 6 * o->nbodies);
        SFX(orc_macros)(o);
        o->step = n;
    }
    return LBM_OK;
}

static int SFX(orc_get_macros)(void* p, double* rho, double* u) {
    SFX(orc)* o = (SFX(orc)*)p;
    int64_t N = o->N;
    for (int64_t c = 0; c < N; ++c) {
        if (rho) rho[c] = (double)o->rho[c];
        if (u) {
            u[c] = (double)o->ux[c];
            u[N + c] = (double)o->uy[c];
            u[2 * N + c] = (double)o->uz[c];
        }
    }
    return LBM_OK;
}

static int SFX(orc_get_populations)(void* p, double* f) {
    SFX(orc)* o = (SFX(orc)*)p;
    for (int64_t i = 0; i < (int64_t)o->q * o->N; ++i) f[i] = (double)o->f_old[i];
    return LBM_OK;
}

static int SFX(orc_get_flags)(void* p, uint8_t* flags) {
    SFX(orc)* o = (SFX(orc)*)p;
    memcpy(flags, o->flags, (size_t)o->N);
    return LBM_OK;
}

static int SFX(orc_get_links)(void* p, int32_t* cell, int8_t* dir, int64_t cap, int64_t* nnz) {
    SFX(orc)* o = (SFX(orc)*)p;
    if (nnz) *nnz = o->nnz;
    if (!cell && !dir) return LBM_OK;
    if (o->nnz > cap) return LBM_ERR_CAPACITY;
    for (int64_t k = 0; k < o->nnz; ++k) {
        if (cell) cell[k] = o->bb_cell[k];
        if (dir) dir[k] = o->bb_dir[k];
    }
    return LBM_OK;
}

static int SFX(orc_conservation)(void* p, double* mass, double* mom) {
    SFX(orc)* o = (SFX(orc)*)p;
    double m = 0, px = 0, py = 0, pz = 0;
    for (int64_t c = 0; c < o->N; ++c) {
        if (o->flags[c]) continue;
        double r = (double)o->rho[c];
        m += r;
        px += r * (double)o->ux[c];
        py += r * (double)o->uy[c];
        pz += r * (double)o->uz[c];
    }
    if (mass) *mass = m;
    if (mom) {
        mom[0] = px;
        mom[1] = py;
        mom[2] = pz;
    }
    return LBM_OK;
}
