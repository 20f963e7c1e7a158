/*
 * lbm_oracle.c -- TEST INFRASTRUCTURE ONLY: the oracle.
 *
 * CPU restatement of the reference's per-step LBM path (EricVaizman/LBM3D_BB,
 * Julia/CUDA.jl), in FP64 and FP32, with an API that mirrors include/lbm_hip.h
 * so a test drives both through the same script.  Only tests/,
 * __graft_entry__.smoke() and bench.py's CPU-baseline leg load it, as the
 * checker; the shipped library never does.  The reference itself cannot be
 * built or run here (Julia + CUDA.jl on NVIDIA); every function below cites the
 * reference lines it restates, and tests/ pins each against closed forms,
 * invariants and brute-force enumerations.
 *
 * Build: oracle/Makefile (gcc -O3 -fopenmp -ffp-contract=off).
 */
#include <stdio.h>

#include "oracle_geom.h"

#define REAL double
#define SFX(n) n##_f64
#include "oracle_impl.h"
#undef REAL
#undef SFX

#define REAL float
#define SFX(n) n##_f32
#include "oracle_impl.h"
#undef REAL
#undef SFX

typedef struct {
    int prec; /* 0 = FP64, 1 = FP32 */
    void* impl;
} orc_handle;

#define DISPATCH(h, fn, ...) ((h)->prec == 0 ? fn##_f64((h)->impl, __VA_ARGS__) : fn##_f32((h)->impl, __VA_ARGS__))

int orc_create(const lbm_config* cfg, int prec, orc_handle** out) {
    if (!cfg || !out || (prec != 0 && prec != 1)) return LBM_ERR_ARG;
    if (cfg->nx < 2 || cfg->ny < 2 || cfg->nz < 2) return LBM_ERR_ARG;
    if (cfg->q != 0 && cfg->q != 19 && cfg->q != 27) return LBM_ERR_ARG;
    for (int a = 0; a < 3; ++a) {
        int lo = cfg->faces[2 * a].kind == LBM_FACE_PERIODIC, hi = cfg->faces[2 * a + 1].kind == LBM_FACE_PERIODIC;
        if (lo != hi) return LBM_ERR_ARG;
    }
    orc_handle* h = (orc_handle*)calloc(1, sizeof(orc_handle));
    if (!h) return LBM_ERR_ARG;
    h->prec = prec;
    h->impl = prec == 0 ? orc_create_f64(cfg) : orc_create_f32(cfg);
    if (!h->impl) {
        free(h);
        return LBM_ERR_ARG;
    }
    *out = h;
    return LBM_OK;
}

int orc_destroy(orc_handle* h) {
    if (!h) return LBM_OK;
    if (h->prec == 0) orc_destroy_f64(h->impl);
    else orc_destroy_f32(h->impl);
    free(h);
    return LBM_OK;
}

int orc_set_bodies(orc_handle* h, int32_t n, const lbm_body* b) { return DISPATCH(h, orc_set_bodies, n, b); }
int orc_set_motion(orc_handle* h, int32_t n, const lbm_motion* m) { return DISPATCH(h, orc_set_motion, n, m); }
int orc_set_flags(orc_handle* h, const uint8_t* f) { return DISPATCH(h, orc_set_flags, f); }
int orc_update_geometry(orc_handle* h) {
    return h->prec == 0 ? orc_update_geometry_f64(h->impl) : orc_update_geometry_f32(h->impl);
}
int orc_get_flags(orc_handle* h, uint8_t* f) { return DISPATCH(h, orc_get_flags, f); }
int orc_get_links(orc_handle* h, int32_t* cell, int8_t* dir, int64_t cap, int64_t* nnz) {
    return DISPATCH(h, orc_get_links, cell, dir, cap, nnz);
}
int orc_init_equilibrium(orc_handle* h, const double* rho, const double* u) { return DISPATCH(h, orc_init, rho, u); }
int orc_set_populations(orc_handle* h, const double* f) { return DISPATCH(h, orc_set_populations, f); }
int orc_step(orc_handle* h, int32_t nsteps, double* force_out) { return DISPATCH(h, orc_step, nsteps, force_out); }
int orc_get_macros(orc_handle* h, double* rho, double* u) { return DISPATCH(h, orc_get_macros, rho, u); }
int orc_get_populations(orc_handle* h, double* f) { return DISPATCH(h, orc_get_populations, f); }
int orc_conservation(orc_handle* h, double* mass, double* mom) { return DISPATCH(h, orc_conservation, mass, mom); }
/* lbm_sync's counterpart: the oracle runs synchronously. */
int orc_sync(orc_handle* h) { return h ? LBM_OK : LBM_ERR_ARG; }

/* Pose of a motion law at step n (1-based), exposed so tests can pin it. */
int orc_body_at(const lbm_motion* m, int64_t n, lbm_body* out) {
    if (!m || !out) return LBM_ERR_ARG;
    orc_body_from_law(m, n, out);
    return LBM_OK;
}

/* The FP32 inside test alone, for brute-force and closed-form checks. */
int orc_inside_test(const lbm_body* b, float px, float py, float pz, const int32_t per[3], const int32_t L[3]) {
    orc_body ob;
    orc_prepare_body(b, &ob);
    int p[3] = {per[0], per[1], per[2]}, l[3] = {L[0], L[1], L[2]};
    return orc_inside(&ob, px, py, pz, p, l);
}

/* The FP32 level set alone (interpolated bounce-back wall fractions). */
float orc_phi_test(const lbm_body* b, float px, float py, float pz, const int32_t per[3], const int32_t L[3]) {
    orc_body ob;
    orc_prepare_body(b, &ob);
    int p[3] = {per[0], per[1], per[2]}, l[3] = {L[0], L[1], L[2]};
    return orc_phi(&ob, px, py, pz, p, l);
}

/* The FP32 NACA 4-digit lower / upper surface at chordwise position x. */
void orc_naca_test(float x, float chord, float m, float p, float t, float x_int, float* yl, float* yu) {
    orc_naca_lines(x, chord, m, p, t, x_int, yl, yu);
}

/* Plants one known mistake (ORC_PLANT_*, oracle_geom.h) in every later step, or
 * removes it (0); tests use it to show that a pin fails on a wrong oracle. */
int orc_plant(int code) {
    if (code < ORC_PLANT_NONE || code > ORC_PLANT_WALL_HALF) return LBM_ERR_ARG;
    orc_planted = code;
    return LBM_OK;
}

int orc_num_threads(void) {
#ifdef _OPENMP
    extern int omp_get_max_threads(void);
    return omp_get_max_threads();
#else
    return 1;
#endif
}
