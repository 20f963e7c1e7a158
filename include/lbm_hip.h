/*
 * lbm_hip.h -- C ABI of the MI355X (gfx950) D3Q19 lattice-Boltzmann hot path.
 *
 * One call of lbm_step() replaces the body of the reference time loop
 * (src/MainCore.jl:245-379): motion! -> collision! -> streaming! -> periodic_BC!
 * -> solid_BC! -> velocity_BC! -> compute_aero! -> compute_macros!.  The
 * reference has no FFI layer: its operators are Julia bang-functions on CuArrays
 * (collision! src/Collision.jl:14-31, streaming! src/Streaming.jl:116-124,
 * periodic_BC! src/OtherBC.jl:87-94, solid_BC! src/SolidBC.jl:129-142,
 * velocity_BC! src/OtherBC.jl:174-194, compute_aero! src/MEA.jl:14-19,
 * compute_macros! src/ComputeMacro.jl:53-63, motion! src/Motion.jl:41-56).  The
 * entry points below are what a Julia `ccall` shim binds in their place.
 *
 * Conventions
 *   - every function returns 0 on success and a negative lbm_status on error;
 *     lbm_last_error() returns the message of the last failing call;
 *   - host buffers are caller-owned, device memory is context-owned;
 *   - cells are numbered cell = x + Nx*(y + Ny*z), 0-based, x fastest (the
 *     reference's global linear index (z-1)NxNy + (y-1)Nx + x minus one,
 *     src/Utilities.jl:766);
 *   - directions are the reference's D3Q19 order minus one
 *     (src/Utilities.jl:950-960): 0..17 moving, 18 at rest;
 *   - populations are structure-of-arrays: f[d*N + cell];
 *   - node coordinates are (i + 0.5) lattice units for 0-based index i (the
 *     reference's (i - 1/2)h for 1-based i, src/Utilities.jl:550-551, h = 1);
 *   - calls are ordered on the context's own HIP stream; getters synchronise;
 *   - one context per host thread; the library keeps no global state.
 */
#ifndef LBM_HIP_H
#define LBM_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LBM_Q 19  /* populations per cell of the default lattice; a D3Q27 context holds 27 */
#define LBM_MAX_BODIES 127

typedef enum {
    LBM_OK = 0,
    LBM_ERR_ARG = -1,      /* invalid argument or configuration          */
    LBM_ERR_HIP = -2,      /* a HIP runtime call failed                  */
    LBM_ERR_STATE = -3,    /* call out of order (e.g. step before init)  */
    LBM_ERR_CAPACITY = -4, /* caller buffer too small                    */
    LBM_ERR_NUMERIC = -5   /* lbm_step: a body's force/torque record is  */
                           /* not finite (NaN or Inf reached its links)  */
} lbm_status;

/* Collision operator (src/Collision.jl:50-69 "BGK" / "TRT"). */
typedef enum { LBM_BGK = 0, LBM_TRT = 1 } lbm_collision;

/* Population storage in HBM.  FP16S/BF16S store f_i - w_i (shifted by the
 * lattice weight, rho0 = 1) in 16 bits and collide in FP32. */
typedef enum { LBM_FP64 = 0, LBM_FP32 = 1, LBM_FP16S = 2, LBM_BF16S = 3 } lbm_storage;

/* Face boundary conditions.
 *   PERIODIC       wrap (both faces of an axis must be periodic);
 *   VELOCITY       moving-wall bounce-back f_d = f*_opp + 6 w_d rho (c_d . u)
 *                  (src/OtherBC.jl:138-149); u = 0 is a no-slip wall;
 *   PRESSURE_ABB   anti-bounce-back with the local rho, u (src/OtherBC.jl:152-166);
 *   ZERO_GRADIENT  unknown populations copied from the face plane itself
 *                  (source coordinate clamped onto the face).
 * Precedence at edges and corners (the reference races there,
 * src/OtherBC.jl:138-166): a VELOCITY face wins over PRESSURE_ABB, which wins
 * over ZERO_GRADIENT; within one kind the lower axis (x < y < z) wins. */
typedef enum {
    LBM_FACE_PERIODIC = 0,
    LBM_FACE_VELOCITY = 1,
    LBM_FACE_PRESSURE_ABB = 2,
    LBM_FACE_ZERO_GRADIENT = 3
} lbm_face_kind;

typedef struct {
    int32_t kind;   /* lbm_face_kind */
    int32_t pad_;
    double u[3];    /* wall / inlet velocity for LBM_FACE_VELOCITY */
} lbm_face;

/* Wall point used for the moving-wall velocity of a boundary link (x, d). */
typedef enum {
    LBM_WALL_PHYSICAL = 0,  /* x + c_d/2, between the fluid node and the solid node */
    LBM_WALL_REFERENCE = 1  /* x - c_d/2, as src/SolidBC.jl:195-196 samples it      */
} lbm_wall_point;

/* Solid boundary rule of the body links (solid_BC!'s bc argument,
 * src/SolidBC.jl:129-153, "SBB" / "IBB").
 *   SBB  simple (halfway) bounce-back with the moving-wall correction
 *        f_opp(x) = f*_d(x) - 6 w_d rho (c_d . u_w)   (src/SolidBC.jl:187-208);
 *   IBB  linear interpolated bounce-back (src/SolidBC.jl:253-350) with the wall
 *        fraction q = phi(x) / (phi(x) - phi(x + c_d)) of the body's level set
 *        on the link from the fluid node x to the solid node x + c_d, clamped to
 *        [1e-3, 1 - 1e-3]:
 *          q <  1/2: f_opp(x) = 2q f*_d(x) + (1 - 2q) f*_d(x - c_d) - corr
 *          q >= 1/2: f_opp(x) = (f*_d(x) + (2q - 1) f*_opp(x) - corr) / (2q)
 *        corr = 6 w_d rho (c_d . u_w(x + q c_d)).  A link whose level set does
 *        not change sign, or (q < 1/2) whose second fluid node x - c_d is solid
 *        or outside the domain, takes the SBB rule.  Needs analytic bodies
 *        (lbm_set_bodies / lbm_set_motion), not lbm_set_flags. */
typedef enum { LBM_SOLID_SBB = 0, LBM_SOLID_IBB = 1 } lbm_solid_bc;

typedef struct {
    int32_t nx, ny, nz;
    int32_t collision;    /* lbm_collision */
    int32_t storage;      /* lbm_storage */
    int32_t wall_point;   /* lbm_wall_point */
    int32_t refill;       /* 0: none (reference); 1: fresh fluid nodes get f_eq of the
                             average of their fluid neighbours' rho, u */
    int32_t solid_bc;     /* lbm_solid_bc */
    int32_t q;            /* lattice: 19 (D3Q19) or 27 (D3Q27), src/Utilities.jl:945-993; 0 means 19 */
    int32_t pad_;
    double omega_plus;    /* omega_+ = 1/tau_+ (BGK: omega)        src/Parameters.jl:69-72 */
    double omega_minus;   /* omega_- = 1/tau_- (TRT only)          src/Parameters.jl:70-72 */
    double g[3];          /* body acceleration, Guo forcing with F = rho g (0 = off) */
    double rho0;          /* uniform start density (src/Initialization.jl:41) */
    double u0[3];         /* uniform start velocity */
    lbm_face faces[6];    /* x-, x+, y-, y+, z-, z+ */
} lbm_config;

/* Body shapes.  SPHERE and ELLIPSOID are 3-D; CYLINDER and NACA4 are the
 * reference's planar shapes, extruded along z (src/Geometry.jl:14-39; the
 * level sets of the interpolated bounce-back, phi_cylinder_lat and
 * phi_airfoil_lat, src/Geometry.jl:368-396):
 *   CYLINDER  semi[0] = R; (x - x_c)^2 + (y - y_c)^2 <= R^2 + 1e-6 R^2
 *             (src/Geometry.jl:71-73); orientation ignored;
 *   NACA4     semi[0] = chord; naca = {m, p, t, x_int, x_le_offset} from the
 *             4-digit code (_parse_naca4, src/Geometry.jl:358-366): m = d1/100,
 *             p = d2/10, t = d3d4/100; x_int is the trailing-edge squeeze
 *             (1.0089304 in the reference) and x_le_offset the distance from
 *             the pivot (the body centre) to the leading edge along the chord
 *             (0.25 chord in the reference, src/Geometry.jl:30-35).  The body
 *             frame is R^T (x - centre); R must be a rotation about z by the
 *             pitch angle theta (_rot_to_body, src/Geometry.jl:321-323).
 *             Inside test src/Geometry.jl:75-82, thickness and camber lines
 *             src/Geometry.jl:325-356. */
typedef enum { LBM_SPHERE = 0, LBM_ELLIPSOID = 1, LBM_CYLINDER = 2, LBM_NACA4 = 3 } lbm_shape;

/* One rigid body.  The mask pose (centre, semi-axes, rotation) is evaluated in
 * FP32 as src/Geometry.jl:66-82 does; the half-step pose and kinematics
 * (centre, rotation, linear and angular velocity at t_{n+1/2},
 * src/Motion.jl:62-63,96-116) feed the moving-wall correction, the torque arm
 * and the interpolated bounce-back's wall fraction. */
typedef struct {
    int32_t shape;            /* lbm_shape */
    int32_t pad_;
    float center[3];          /* mask pose at t_n, lattice units */
    float semi[3];            /* sphere, cylinder: semi[0] = R; ellipsoid: a, b, c; NACA4: semi[0] = chord */
    float rot[9];             /* row-major R with x_lattice = R x_body (ellipsoid, NACA4) */
    float naca[5];            /* NACA4: m, p, t, x_int, x_le_offset */
    double center_half[3];    /* centre at t_{n+1/2} */
    double vel[3];            /* U at t_{n+1/2} */
    double omega[3];          /* Omega at t_{n+1/2} (rad/step) */
    double rot_half[9];       /* R at t_{n+1/2} (IBB wall fraction) */
} lbm_body;

/* Prescribed motion law evaluated by lbm_step() itself: centre
 * c(t) = center0 + vel t + amp sin(freq t + phase) and a constant angular
 * velocity about the centre.  The pose for step n (1-based) is taken at
 * t = (n - 1 + offset) with offset 0 for the mask and 1/2 for the kinematics
 * (src/Motion.jl:26-36).  The reference's heaving cylinder
 * (src/MainCore.jl:72-80) is amp = (0, -0.4 R, 0), freq = 2 pi 0.154 U / (2R),
 * phase = 0. */
typedef struct {
    int32_t shape;
    int32_t pad_;
    double center0[3];        /* centre at t = 0 */
    double semi[3];
    double rot0[9];           /* orientation at t = 0, row-major */
    double vel[3];            /* constant linear velocity (cells/step) */
    double omega[3];          /* constant angular velocity (rad/step) */
    double amp[3];            /* oscillation amplitude (cells) */
    double freq;              /* oscillation angular frequency (rad/step) */
    double phase;             /* oscillation phase (rad) */
    double naca[5];           /* NACA4 parameters, as in lbm_body */
} lbm_motion;

typedef struct lbm_ctx lbm_ctx;

/* Replaces worker_alloc + init_bb_buffers! + the LBMParams derivation consumer
 * (src/Initialization.jl:89-130, src/MainCore.jl:181-206). */
int lbm_create(const lbm_config* cfg, lbm_ctx** out);
int lbm_destroy(lbm_ctx* ctx);
/* Copies the last error message (NUL-terminated, truncated to n). */
int lbm_last_error(const lbm_ctx* ctx, char* buf, size_t n);

/* Bodies for the next step: replaces the pose part of motion!
 * (src/Motion.jl:61-100); the mask and link list are rebuilt on the device by
 * the next lbm_step (generate_mask_s_local! / generate_mask_bb_local! /
 * compile_bb_data!, src/Geometry.jl:92-311).  Clears any motion law. */
int lbm_set_bodies(lbm_ctx* ctx, int32_t n, const lbm_body* bodies);
/* Prescribed motion law (replaces MotionLaw/motion_at, src/Motion.jl:14-36). */
int lbm_set_motion(lbm_ctx* ctx, int32_t n, const lbm_motion* laws);
/* Explicit solid flags (0 fluid, k+1 solid of body k); no bodies are kept. */
int lbm_set_flags(lbm_ctx* ctx, const uint8_t* flags);
/* Rebuilds the geometry for the pending bodies now (normally done by lbm_step). */
int lbm_update_geometry(lbm_ctx* ctx);
int lbm_get_flags(lbm_ctx* ctx, uint8_t* flags);
/* Boundary links in canonical order: ascending cell, then ascending direction,
 * as kernel_fill_bb_csr! writes them (src/Geometry.jl:233-261).  A link (cell, d)
 * has a solid node at cell + c_d. */
int lbm_get_links(lbm_ctx* ctx, int32_t* cell, int8_t* dir, int64_t cap, int64_t* nnz);

/* Equilibrium start (initialize_f, src/Initialization.jl:57-83), then the first
 * collision.  rho: N values or NULL (rho0); u: 3N values (ux | uy | uz) or NULL (u0). */
int lbm_init_equilibrium(lbm_ctx* ctx, const double* rho, const double* u);

/* Runs nsteps reference time steps.  force_out (nullable) receives
 * nsteps * nbodies * 6 doubles: per step and body (Fx, Fy, Fz, Tx, Ty, Tz), the
 * momentum-exchange force of compute_aero! (src/MEA.jl:76-116, accumulated in
 * FP64 per body) and its torque about the half-step centre. */
int lbm_step(lbm_ctx* ctx, int32_t nsteps, double* force_out);

/* State of the last completed step (compute_macros!, src/ComputeMacro.jl:11-48):
 * rho (N) and u (3N, ux | uy | uz).  Either pointer may be NULL. */
int lbm_get_macros(lbm_ctx* ctx, double* rho, double* u);
/* Post-stream, post-boundary populations f_old of the last completed step,
 * 19*N doubles, SoA. */
int lbm_get_populations(lbm_ctx* ctx, double* f);
/* Sets f_old (19*N doubles, SoA) as the state after step 0 and collides it. */
int lbm_set_populations(lbm_ctx* ctx, const double* f);
/* FP64 sums over fluid cells of rho and rho*u for the last completed step. */
int lbm_conservation(lbm_ctx* ctx, double* mass, double* mom3);
/* lbm_conservation without any allocation (no counterpart in the reference):
 * for callers that cannot allocate device memory (HIP graph capture, serving
 * loops, harnesses that forbid hipMalloc).  The same sums, formed from the
 * stored post-collision state (collision conserves rho and adds exactly rho g
 * to rho u), so with FP64/FP32 storage they equal lbm_conservation's to
 * rounding and with 16-bit storage they differ by that storage's rounding of
 * f*.  Its partial sums live in a library-owned device array of
 * 1024 x 4 + 4 doubles per GPU, which bounds no lattice size; calls that
 * overlap on one GPU share that array and race.  Call from the context's
 * device; synchronises the context stream. */
int lbm_conservation_static(lbm_ctx* ctx, double* mass, double* mom3);

/* Introspection: N, number of boundary (link-carrying) nodes, number of links,
 * completed steps, device bytes held. */
int lbm_info(lbm_ctx* ctx, int64_t* ncells, int64_t* nbb_nodes, int64_t* nlinks,
             int64_t* steps, int64_t* device_bytes);
/* Per-kernel HIP-event timing on the context stream: an event pair around every
 * enable-th step kernel launch (0 = off, 1 = every launch), at most 4096 pairs
 * per lbm_step call. */
int lbm_profile(lbm_ctx* ctx, int32_t enable);
/* Totals since the last enable: summed duration (ms) and launch count of the
 * fused stream-collide kernel and of the boundary-link kernel. */
int lbm_profile_read(lbm_ctx* ctx, double* bulk_ms, int64_t* bulk_launches,
                     double* link_ms, int64_t* link_launches);
/* Waits for all work queued on the context stream. */
int lbm_sync(lbm_ctx* ctx);

#ifdef __cplusplus
}
#endif

#endif /* LBM_HIP_H */
