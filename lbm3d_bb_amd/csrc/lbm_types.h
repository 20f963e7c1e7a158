// lbm_types.h -- data shared by the host code and the gfx950 kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/lbm_hip.h"

namespace lbm {

constexpr int kBlock = 256;
constexpr int kWaves = kBlock / 64;
constexpr int kMaxBoxes = 12 + LBM_MAX_BODIES;  // domain edges, bodies
constexpr int kScanBoxes = 32;                   // boxes a fix-up block scans with independent loads
// Bodies the fix-up blocks stage in LDS (new and old pose) so the analytic
// solidity tests and the link kinematics read LDS instead of global memory.
constexpr int kLdsBodies = 16;
// Interpolated bounce-back: the wall fraction q is clamped to [kEpsQ, 1 - kEpsQ]
// (EPS_Q of src/SolidBC.jl:321, which the reference leaves undefined).
constexpr float kEpsQ = 1e-3f;

// Flag byte: low 7 bits 0 (fluid) or k+1 (solid of body k); bit 7 marks a fluid
// boundary node of static geometry, finished by the fix-up blocks.
constexpr uint8_t kFlagBody = 0x7f;
constexpr uint8_t kFlagFixup = 0x80;

// Body as the kernels see it: FP32 mask pose (inside test) and FP64 half-step
// pose and kinematics (moving-wall velocity, torque arm, wall fraction).
struct DevBody {
    int shape;
    float c[3];
    float r2;      // sphere, cylinder: R * R
    float ia2[3];  // ellipsoid: 1 / a^2, 1 / b^2, 1 / c^2
    float rot[9];
    float reach;   // 0: analytic shape; < 0: unknown (explicit flags)
    float semi[3];
    float naca[5];  // NACA4: m, p, t, x_int, x_le_offset
    float rh[9];   // rotation at the half step
    float ext[3];  // per axis: beyond this offset of the cell centre no neighbour is inside
    double ch[3];
    double v[3];
    double om[3];
};

// A box of cells: lo may lie outside [0, L) on a periodic axis (wrapped).
// kind 1 = body region, 2 = every cell on the general path (domain edges and
// the x faces with a velocity or anti-bounce-back rule).
struct Box {
    int lo[3];
    int ext[3];
    int kind;
    int first_block;
    int face;  // face slab: 2 * axis + side (0..5); body region: -1
    int ovl;   // 1: may share cells with an earlier fix-up box (cells there belong to the first)
};

// Scale of the 16-bit population formats: they store (f_d - w_d) k, the
// deviation from the rest state w_d (rho0 = 1) times a power of two k that
// keeps small deviations out of FP16's subnormal range (exact scaling, one
// rounding as before).
struct Shift {
    float k;
    float ik;       // 1 / k
    uint32_t seed;  // per step: random stream of the stochastic rounding of stores
};

struct Grid {
    int nx, ny, nz;
    int per;  // bit a: axis a periodic
    int64_t N;
    int64_t P;  // population plane stride (elements): N plus a pad that staggers the planes' channel mapping
    int64_t nxy;
    int face_kind[6];
    const double* face_u;  // device [6][3]: wall / inlet velocity of each face
    int wall_point;
    int refill;
    double rho0;
    // Face moments: for every cell on a velocity or anti-bounce-back face, the
    // (rho, u) its rule reads next step, four compute-type values per slot,
    // written when the cell's f* is produced.  face_base[f] is the first slot
    // of face f (-1: periodic or zero-gradient face, no slots).
    int face_base[6];
    Shift sh;        // 16-bit storage reference state
    void* side_in;   // per launch: moments of the face cells of fin's f*
    void* side_out;  // per launch: the same for fout
};

template <typename C>
struct Phys {
    C wp, wm;
    C g[3];
};

// What the bulk rows need of a step's geometry, in one compact block of the
// step record: the mode, the body count and up to kBulkBoxes body regions
// (read with one batch of scalar loads instead of a walk over the boxes).
constexpr int kBulkBoxes = 8;
struct BulkGeom {
    int mode;     // 0: static (fix-up bit in the flags), 1: moving (analytic regions)
    int nbodies;
    int nbox;     // body regions listed below; -1: more than kBulkBoxes (read the step record)
    int lo[3][kBulkBoxes];   // lo in [0, L) on periodic axes
    int ext[3][kBulkBoxes];
};

// Geometry of one step for the fix-up blocks of k_step: domain edges and
// general-path x faces (kind 2) first, then body regions (kind 1); a cell in
// several boxes belongs to the first.
struct StepGeom {
    int nbox;
    int nface_blocks;  // unused (0)
    int nfix_blocks;   // fix-up blocks over the boxes
    int first_fix_box; // index of the first box the fix-up blocks walk
    int nbodies;
    int mode;          // 0: solidity from flags (static), 1: analytic (moving bodies)
    int list_blocks;   // k_fixup blocks over the static boundary-node list (mode 0)
    int pad;
    BulkGeom bg;
    Box box[kMaxBoxes];
};

enum { MODE_STEP = 0, MODE_GET = 1, MODE_COLLIDE_ONLY = 2, MODE_GET_RAW = 3 };

// One kernel launch request; the per-storage translation units implement
// launch<S>() for it, so the host code never instantiates kernels itself.
enum { K_BULK = 0, K_STEP = 1, K_EQUILIBRIUM = 2, K_FROM_DOUBLE = 3 };

struct LaunchReq {
    int kind;
    int mode;
    int collision;
    int solid_bc;  // lbm_solid_bc: picks the kernels with the interpolated rule
    int forced;
    Grid g;
    Phys<double> p;
    hipStream_t stream;
    const void* fin;
    void* fout;
    double* gout;
    double* rho_out;
    double* u_out;
    // fix-up
    int grid_blocks;
    const char* rec;
    int nb;
    uint8_t* flags;
    const int32_t* list;
    const int32_t* list_count;
    double* partials;
    int write_flags;
    // start state
    const double* src_rho;
    const double* src_u;
    const double* src_f;
    double rho0;
    double u0[3];
};

struct StoreF64;
struct StoreF32;
struct StoreF16S;
struct StoreBF16S;

template <class S>
int launch(const LaunchReq& r);

// storage-independent kernels (lbm_geom.hip)
int launch_flags_full(uint8_t* flags, const DevBody* bodies, int nb, const Grid& g, hipStream_t st);
int launch_bb_count(const uint8_t* flags, const Grid& g, int32_t* node_counts, int32_t* link_counts, hipStream_t st);
int launch_scan2(int32_t* a, int32_t* b, int n, hipStream_t st);
int launch_bb_fill(const uint8_t* flags, const Grid& g, const int32_t* node_off, const int32_t* link_off,
                   int32_t* nodes, int32_t* link_cell, int8_t* link_dir, hipStream_t st);
int launch_force_finalize(const double* partials, int64_t step_stride, const char* recs, size_t rec_bytes,
                          int nsteps, int n6, double* out, hipStream_t st);
int launch_mark_list(uint8_t* flags, const int32_t* list, int64_t n, int set, int64_t N, hipStream_t st);
int launch_conservation(const double* rho, const double* u, const uint8_t* flags, int64_t N, double* partials,
                        hipStream_t st);

inline int64_t nblocks_of(int64_t n) { return (n + kBlock - 1) / kBlock; }

}  // namespace lbm
