// lbm_capi.hip -- host side of liblbm_hip.so: the C ABI of include/lbm_hip.h.
//
// Replaces the per-worker state and time loop of the reference
// (src/MainCore.jl:165-379): device buffers (worker_alloc, init_bb_buffers!,
// src/Initialization.jl:89-130), the motion/mask/link rebuild (motion!,
// src/Motion.jl:41-122) and the per-step operator sequence.
#include <hip/hip_runtime.h>

#include <hip/hip_fp16.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "lbm_lattice.h"
#include "lbm_q27.h"
#include "lbm_types.h"

using namespace lbm;

namespace {

constexpr size_t kMaxProfileEvents = 8192;  // event pairs held per lbm_step call (4096 samples)
constexpr int kChunk = 64;          // steps whose poses are uploaded with one copy
// Bytes in front of the first population plane: the bulk pass's shifted pair
// loads read one element before a row start (the value is replaced for x = 0).
constexpr size_t kFrontPad = 256;

}  // namespace

struct lbm_ctx {
    lbm_config cfg{};
    Grid g{};
    Phys<double> pd{};
    Phys<float> pf{};
    bool forced = false;
    hipStream_t stream = nullptr;   // k_bulk, copies, getters
    int device = 0;

    void* fbuf[2] = {nullptr, nullptr};      // population planes (fbuf_raw + kFrontPad)
    void* fbuf_raw[2] = {nullptr, nullptr};  // allocations
    void* side[2] = {nullptr, nullptr};      // face moments of fbuf[i]'s f* (Grid::face_base slots)
    double* d_face_u = nullptr;              // [6][3] face velocities
    size_t side_bytes = 0;
    int cur = 0;               // buffer holding f* (pending stream)
    size_t fbytes = 0;         // q planes of P elements
    uint8_t* flags = nullptr;
    int q = 19;                // lattice: D3Q19 or D3Q27

    // D3Q27: flags before the last pose change (the two buffers swap roles at
    // every pose change), and the last step's bodies
    // (inside the device step records) for the getters' replay
    uint8_t* flags_old = nullptr;
    std::vector<DevBody> q27_pose_cur, q27_pose_old;  // the pose whose flags each buffer holds
    std::vector<std::array<float, 3>> q27_reach_cur, q27_reach_old;  // and the bodies' reach then
    bool q27_cur_ok = false, q27_old_ok = false;
    const DevBody* q27_bnew = nullptr;
    const DevBody* q27_bold = nullptr;
    int q27_refill = 0;

    // static boundary-node list (solidity from flags)
    int32_t* list = nullptr;
    int32_t* list_count = nullptr;  // device scalar
    int64_t list_n = 0;
    int64_t list_cap = 0;
    int64_t nlinks = 0;
    bool list_valid = false;
    int32_t* cnt_a = nullptr;
    int32_t* cnt_b = nullptr;
    int64_t cnt_cap = 0;

    // bodies
    int nb = 0;
    bool have_laws = false;
    std::vector<lbm_motion> laws;
    std::vector<lbm_body> pending;
    bool pending_set = false;
    bool flags_only = false;      // geometry from lbm_set_flags (no analytic bodies)
    std::vector<DevBody> mask_bodies;  // pose that flags reflect
    std::vector<std::array<float, 3>> mask_reach;  // the bodies' reach in that pose
    std::vector<DevBody> step_bodies;  // pose + kinematics of the current step
    std::vector<std::array<float, 3>> reach;  // bounding half-extents per body and axis
    bool have_flags_box = false;       // lbm_set_flags: bounding box of the solid cells
    float flags_lo[3] = {0, 0, 0}, flags_hi[3] = {0, 0, 0};
    bool geom_valid = false;

    // per-step records: [kChunk] x (StepGeom | DevBody new[nb] | DevBody old[nb])
    size_t rec_bytes = 0;
    char* d_rec = nullptr;
    char* h_rec[2] = {nullptr, nullptr};
    hipEvent_t rec_ev[2] = {nullptr, nullptr};
    bool rec_ev_used[2] = {false, false};
    int parity = 0;
    int nrec_cap_bodies = -1;

    double* d_partials = nullptr;
    int64_t partials_cap = 0;  // doubles
    double* d_forces = nullptr;  // [kChunk][nb][6]
    double* h_forces[2] = {nullptr, nullptr};

    // last step (for getters)
    const char* last_rec = nullptr;
    int last_grid = 0;
    int last_face_grid = 0;
    bool last_rec_valid = false;

    int64_t step = 0;
    bool initialised = false;
    bool raw_state = false;  // step 0: fbuf[1-cur] holds f_old_0
    // D3Q27 16-bit storage, even nx >= 4: the split step (k27_split: pair and per-cell blocks);
    // LBM_Q27_SPLIT=0 in the environment keeps the single per-cell launch
    bool q27_split = false;

    int profile_stride = 0;  // lbm_profile: time every profile_stride-th step (0 = off)
    std::vector<hipEvent_t> ev_pool;
    size_t ev_used = 0;
    double bulk_ms = 0.0, link_ms = 0.0;
    int64_t bulk_launches = 0, link_launches = 0;

    int64_t device_bytes = 0;
    std::string err;
};

namespace {

int fail(lbm_ctx* c, int code, const std::string& msg) {
    if (c) c->err = msg;
    return code;
}

#define HIPCHK(ctx, call)                                                                           \
    do {                                                                                            \
        hipError_t e_ = (call);                                                                     \
        if (e_ != hipSuccess)                                                                       \
            return fail((ctx), LBM_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_));   \
    } while (0)

// For calls made while temporaries are held: a failure is recorded in rc (the
// first one wins) and the caller goes on to free what it allocated.
#define HIPKEEP(ctx, rc, call)                                                                      \
    do {                                                                                            \
        hipError_t e_ = (call);                                                                     \
        if (e_ != hipSuccess && (rc) == LBM_OK)                                                     \
            (rc) = fail((ctx), LBM_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_));   \
    } while (0)

int dev_alloc(lbm_ctx* c, void** p, size_t bytes) {
    hipError_t e = hipMalloc(p, bytes ? bytes : 16);
    if (e != hipSuccess) return fail(c, LBM_ERR_HIP, std::string("hipMalloc: ") + hipGetErrorString(e));
    c->device_bytes += (int64_t)bytes;
    return LBM_OK;
}

void dev_free(lbm_ctx* c, void* p, size_t bytes) {
    if (p) {
        (void)hipFree(p);
        c->device_bytes -= (int64_t)bytes;
    }
}

// Host pose evaluation with no contraction, so the FP32 mask pose is the same
// bit pattern a plain C evaluation of the formula produces.
void rotation_at(const double om[3], double t, const double R0[9], double out[9]) {
#pragma clang fp contract(off)
    const double w = std::sqrt(om[0] * om[0] + om[1] * om[1] + om[2] * om[2]);
    double Rr[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
    if (w > 0.0) {
        const double k[3] = {om[0] / w, om[1] / w, om[2] / w};
        const double th = w * t, s = std::sin(th), c1 = 1.0 - std::cos(th);
        const double K[9] = {0, -k[2], k[1], k[2], 0, -k[0], -k[1], k[0], 0};
        double K2[9];
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) {
                double a = 0.0;
                for (int m = 0; m < 3; ++m) a += K[i * 3 + m] * K[m * 3 + j];
                K2[i * 3 + j] = a;
            }
        for (int i = 0; i < 9; ++i) Rr[i] += s * K[i] + c1 * K2[i];
    }
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            double a = 0.0;
            for (int m = 0; m < 3; ++m) a += Rr[i * 3 + m] * R0[m * 3 + j];
            out[i * 3 + j] = a;
        }
}

// motion_at (src/Motion.jl:26-36) for step n with dt = 1: mask pose at n-1,
// kinematics at n-1/2.
// The centre is c0 + v t + A sin(w t + phase), its velocity v + A w cos(w t + phase).
lbm_body body_from_law(const lbm_motion& m, int64_t n) {
#pragma clang fp contract(off)
    lbm_body b;
    std::memset(&b, 0, sizeof(b));
    b.shape = m.shape;
    const double t0 = (double)(n - 1), th = (double)(n - 1) + 0.5;
    double R[9], Rh[9];
    rotation_at(m.omega, t0, m.rot0, R);
    rotation_at(m.omega, th, m.rot0, Rh);
    const double s0 = std::sin(m.freq * t0 + m.phase), sh = std::sin(m.freq * th + m.phase);
    const double ch = std::cos(m.freq * th + m.phase);
    for (int k = 0; k < 3; ++k) {
        b.center[k] = (float)(m.center0[k] + m.vel[k] * t0 + m.amp[k] * s0);
        b.semi[k] = (float)m.semi[k];
        b.center_half[k] = m.center0[k] + m.vel[k] * th + m.amp[k] * sh;
        b.vel[k] = m.vel[k] + m.amp[k] * m.freq * ch;
        b.omega[k] = m.omega[k];
    }
    for (int k = 0; k < 9; ++k) {
        b.rot[k] = (float)R[k];
        b.rot_half[k] = Rh[k];
    }
    for (int k = 0; k < 5; ++k) b.naca[k] = (float)m.naca[k];
    return b;
}

std::array<float, 3> reach_of(int shape, const double semi[3], const double naca[5], int nz);

DevBody to_dev(const lbm_body& b, int nz) {
#pragma clang fp contract(off)
    DevBody d;
    std::memset(&d, 0, sizeof(d));
    d.shape = b.shape;
    for (int k = 0; k < 3; ++k) {
        d.c[k] = b.center[k];
        d.ch[k] = b.center_half[k];
        d.v[k] = b.vel[k];
        d.om[k] = b.omega[k];
    }
    for (int k = 0; k < 9; ++k) {
        d.rot[k] = b.rot[k];
        d.rh[k] = (float)b.rot_half[k];
    }
    for (int k = 0; k < 3; ++k) d.semi[k] = b.semi[k];
    for (int k = 0; k < 5; ++k) d.naca[k] = b.naca[k];
    d.reach = 0.0f;
    {
        // bounding half-extents, widened by the neighbour offset and rounding
        const double semi[3] = {b.semi[0], b.semi[1], b.semi[2]};
        double naca[5];
        for (int k = 0; k < 5; ++k) naca[k] = b.naca[k];
        const std::array<float, 3> r = reach_of(b.shape, semi, naca, nz);
        for (int k = 0; k < 3; ++k) d.ext[k] = r[k] * 1.0001f + 1.01f;
        if (b.shape == LBM_CYLINDER || b.shape == LBM_NACA4) d.ext[2] = 3.0e38f;
    }
    if (b.shape == LBM_SPHERE || b.shape == LBM_CYLINDER) {
        const float R = b.semi[0];
        d.r2 = R * R;
    } else if (b.shape == LBM_ELLIPSOID) {
        for (int k = 0; k < 3; ++k) {
            const float a = b.semi[k];
            d.ia2[k] = 1.0f / (a * a);
        }
    }
    return d;
}

// Bounding half-extents of a body about its centre, per axis.  The planar
// shapes span the whole z extent (nz + 4 covers every z offset).
std::array<float, 3> reach_of(int shape, const double semi[3], const double naca[5], int nz) {
    if (shape == LBM_SPHERE) return {(float)semi[0], (float)semi[0], (float)semi[0]};
    if (shape == LBM_ELLIPSOID) {
        const float m = (float)std::max(semi[0], std::max(semi[1], semi[2]));
        return {m, m, m};
    }
    const float zr = (float)nz + 4.0f;
    if (shape == LBM_CYLINDER) return {(float)semi[0], (float)semi[0], zr};
    // airfoil: farthest chord end from the pivot plus the thickness and camber
    const double chord = semi[0], xle = naca[4];
    const float r = (float)(std::max(std::fabs(xle), std::fabs(chord - xle)) + std::fabs(naca[2]) * chord +
                            std::fabs(naca[0]) + 1.0);
    return {r, r, zr};
}

bool same_mask_pose(const DevBody& a, const DevBody& b) {
    if (a.shape != b.shape || std::memcmp(a.c, b.c, sizeof(a.c)) != 0) return false;
    // sphere and cylinder masks ignore the orientation
    if (a.shape == LBM_SPHERE || a.shape == LBM_CYLINDER) return a.r2 == b.r2;
    if (a.shape == LBM_NACA4)
        return std::memcmp(a.semi, b.semi, sizeof(a.semi)) == 0 && std::memcmp(a.naca, b.naca, sizeof(a.naca)) == 0 &&
               std::memcmp(a.rot, b.rot, sizeof(a.rot)) == 0;
    return std::memcmp(a.ia2, b.ia2, sizeof(a.ia2)) == 0 && std::memcmp(a.rot, b.rot, sizeof(a.rot)) == 0;
}

// Shape checks shared by lbm_set_bodies and lbm_set_motion: known shape,
// positive size, and for the planar shapes a rotation about z only.
const char* shape_error(int shape, const double semi[3], const double rot[9], const double omega[3]) {
    if (shape < LBM_SPHERE || shape > LBM_NACA4) return "unknown body shape";
    const int ns = shape == LBM_ELLIPSOID ? 3 : 1;
    for (int k = 0; k < ns; ++k)
        if (!(semi[k] > 0.0)) return "body size must be positive";
    if (shape == LBM_NACA4) {
        const double off[4] = {rot[2], rot[5], rot[6], rot[7]};
        for (double v : off)
            if (std::fabs(v) > 1e-6) return "a NACA4 body rotates about z only";
        if (std::fabs(rot[8] - 1.0) > 1e-6) return "a NACA4 body rotates about z only";
    }
    if ((shape == LBM_NACA4 || shape == LBM_CYLINDER) && omega && (omega[0] != 0.0 || omega[1] != 0.0))
        return "a planar body spins about z only";
    return nullptr;
}


void index_box(const lbm_ctx* c, float lo_f, float hi_f, int axis, int& lo, int& ext) {
    const int L = axis == 0 ? c->g.nx : (axis == 1 ? c->g.ny : c->g.nz);
    int l = (int)std::floor(lo_f) - 2, h = (int)std::ceil(hi_f) + 1;
    if ((c->g.per >> axis) & 1) {
        if (h - l + 1 >= L) {
            l = 0;
            h = L - 1;
        } else {
            // bring lo into [0, L): the kernel wraps lo + i < 2L with one subtraction
            const int sh = ((l % L) + L) % L - l;
            l += sh;
            h += sh;
        }
    } else {
        l = std::max(l, 0);
        h = std::min(h, L - 1);
    }
    lo = l;
    ext = std::max(h - l + 1, 0);
}

// Domain edges: the lines where two non-periodic faces meet.  With three
// non-periodic axes the corner cells lie on three lines: the lines of the
// first axis pair keep them and the later lines stop one cell short, so the
// edge boxes share no cell.
void add_edge_boxes(const lbm_ctx* c, StepGeom& sg) {
    const int L[3] = {c->g.nx, c->g.ny, c->g.nz};
    int pairs = 0;
    for (int a = 0; a < 3; ++a)
        for (int b = a + 1; b < 3; ++b) {
            if (((c->g.per >> a) & 1) || ((c->g.per >> b) & 1)) continue;
            const int o = 3 - a - b;
            const bool trim = pairs++ > 0 && !((c->g.per >> o) & 1);
            for (int sa = 0; sa < 2; ++sa)
                for (int sb = 0; sb < 2; ++sb) {
                    Box e{};
                    e.lo[a] = sa ? L[a] - 1 : 0;
                    e.ext[a] = 1;
                    e.lo[b] = sb ? L[b] - 1 : 0;
                    e.ext[b] = 1;
                    e.lo[o] = trim ? 1 : 0;
                    e.ext[o] = trim ? L[o] - 2 : L[o];
                    e.kind = 2;
                    e.face = -1;
                    sg.box[sg.nbox++] = e;
                }
        }
}

// Whether two boxes share a cell (wrapping on periodic axes).
bool boxes_meet(const lbm_ctx* c, const Box& a, const Box& b) {
    const int L[3] = {c->g.nx, c->g.ny, c->g.nz};
    for (int ax = 0; ax < 3; ++ax) {
        if (a.ext[ax] <= 0 || b.ext[ax] <= 0) return false;
        if ((c->g.per >> ax) & 1) {
            const int r = (((b.lo[ax] - a.lo[ax]) % L[ax]) + L[ax]) % L[ax];
            if (!(r < a.ext[ax] || L[ax] - r < b.ext[ax])) return false;
        } else if (!(a.lo[ax] < b.lo[ax] + b.ext[ax] && b.lo[ax] < a.lo[ax] + a.ext[ax])) {
            return false;
        }
    }
    return true;
}

// Block numbering: face slabs (k_faces) and edge/body boxes (k_fixup) each from 0.
void finish_boxes(const lbm_ctx* c, StepGeom& sg) {
    int fblocks = 0, xblocks = 0;
    sg.first_fix_box = sg.nbox;
    for (int i = 0; i < sg.nbox; ++i) {
        const int64_t vol = (int64_t)sg.box[i].ext[0] * sg.box[i].ext[1] * sg.box[i].ext[2];
        if (sg.box[i].kind == 0) {
            sg.box[i].first_block = fblocks;
            fblocks += (int)nblocks_of(vol);
        } else {
            if (sg.first_fix_box == sg.nbox) sg.first_fix_box = i;
            sg.box[i].first_block = xblocks;
            xblocks += (int)nblocks_of(vol);
        }
    }
    sg.nface_blocks = fblocks;
    sg.nfix_blocks = xblocks;
    for (int i = sg.first_fix_box; i < sg.nbox; ++i) {
        sg.box[i].ovl = 0;
        for (int m = sg.first_fix_box; m < i && !sg.box[i].ovl; ++m)
            if (boxes_meet(c, sg.box[m], sg.box[i])) sg.box[i].ovl = 1;
    }
}

LaunchReq base_req(const lbm_ctx* c, int kind, int mode) {
    LaunchReq r;
    std::memset(&r, 0, sizeof(r));
    r.kind = kind;
    r.mode = mode;
    r.collision = c->cfg.collision;
    r.solid_bc = c->cfg.solid_bc;
    r.forced = c->forced ? 1 : 0;
    r.g = c->g;
    r.p = c->pd;
    r.stream = c->stream;
    r.nb = c->nb;
    r.flags = c->flags;
    r.list = c->list;
    r.list_count = c->list_count;
    r.partials = c->d_partials;
    return r;
}

// the face-moment buffer that belongs to population buffer p
void* side_of(const lbm_ctx* c, const void* p) {
    if (p == c->fbuf[0]) return c->side[0];
    if (p == c->fbuf[1]) return c->side[1];
    return nullptr;
}

int do_launch(lbm_ctx* c, const LaunchReq& r0) {
    LaunchReq r = r0;
    r.g.side_in = side_of(c, r.fin);
    r.g.side_out = side_of(c, r.fout);
    int e = 0;
    switch (c->cfg.storage) {
        case LBM_FP64: e = launch<StoreF64>(r); break;
        case LBM_FP32: e = launch<StoreF32>(r); break;
        case LBM_FP16S: e = launch<StoreF16S>(r); break;
        case LBM_BF16S: e = launch<StoreBF16S>(r); break;
        default: return fail(c, LBM_ERR_ARG, "unknown storage");
    }
    if (e) return fail(c, LBM_ERR_HIP, std::string("kernel launch: ") + hipGetErrorString((hipError_t)e));
    return LBM_OK;
}

int launch_bulk(lbm_ctx* c, int mode, const void* fin, void* fout, double* gout, double* rho, double* u,
                const char* rec) {
    LaunchReq r = base_req(c, K_BULK, mode);
    r.rec = rec;
    r.fin = fin;
    r.fout = fout;
    r.gout = gout;
    r.rho_out = rho;
    r.u_out = u;
    return do_launch(c, r);
}

// One time step (or its pull-only replay for the getters): nfix fix-up blocks
// followed by the bulk blocks, in a single launch.
// The bulk rows' view of a step record.
BulkGeom bulk_geom_of(const StepGeom& sg) {
    BulkGeom bg;
    std::memset(&bg, 0, sizeof(bg));
    bg.mode = sg.mode;
    bg.nbodies = sg.nbodies;
    int n = 0;
    for (int i = 0; i < sg.nbox; ++i) {
        if (sg.box[i].kind != 1) continue;
        if (n == kBulkBoxes) {
            n = -1;
            break;
        }
        for (int a = 0; a < 3; ++a) {
            bg.lo[a][n] = sg.box[i].lo[a];
            bg.ext[a][n] = sg.box[i].ext[a];
        }
        ++n;
    }
    bg.nbox = n;
    return bg;
}

int launch_step(lbm_ctx* c, int mode, int nfix, const void* fin, void* fout, double* gout, double* rho, double* u,
                const char* rec, int write_flags, double* partials) {
    LaunchReq r = base_req(c, K_STEP, mode);
    r.fin = fin;
    r.fout = fout;
    r.gout = gout;
    r.rho_out = rho;
    r.u_out = u;
    r.grid_blocks = nfix;
    r.rec = rec;
    r.write_flags = write_flags;
    r.partials = partials;
    return do_launch(c, r);
}

#define GEOCHK(ctx, call)                                                                                   \
    do {                                                                                                    \
        int e_ = (call);                                                                                    \
        if (e_) return fail((ctx), LBM_ERR_HIP, std::string(#call) + ": " + hipGetErrorString((hipError_t)e_)); \
    } while (0)

int ensure_records(lbm_ctx* c) {
    if (c->nrec_cap_bodies == c->nb && c->d_rec) return LBM_OK;
    if (c->d_rec) {
        HIPCHK(c, hipStreamSynchronize(c->stream));
        dev_free(c, c->d_rec, (size_t)kChunk * c->rec_bytes);
        for (int i = 0; i < 2; ++i) {
            (void)hipHostFree(c->h_rec[i]);
            (void)hipHostFree(c->h_forces[i]);
            c->h_rec[i] = nullptr;
            c->h_forces[i] = nullptr;
        }
        dev_free(c, c->d_forces, 0);
        c->d_forces = nullptr;
    }
    c->rec_bytes = (sizeof(StepGeom) + 2 * (size_t)std::max(c->nb, 1) * sizeof(DevBody) + 255) / 256 * 256;
    int rc = dev_alloc(c, (void**)&c->d_rec, (size_t)kChunk * c->rec_bytes);
    if (rc) return rc;
    const size_t fbytes = (size_t)kChunk * std::max(c->nb, 1) * 6 * sizeof(double);
    rc = dev_alloc(c, (void**)&c->d_forces, fbytes);
    if (rc) return rc;
    for (int i = 0; i < 2; ++i) {
        HIPCHK(c, hipHostMalloc((void**)&c->h_rec[i], (size_t)kChunk * c->rec_bytes, hipHostMallocDefault));
        HIPCHK(c, hipHostMalloc((void**)&c->h_forces[i], fbytes, hipHostMallocDefault));
        std::memset(c->h_rec[i], 0, (size_t)kChunk * c->rec_bytes);
    }
    c->nrec_cap_bodies = c->nb;
    c->last_rec_valid = false;
    return LBM_OK;
}

int ensure_partials(lbm_ctx* c, int64_t blocks) {
    const int64_t need = blocks * std::max(c->nb, 1) * 6;  // blocks = slots over all steps of a chunk
    if (need <= c->partials_cap) return LBM_OK;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    dev_free(c, c->d_partials, (size_t)c->partials_cap * sizeof(double));
    c->d_partials = nullptr;
    const int64_t cap = std::max(need, c->partials_cap * 2);
    int rc = dev_alloc(c, (void**)&c->d_partials, (size_t)cap * sizeof(double));
    if (rc) return rc;
    c->partials_cap = cap;
    return LBM_OK;
}

// Ordered compaction of the boundary nodes (and optionally the links) from flags.
// Per-block count buffers of at least n entries.
int ensure_counts(lbm_ctx* c, int64_t n) {
    if (n <= c->cnt_cap) return LBM_OK;
    dev_free(c, c->cnt_a, (size_t)c->cnt_cap * 4);
    dev_free(c, c->cnt_b, (size_t)c->cnt_cap * 4);
    c->cnt_a = c->cnt_b = nullptr;
    c->cnt_cap = 0;
    int rc = dev_alloc(c, (void**)&c->cnt_a, (size_t)n * 4);
    if (rc) return rc;
    rc = dev_alloc(c, (void**)&c->cnt_b, (size_t)n * 4);
    if (rc) return rc;
    c->cnt_cap = n;
    return LBM_OK;
}

int build_list(lbm_ctx* c, int32_t* d_link_cell, int8_t* d_link_dir, int64_t* nlinks_out) {
    const int64_t nblk = nblocks_of(c->g.N);
    int rc0 = ensure_counts(c, nblk + 1);
    if (rc0) return rc0;
    GEOCHK(c, launch_bb_count(c->flags, c->g, c->cnt_a, c->cnt_b, c->stream));
    GEOCHK(c, launch_scan2(c->cnt_a, c->cnt_b, (int)nblk, c->stream));
    int32_t tot[2];
    HIPCHK(c, hipMemcpyAsync(&tot[0], c->cnt_a + nblk, 4, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipMemcpyAsync(&tot[1], c->cnt_b + nblk, 4, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    const int64_t nn = tot[0];
    if (nn > c->list_cap) {
        dev_free(c, c->list, (size_t)c->list_cap * 4);
        c->list = nullptr;
        c->list_cap = std::max<int64_t>(nn, 1024);
        int rc = dev_alloc(c, (void**)&c->list, (size_t)c->list_cap * 4);
        if (rc) return rc;
    }
    if (d_link_cell && *nlinks_out < tot[1]) {
        *nlinks_out = tot[1];
        return LBM_ERR_CAPACITY;
    }
    GEOCHK(c, launch_bb_fill(c->flags, c->g, c->cnt_a, c->cnt_b, c->list, d_link_cell, d_link_dir, c->stream));
    HIPCHK(c, hipMemcpyAsync(c->list_count, c->cnt_a + nblk, 4, hipMemcpyDeviceToDevice, c->stream));
    // the listed nodes carry the fix-up bit (k_step's bulk blocks skip them)
    GEOCHK(c, launch_mark_list(c->flags, c->list, nn, 0, c->g.N, c->stream));
    GEOCHK(c, launch_mark_list(c->flags, c->list, nn, 1, c->g.N, c->stream));
    HIPCHK(c, hipGetLastError());
    c->list_n = nn;
    c->nlinks = tot[1];
    if (nlinks_out) *nlinks_out = tot[1];
    return LBM_OK;
}

// Full build of flags (from the analytic bodies) and of the static node list.
int full_geometry(lbm_ctx* c, const std::vector<DevBody>& bodies) {
    DevBody* d_b = nullptr;
    int rc = dev_alloc(c, (void**)&d_b, std::max<size_t>(bodies.size(), 1) * sizeof(DevBody));
    if (rc) return rc;
    if (!bodies.empty())
        HIPKEEP(c, rc, hipMemcpyAsync(d_b, bodies.data(), bodies.size() * sizeof(DevBody), hipMemcpyHostToDevice,
                                      c->stream));
    if (rc == LBM_OK) HIPKEEP(c, rc, (hipError_t)launch_flags_full(c->flags, d_b, (int)bodies.size(), c->g, c->stream));
    HIPKEEP(c, rc, hipGetLastError());
    HIPKEEP(c, rc, hipStreamSynchronize(c->stream));
    dev_free(c, d_b, std::max<size_t>(bodies.size(), 1) * sizeof(DevBody));
    if (rc) return rc;
    c->mask_bodies = bodies;
    c->mask_reach = c->reach;
    c->geom_valid = true;
    if (c->q == 27) {  // the D3Q27 step finds its links from the flags
        c->q27_pose_cur = bodies;
        c->q27_reach_cur = c->reach;
        c->q27_cur_ok = true;
        return LBM_OK;
    }
    rc = build_list(c, nullptr, nullptr, nullptr);
    if (rc) return rc;
    c->list_valid = true;
    return LBM_OK;
}

// Bodies for step n (1-based) from the law or the pending set.
void bodies_for_step(lbm_ctx* c, int64_t n, std::vector<DevBody>& out) {
    out.clear();
    if (c->have_laws) {
        for (const auto& m : c->laws) out.push_back(to_dev(body_from_law(m, n), c->g.nz));
    } else if (c->pending_set) {
        for (const auto& b : c->pending) out.push_back(to_dev(b, c->g.nz));
    } else {
        out = c->step_bodies;
    }
}

constexpr int kStopChunk = 1;

// Fills the record of one step.  A step that needs a synchronous rebuild (full
// flags, or the static list after a motion phase) returns kStopChunk unless
// allow_sync, so that the rebuild sees the flags of every earlier step.
// Returns the fix-up grid size via *grid.
// faces_only: keep the box only if it reaches a non-periodic face (static
// bodies: the fix-up blocks walk a region only for its face cells).
void add_body_box(const lbm_ctx* c, StepGeom& sg, const float lo_f[3], const float hi_f[3], bool faces_only = false) {
    Box b{};
    for (int a = 0; a < 3; ++a) index_box(c, lo_f[a], hi_f[a], a, b.lo[a], b.ext[a]);
    b.kind = 1;
    b.face = -1;
    if ((int64_t)b.ext[0] * b.ext[1] * b.ext[2] <= 0) return;
    if (faces_only) {
        const int L[3] = {c->g.nx, c->g.ny, c->g.nz};
        bool face = false;
        for (int a = 0; a < 3; ++a)
            if (!((c->g.per >> a) & 1) && (b.lo[a] == 0 || b.lo[a] + b.ext[a] == L[a])) face = true;
        if (!face) return;
    }
    sg.box[sg.nbox++] = b;
}

int prepare_step(lbm_ctx* c, int64_t n, char* rec, int* grid, int* face_grid, bool allow_sync) {
    StepGeom sg;
    std::memset(&sg, 0, sizeof(sg));
    add_edge_boxes(c, sg);
    sg.nbodies = c->nb;
    std::vector<DevBody> bnew;
    if (!c->flags_only) bodies_for_step(c, n, bnew);
    else bnew = c->step_bodies;
    std::vector<DevBody> bold;
    bool moving = false;
    if (c->nb == 0 && !c->flags_only && !c->geom_valid) {
        // no bodies (any more): clear the flags once
        if (!allow_sync) return kStopChunk;
        int rc = full_geometry(c, bnew);
        if (rc) return rc;
    }
    if (c->nb > 0 && !c->flags_only) {
        if (!c->geom_valid) {
            if (!allow_sync) return kStopChunk;
            int rc = full_geometry(c, bnew);
            if (rc) return rc;
        } else {
            for (int i = 0; i < c->nb; ++i)
                if (!same_mask_pose(bnew[i], c->mask_bodies[i])) moving = true;
        }
    }
    if (!moving && c->nb > 0 && !c->list_valid && !allow_sync) return kStopChunk;
    if (moving) {
        sg.mode = 1;
        bold = c->mask_bodies;
        // body regions: union of the old and new bounding boxes, dilated by the
        // larger reach of the two poses (lbm_set_bodies may change a shape)
        const bool old_reach = (int)c->mask_reach.size() == c->nb;
        for (int i = 0; i < c->nb; ++i) {
            float lo_f[3], hi_f[3];
            for (int a = 0; a < 3; ++a) {
                const float rr = old_reach ? std::max(c->reach[i][a], c->mask_reach[i][a]) : c->reach[i][a];
                lo_f[a] = std::min(bnew[i].c[a], bold[i].c[a]) - rr;
                hi_f[a] = std::max(bnew[i].c[a], bold[i].c[a]) + rr;
            }
            add_body_box(c, sg, lo_f, hi_f);
        }
        finish_boxes(c, sg);
        sg.list_blocks = 0;
        c->mask_bodies = bnew;
        c->mask_reach = c->reach;
        c->list_valid = false;
    } else {
        sg.mode = 0;
        if (c->nb > 0 && !c->list_valid) {
            int rc = build_list(c, nullptr, nullptr, nullptr);
            if (rc) return rc;
            c->list_valid = true;
        }
        // body regions that reach a face (their face cells go to the general path)
        if (c->flags_only) {
            if (c->have_flags_box) add_body_box(c, sg, c->flags_lo, c->flags_hi, true);
        } else {
            for (int i = 0; i < c->nb; ++i) {
                float lo_f[3], hi_f[3];
                for (int a = 0; a < 3; ++a) {
                    lo_f[a] = bnew[i].c[a] - c->reach[i][a];
                    hi_f[a] = bnew[i].c[a] + c->reach[i][a];
                }
                add_body_box(c, sg, lo_f, hi_f, true);
            }
        }
        finish_boxes(c, sg);
        sg.list_blocks = c->nb > 0 ? (int)nblocks_of(c->list_n) : 0;
        bold = bnew;
    }
    c->step_bodies = bnew;
    sg.bg = bulk_geom_of(sg);
    std::memcpy(rec, &sg, sizeof(sg));
    DevBody* bn = (DevBody*)(rec + sizeof(StepGeom));
    for (int i = 0; i < c->nb; ++i) {
        bn[i] = bnew[i];
        bn[c->nb + i] = bold[i];
    }
    *grid = sg.nfix_blocks + sg.list_blocks;
    *face_grid = sg.mode;  // the caller needs the step's mode (k_faces no longer exists)
    return LBM_OK;
}

hipEvent_t next_event(lbm_ctx* c) {
    if (c->ev_used == c->ev_pool.size()) {
        hipEvent_t e;
        if (hipEventCreate(&e) != hipSuccess) return nullptr;
        c->ev_pool.push_back(e);
    }
    return c->ev_pool[c->ev_used++];
}

q27::StepArgs q27_args(const lbm_ctx* c, int mode, const void* fin, void* fout) {
    q27::StepArgs a;
    std::memset(&a, 0, sizeof(a));
    a.mode = mode;
    a.storage = c->cfg.storage;
    a.trt = c->cfg.collision == LBM_TRT;
    a.forced = c->forced ? 1 : 0;
    a.ibb = c->cfg.solid_bc == LBM_SOLID_IBB ? 1 : 0;
    a.g = c->g;
    a.wp = c->pd.wp;
    a.wm = c->pd.wm;
    for (int k = 0; k < 3; ++k) a.grav[k] = c->pd.g[k];
    a.stream = c->stream;
    a.fin = fin;
    a.fout = fout;
    a.fl_new = c->flags;
    a.fl_old = c->flags;
    a.nb = c->nb;
    a.face_u = c->d_face_u;
    a.g.side_in = side_of(c, fin);
    a.g.side_out = side_of(c, fout);
    a.split = c->q27_split ? 1 : 0;
    return a;
}

// Boxes of the bodies of a split 16-bit step (q27::SplitRows): for every body
// the cells with |x_a + 1/2 - c_a| <= ext_a on each axis in the new and in the
// old pose (the device's near test, minimum image on periodic axes), one cell
// wider on each side, the x bounds widened to an even first and an odd last
// cell.  n = -1 (no split this step): a body without an analytic extent
// (explicit flags), a non-finite pose or more boxes than fit.
q27::SplitRows split_rows(const lbm_ctx* c, const DevBody* bnew, const DevBody* bold, int nb) {
    q27::SplitRows r;
    std::memset(&r, 0, sizeof(r));
    const int L[3] = {c->g.nx, c->g.ny, c->g.nz};
    auto fail_split = [&r]() {
        r.n = -1;
        return r;
    };
    for (int b = 0; b < nb; ++b) {
        const DevBody* poses[2] = {&bnew[b], bold ? &bold[b] : nullptr};
        // the two poses in one box unless a periodic wrap lies between them
        bool merge = poses[1] != nullptr;
        for (int a = 0; a < 3 && merge; ++a)
            if (((c->g.per >> a) & 1) && std::fabs((double)poses[0]->c[a] - (double)poses[1]->c[a]) > 0.25 * L[a])
                merge = false;
        for (int k = 0; k < 2; ++k) {
            if (poses[k] == nullptr || (k == 1 && merge)) break;
            int lo[3], hi[3];
            for (int a = 0; a < 3; ++a) {
                double lf = 1e300, hf = -1e300;
                for (int m = 0; m < 2; ++m) {
                    const DevBody* B = (m == 0 || merge) ? poses[m == 0 ? k : 1] : nullptr;
                    if (B == nullptr) continue;
                    if (B->reach < 0.0f) return fail_split();
                    lf = std::min(lf, (double)B->c[a] - 0.5 - (double)B->ext[a]);
                    hf = std::max(hf, (double)B->c[a] - 0.5 + (double)B->ext[a]);
                }
                if (!std::isfinite(lf) || !std::isfinite(hf)) return fail_split();
                lo[a] = (int)std::floor(lf) - 1;
                hi[a] = (int)std::ceil(hf) + 1;
            }
            // x: pair-aligned; a range that wraps on a periodic x is two boxes
            lo[0] = lo[0] >= 0 ? lo[0] & ~1 : -((-lo[0] + 1) & ~1);
            hi[0] |= 1;
            int xs[2][2];
            int nxr = 0;
            if (hi[0] - lo[0] + 1 >= L[0] || !(c->g.per & 1)) {
                xs[0][0] = std::max(lo[0], 0);
                xs[0][1] = std::min(hi[0], L[0] - 1);
                if (hi[0] - lo[0] + 1 >= L[0]) {
                    xs[0][0] = 0;
                    xs[0][1] = L[0] - 1;
                }
                nxr = xs[0][1] >= xs[0][0] ? 1 : 0;
            } else {
                const int l = ((lo[0] % L[0]) + L[0]) % L[0], h = l + (hi[0] - lo[0]);
                if (h < L[0]) {
                    xs[0][0] = l;
                    xs[0][1] = h;
                    nxr = 1;
                } else {
                    xs[0][0] = l;
                    xs[0][1] = L[0] - 1;
                    xs[1][0] = 0;
                    xs[1][1] = h - L[0];
                    nxr = 2;
                }
            }
            // y, z: wrapped start and extent (periodic) or clamped (else)
            int y0 = 0, ey = 0, z0 = 0, ez = 0;
            for (int a = 1; a < 3; ++a) {
                int l = lo[a], h = hi[a], e;
                if ((c->g.per >> a) & 1) {
                    e = std::min(h - l + 1, L[a]);
                    l = ((l % L[a]) + L[a]) % L[a];
                } else {
                    l = std::max(l, 0);
                    h = std::min(h, L[a] - 1);
                    e = std::max(h - l + 1, 0);
                }
                (a == 1 ? y0 : z0) = l;
                (a == 1 ? ey : ez) = e;
            }
            if (ey == 0 || ez == 0) continue;  // entirely beyond a face
            for (int i = 0; i < nxr; ++i) {
                if (r.n == q27::kSplitBoxes) return fail_split();
                r.x0[r.n] = xs[i][0];
                r.ex[r.n] = xs[i][1] - xs[i][0] + 1;
                r.y0[r.n] = y0;
                r.ey[r.n] = ey;
                r.z0[r.n] = z0;
                r.ez[r.n] = ez;
                // blocks an x range of ex cells can span in one row
                r.bpx[r.n] = (r.ex[r.n] + kBlock - 2) / kBlock + 1;
                ++r.n;
            }
        }
    }
    // enumeration of the per-cell blocks: each box's rows, then the face rows
    const bool yf = !((c->g.per >> 1) & 1), zf = !((c->g.per >> 2) & 1);
    r.bpr = L[0] % kBlock == 0 ? L[0] / kBlock : L[0] / kBlock + 2;
    r.nface = (zf ? 2 * L[1] : 0) + (yf ? 2 * (zf ? L[2] - 2 : L[2]) : 0);
    int64_t v = 0;
    for (int k = 0; k < r.n; ++k) {
        r.base[k] = (int)v;
        v += (int64_t)r.ey[k] * r.ez[k] * r.bpx[k];
    }
    r.base[r.n] = (int)v;
    r.total = v + (int64_t)r.nface * r.bpr;
    if (r.total >= ((int64_t)1 << 30)) r.n = -1;
    return r;
}

int q27_step_launch(lbm_ctx* c, const q27::StepArgs& a) {
    const int e = q27_launch_step(a);
    if (e) return fail(c, LBM_ERR_HIP, std::string("kernel launch: ") + hipGetErrorString((hipError_t)e));
    return LBM_OK;
}

// D3Q27 steps: per step the bodies of the new pose (and of the old one when it
// changed) go into the step record; a pose change rebuilds the flags (the old
// ones kept for the refill) before the step's single launch.
// the flag boxes of a step record sit behind its 2 nb bodies, in the room the
// D3Q19 step geometry takes
static_assert(sizeof(StepGeom) >= LBM_MAX_BODIES * 6 * sizeof(int32_t), "step record too small for the flag boxes");

// Seed of step n's stochastic rounding (16-bit storage): distinct per step,
// fixed by n alone, so a run repeats bitwise.
uint32_t step_seed(int64_t n) {
    uint32_t x = (uint32_t)n * 0x9e3779b9u + (uint32_t)(n >> 32) + 0x85ebca6bu;
    x ^= x >> 16;
    x *= 0x21f0aaadu;
    x ^= x >> 15;
    x *= 0xd35a2d97u;
    x ^= x >> 15;
    return x;
}

int run_chunk27(lbm_ctx* c, int nsteps, double* force_out, int64_t* done) {
    // block partials of every step of the chunk: at most 1 GiB of them
    {
        const int64_t per_step = nblocks_of(c->g.N) * std::max(c->nb, 1) * 6 * (int64_t)sizeof(double);
        nsteps = (int)std::max<int64_t>(1, std::min<int64_t>(nsteps, ((int64_t)1 << 30) / per_step));
    }
    const int par = c->parity;
    if (c->rec_ev_used[par]) HIPCHK(c, hipEventSynchronize(c->rec_ev[par]));
    const int nb = c->nb;
    std::vector<char> moved((size_t)nsteps, 0), flag_boxes((size_t)nsteps, 0);
    std::vector<DevBody> bnew;
    const int L[3] = {c->g.nx, c->g.ny, c->g.nz};
    int max_box = 0;
    int nprep = 0;
    for (int s = 0; s < nsteps; ++s) {
        bool moving = false;
        if (c->flags_only) {
            bnew = c->step_bodies;
        } else {
            bodies_for_step(c, c->step + 1 + s, bnew);
            if (!c->geom_valid) {  // (with no bodies: the flags are cleared once)
                if (s > 0) break;  // the full build runs before any launch of the chunk
                int rc = full_geometry(c, bnew);
                if (rc) return rc;
            } else {
                for (int i = 0; i < nb; ++i)
                    if (!same_mask_pose(bnew[i], c->mask_bodies[i])) moving = true;
            }
        }
        DevBody* rb = (DevBody*)(c->h_rec[par] + (size_t)s * c->rec_bytes);
        for (int i = 0; i < nb; ++i) {
            rb[i] = bnew[i];
            rb[nb + i] = moving ? c->mask_bodies[i] : bnew[i];
        }
        if (moving) {
            c->mask_bodies = bnew;
            c->mask_reach = c->reach;
            // The buffers swap: the one that becomes current holds the flags of
            // its own (older) pose, which differ from the new pose's only inside
            // each body's box around both poses; the step record carries the boxes.
            std::swap(c->q27_pose_cur, c->q27_pose_old);
            std::swap(c->q27_reach_cur, c->q27_reach_old);
            std::swap(c->q27_cur_ok, c->q27_old_ok);
            if (c->q27_cur_ok && (int)c->q27_pose_cur.size() == nb && (int)c->q27_reach_cur.size() == nb &&
                (int)c->reach.size() == nb) {
                int32_t* bx = (int32_t*)(c->h_rec[par] + (size_t)s * c->rec_bytes + 2 * (size_t)nb * sizeof(DevBody));
                for (int i = 0; i < nb; ++i) {
                    int64_t vol = 1;
                    for (int a = 0; a < 3; ++a) {
                        // the larger reach of the two poses (lbm_set_bodies may change a shape)
                        const float r = std::max(c->reach[i][a], c->q27_reach_cur[i][a]);
                        const float lo_f = std::min(c->q27_pose_cur[i].c[a], bnew[i].c[a]) - r - 2.0f;
                        const float hi_f = std::max(c->q27_pose_cur[i].c[a], bnew[i].c[a]) + r + 2.0f;
                        int lo = 0, ext = L[a];
                        if (std::isfinite(lo_f) && std::isfinite(hi_f) && hi_f - lo_f + 1.0f < (float)L[a]) {
                            lo = (int)std::floor(lo_f);
                            const int hi = (int)std::ceil(hi_f);
                            if ((c->g.per >> a) & 1) {
                                ext = std::min(hi - lo + 1, L[a]);
                                lo = ((lo % L[a]) + L[a]) % L[a];
                            } else {
                                lo = std::max(lo, 0);
                                ext = std::max(std::min(hi, L[a] - 1) - lo + 1, 0);
                            }
                        }
                        bx[6 * i + a] = lo;
                        bx[6 * i + 3 + a] = ext;
                        vol *= ext;
                    }
                    max_box = (int)std::max<int64_t>(max_box, vol);
                }
                flag_boxes[(size_t)s] = 1;
            }
            c->q27_pose_cur = bnew;
            c->q27_reach_cur = c->reach;
            c->q27_cur_ok = true;
        }
        c->step_bodies = bnew;
        moved[(size_t)s] = moving ? 1 : 0;
        ++nprep;
    }
    HIPCHK(c, hipMemcpyAsync(c->d_rec, c->h_rec[par], (size_t)nprep * c->rec_bytes, hipMemcpyHostToDevice, c->stream));
    const int64_t nblk = nblocks_of(c->g.N);
    const int64_t stride = nblk * std::max(nb, 1) * 6;
    int rc = ensure_partials(c, nblk * std::max(nprep, 1));
    if (rc) return rc;
    const int n6 = nb * 6;
    for (int s = 0; s < nprep; ++s) {
        const DevBody* dn = (const DevBody*)(c->d_rec + (size_t)s * c->rec_bytes);
        const bool mv = moved[(size_t)s] != 0;
        if (mv) {
            // the flags of the old pose stay in the other buffer for the refill
            std::swap(c->flags, c->flags_old);
            if (flag_boxes[(size_t)s])
                GEOCHK(c, q27_launch_flags_boxes(c->flags, dn, nb, c->g,
                                                 (const int32_t*)((const char*)dn + 2 * (size_t)nb * sizeof(DevBody)),
                                                 max_box, c->stream));
            else
                GEOCHK(c, launch_flags_full(c->flags, dn, nb, c->g, c->stream));
        }
        c->g.sh.seed = step_seed(c->step + 1 + s);
        q27::StepArgs a = q27_args(c, MODE_STEP, c->fbuf[c->cur], c->fbuf[1 - c->cur]);
        a.fl_old = mv ? c->flags_old : c->flags;
        a.bnew = dn;
        a.bold = mv ? dn + nb : nullptr;
        a.refill = (c->cfg.refill && mv) ? 1 : 0;
        a.partials = c->d_partials + (size_t)s * stride;
        if (a.split) {
            const DevBody* hb = (const DevBody*)(c->h_rec[par] + (size_t)s * c->rec_bytes);
            a.rows = split_rows(c, hb, mv ? hb + nb : nullptr, nb);
            if (a.rows.n < 0) a.split = 0;
        }
        hipEvent_t e0 = nullptr, e1 = nullptr;
        const bool sample = c->profile_stride > 0 && ((c->step + s) % c->profile_stride == 0) &&
                            c->ev_used + 2 <= kMaxProfileEvents;
        if (sample) {
            e0 = next_event(c);
            e1 = next_event(c);
            HIPCHK(c, hipEventRecord(e0, c->stream));
        }
        rc = q27_step_launch(c, a);
        if (rc) return rc;
        if (sample) HIPCHK(c, hipEventRecord(e1, c->stream));
        c->cur = 1 - c->cur;
        c->q27_bnew = a.bnew;
        c->q27_bold = a.bold;
        c->q27_refill = a.refill;
        c->last_rec_valid = true;
        c->raw_state = false;
    }
    if (n6 > 0 && nprep > 0)
        GEOCHK(c, q27_launch_force_sum(c->d_partials, nblk, stride, n6, nprep, c->d_rec, c->rec_bytes, c->g,
                                       c->d_forces, c->stream));
    // the records always come back: lbm_step checks them for non-finite values
    if (n6 > 0) {
        HIPCHK(c, hipMemcpyAsync(c->h_forces[par], c->d_forces, (size_t)nprep * n6 * sizeof(double),
                                 hipMemcpyDeviceToHost, c->stream));
    }
    HIPCHK(c, hipEventRecord(c->rec_ev[par], c->stream));
    c->rec_ev_used[par] = true;
    c->parity = 1 - par;
    c->step += nprep;
    c->pending_set = false;
    *done = nprep;
    return LBM_OK;
}

int run_chunk(lbm_ctx* c, int nsteps, double* force_out, int64_t* done) {
    if (c->q == 27) return run_chunk27(c, nsteps, force_out, done);
    // the pinned buffers of this parity are free once their last copy finished
    const int par = c->parity;
    if (c->rec_ev_used[par]) HIPCHK(c, hipEventSynchronize(c->rec_ev[par]));
    std::vector<int> grids(nsteps), fgrids(nsteps);
    int nprep = 0;
    for (int s = 0; s < nsteps; ++s) {
        int rc = prepare_step(c, c->step + 1 + s, c->h_rec[par] + (size_t)s * c->rec_bytes, &grids[s], &fgrids[s],
                              s == 0);
        if (rc == kStopChunk) break;
        if (rc) return rc;
        ++nprep;
    }
    HIPCHK(c, hipMemcpyAsync(c->d_rec, c->h_rec[par], (size_t)nprep * c->rec_bytes, hipMemcpyHostToDevice, c->stream));
    int maxgrid = 0;
    for (int s = 0; s < nprep; ++s) maxgrid = std::max(maxgrid, grids[s]);
    // block partials of every step of the chunk, summed once per chunk
    const int64_t stride = (int64_t)std::max(maxgrid, 1) * std::max(c->nb, 1) * 6;
    int rc = ensure_partials(c, (int64_t)std::max(maxgrid, 1) * nprep);
    if (rc) return rc;
    const int n6 = c->nb * 6;
    for (int s = 0; s < nprep; ++s) {
        const char* rec = c->d_rec + (size_t)s * c->rec_bytes;
        void* fin = c->fbuf[c->cur];
        void* fout = c->fbuf[1 - c->cur];
        hipEvent_t e0 = nullptr, e1 = nullptr;
        const bool sample = c->profile_stride > 0 && ((c->step + s) % c->profile_stride == 0) &&
                            c->ev_used + 2 <= kMaxProfileEvents;
        if (sample) {
            e0 = next_event(c);
            e1 = next_event(c);
            HIPCHK(c, hipEventRecord(e0, c->stream));
        }
        c->g.sh.seed = step_seed(c->step + 1 + s);
        int lr = launch_step(c, MODE_STEP, grids[s], fin, fout, nullptr, nullptr, nullptr, rec, 1,
                             c->d_partials + (size_t)s * stride);
        if (lr) return lr;
        if (sample) HIPCHK(c, hipEventRecord(e1, c->stream));
        c->cur = 1 - c->cur;
        c->last_rec = rec;
        c->last_grid = grids[s];
        c->last_rec_valid = true;
        c->raw_state = false;
    }
    if (n6 > 0 && nprep > 0)
        GEOCHK(c, launch_force_finalize(c->d_partials, stride, c->d_rec, c->rec_bytes, nprep, n6, c->d_forces,
                                        c->stream));
    // the records always come back: lbm_step checks them for non-finite values
    if (n6 > 0) {
        HIPCHK(c, hipMemcpyAsync(c->h_forces[par], c->d_forces, (size_t)nprep * n6 * sizeof(double),
                                 hipMemcpyDeviceToHost, c->stream));
    }
    HIPCHK(c, hipEventRecord(c->rec_ev[par], c->stream));
    c->rec_ev_used[par] = true;
    c->parity = 1 - par;
    c->step += nprep;
    c->pending_set = false;
    *done = nprep;
    return LBM_OK;
}

int collide_only(lbm_ctx* c) {
    c->g.sh.seed = step_seed(c->step);
    if (c->q == 27) return q27_step_launch(c, q27_args(c, MODE_COLLIDE_ONLY, c->fbuf[1 - c->cur], c->fbuf[c->cur]));
    return launch_bulk(c, MODE_COLLIDE_ONLY, c->fbuf[1 - c->cur], c->fbuf[c->cur], nullptr, nullptr, nullptr, nullptr);
}

// Pulled state f_old and moments of the last completed step.
int get_state(lbm_ctx* c, double* d_f, double* d_rho, double* d_u) {
    if (!c->initialised) return fail(c, LBM_ERR_STATE, "not initialised");
    const void* prev = c->fbuf[1 - c->cur];
    if (c->q == 27) {
        if (!c->raw_state && !c->last_rec_valid) return fail(c, LBM_ERR_STATE, "no completed step to read");
        q27::StepArgs a = q27_args(c, c->raw_state ? MODE_GET_RAW : MODE_GET, prev, nullptr);
        a.gout = d_f;
        a.rho_out = d_rho;
        a.u_out = d_u;
        a.bnew = c->q27_bnew;
        a.bold = c->q27_bold;
        a.fl_old = c->q27_bold ? c->flags_old : c->flags;
        a.refill = c->q27_refill;
        return q27_step_launch(c, a);
    }
    if (c->raw_state) return launch_bulk(c, MODE_GET_RAW, prev, nullptr, d_f, d_rho, d_u, nullptr);
    if (!c->last_rec_valid) return fail(c, LBM_ERR_STATE, "no completed step to read");
    return launch_step(c, MODE_GET, c->last_grid, prev, nullptr, d_f, d_rho, d_u, c->last_rec, 0, nullptr);
}

// ---------------------------------------------------------------------------
// Allocation-free conservation sums (lbm_conservation_static).
//
// The state the context holds is f* = C(f_old) for every cell, and the
// collision conserves rho and adds exactly rho g to rho u (the Guo source), so
// the fluid cells' mass and momentum of f_old are sum f* and
// sum c f* - M g / 2 (u carries the half-force shift, src/ComputeMacro.jl:26-45
// plus the forcing).  One pass over the stored f* gives them, with no
// per-cell moment arrays: a fixed grid of kConsBlocks blocks strides over the
// cells, each block writes its four partial sums (every block writes, so
// nothing needs clearing), and one block adds the partials in block order.
// The partials and the result live in `scratch`, a file-scope device array:
// each GPU holds its own copy for the life of the process, and its device
// address is looked up once per device when the first context is created
// there.  Calls that overlap on one GPU share it and race.
// ---------------------------------------------------------------------------
constexpr int kConsBlocks = 1024;
constexpr int kConsThreads = 256;
constexpr int kScratchDoubles = kConsBlocks * 4 + 4;
static __device__ __align__(256) double scratch[kScratchDoubles];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch[kMaxDevices] = {};

// Device address of `scratch` on the current device (once per device).
int scratch_init(lbm_ctx* c) {
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (c->device < 0 || c->device >= kMaxDevices) return fail(c, LBM_ERR_ARG, "device index beyond the scratch table");
    if (g_scratch[c->device]) return LBM_OK;
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
        return fail(c, LBM_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch[c->device] = static_cast<double*>(p);
    return LBM_OK;
}

double* scratch_of(int device) {
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    return device >= 0 && device < kMaxDevices ? g_scratch[device] : nullptr;
}

struct ConsDirs {
    int8_t cx[27], cy[27], cz[27];
    float w[27];  // weights in FP32, for the shifted 16-bit formats
};

// One stored population as the storage format's load returns it, in FP64.
template <int ST>
__device__ __forceinline__ double stored_pop(const void* f, int64_t i, float w, float ik) {
    if (ST == LBM_FP64) return static_cast<const double*>(f)[i];
    if (ST == LBM_FP32) return (double)static_cast<const float*>(f)[i];
    const uint16_t h = static_cast<const uint16_t*>(f)[i];
    if (ST == LBM_FP16S) return (double)__fmaf_rn(__half2float(__ushort_as_half(h)), ik, w);
    return (double)(__uint_as_float(uint32_t(h) << 16) + w);
}

template <int ST>
__global__ __launch_bounds__(kConsThreads) void k_cons_partials(const void* __restrict__ f, int64_t P,
                                                                const uint8_t* __restrict__ flags, int64_t N, int q,
                                                                ConsDirs t, float ik, double* __restrict__ part) {
    double s[4] = {0.0, 0.0, 0.0, 0.0};
    for (int64_t cell = (int64_t)blockIdx.x * kConsThreads + threadIdx.x; cell < N;
         cell += (int64_t)gridDim.x * kConsThreads) {
        if (flags[cell] & kFlagBody) continue;  // solid (the fix-up bit marks fluid boundary nodes)
        double r = 0.0, jx = 0.0, jy = 0.0, jz = 0.0;
        for (int d = 0; d < q; ++d) {
            const double v = stored_pop<ST>(f, (int64_t)d * P + cell, t.w[d], ik);
            r += v;
            jx += (double)t.cx[d] * v;
            jy += (double)t.cy[d] * v;
            jz += (double)t.cz[d] * v;
        }
        s[0] += r;
        s[1] += jx;
        s[2] += jy;
        s[3] += jz;
    }
    __shared__ double lds[kConsThreads / 64][4];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        double v = s[k];
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
        if (lane == 0) lds[wave][k] = v;
    }
    __syncthreads();
    if (threadIdx.x < 4) {
        double v = lds[0][threadIdx.x];
        for (int w = 1; w < kConsThreads / 64; ++w) v += lds[w][threadIdx.x];
        part[(int64_t)blockIdx.x * 4 + threadIdx.x] = v;
    }
}

// out = (mass, momentum - mass g / 2) from the partials, added in block order
// (one wave; lanes 0..3 own the four sums)
__global__ void k_cons_final(const double* __restrict__ part, int nblocks, double gx, double gy, double gz,
                             double* __restrict__ out) {
    const int k = threadIdx.x;
    double v = 0.0;
    if (k < 4)
        for (int b = 0; b < nblocks; ++b) v += part[(int64_t)b * 4 + k];
    const double mass = __shfl(v, 0, 64);
    const double g = k == 1 ? gx : (k == 2 ? gy : gz);
    if (k < 4) out[k] = k == 0 ? v : v - 0.5 * mass * g;
}

}  // namespace

extern "C" {

// D3Q27 links from the flags: per-block counts, their scan, then the fill.
static int get_links27(lbm_ctx* c, int32_t* cell, int8_t* dir, int64_t cap, int64_t* nnz) {
    const int64_t nblk = nblocks_of(c->g.N);
    int rc = ensure_counts(c, nblk + 1);
    if (rc) return rc;
    GEOCHK(c, q27_launch_link_count(c->flags, c->g, c->cnt_a, c->stream));
    HIPCHK(c, hipMemsetAsync(c->cnt_b, 0, (size_t)(nblk + 1) * 4, c->stream));
    GEOCHK(c, launch_scan2(c->cnt_a, c->cnt_b, (int)nblk, c->stream));
    int32_t tot = 0;
    HIPCHK(c, hipMemcpyAsync(&tot, c->cnt_a + nblk, 4, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    const int64_t need = tot;
    if (nnz) *nnz = need;
    if (!cell && !dir) return LBM_OK;
    if (need > cap) return fail(c, LBM_ERR_CAPACITY, "link buffer too small");
    int32_t* d_cell = nullptr;
    int8_t* d_dir = nullptr;
    rc = dev_alloc(c, (void**)&d_cell, (size_t)std::max<int64_t>(need, 1) * 4);
    if (rc == LBM_OK) rc = dev_alloc(c, (void**)&d_dir, (size_t)std::max<int64_t>(need, 1));
    if (rc == LBM_OK) {
        const int e = q27_launch_link_fill(c->flags, c->g, c->cnt_a, d_cell, d_dir, c->stream);
        if (e) rc = fail(c, LBM_ERR_HIP, std::string("link fill: ") + hipGetErrorString((hipError_t)e));
    }
    if (rc == LBM_OK && need) {
        if (cell && hipMemcpyAsync(cell, d_cell, (size_t)need * 4, hipMemcpyDeviceToHost, c->stream) != hipSuccess)
            rc = fail(c, LBM_ERR_HIP, "link copy");
        if (dir && hipMemcpyAsync(dir, d_dir, (size_t)need, hipMemcpyDeviceToHost, c->stream) != hipSuccess)
            rc = fail(c, LBM_ERR_HIP, "link copy");
    }
    if (hipStreamSynchronize(c->stream) != hipSuccess && rc == LBM_OK) rc = fail(c, LBM_ERR_HIP, "link sync");
    dev_free(c, d_cell, (size_t)std::max<int64_t>(need, 1) * 4);
    dev_free(c, d_dir, (size_t)std::max<int64_t>(need, 1));
    return rc;
}

int lbm_create(const lbm_config* cfg, lbm_ctx** out) {
    if (!cfg || !out) return LBM_ERR_ARG;
    *out = nullptr;
    if (cfg->collision != LBM_BGK && cfg->collision != LBM_TRT) return LBM_ERR_ARG;
    if (cfg->storage < LBM_FP64 || cfg->storage > LBM_BF16S) return LBM_ERR_ARG;
    if (cfg->solid_bc != LBM_SOLID_SBB && cfg->solid_bc != LBM_SOLID_IBB) return LBM_ERR_ARG;
    const int q = cfg->q == 0 ? 19 : cfg->q;
    if (q != 19 && q != 27) return LBM_ERR_ARG;
    const int L[3] = {cfg->nx, cfg->ny, cfg->nz};
    for (int a = 0; a < 3; ++a) {
        const int k0 = cfg->faces[2 * a].kind, k1 = cfg->faces[2 * a + 1].kind;
        if (k0 < 0 || k0 > 3 || k1 < 0 || k1 > 3) return LBM_ERR_ARG;
        if ((k0 == LBM_FACE_PERIODIC) != (k1 == LBM_FACE_PERIODIC)) return LBM_ERR_ARG;
        if (L[a] < (k0 == LBM_FACE_PERIODIC ? 2 : 3)) return LBM_ERR_ARG;
    }
    const int64_t N = (int64_t)cfg->nx * cfg->ny * cfg->nz;
    if (N >= ((int64_t)1 << 31)) return LBM_ERR_ARG;  // cell indices are int32 in the lists
    // 16-bit storage on D3Q19: a bulk thread owns an aligned pair of x cells
    if (q == 19 && (cfg->storage == LBM_FP16S || cfg->storage == LBM_BF16S) && (cfg->nx % 2) != 0) return LBM_ERR_ARG;
    if (!(cfg->omega_plus > 0.0 && cfg->omega_plus < 2.0)) return LBM_ERR_ARG;
    if (cfg->collision == LBM_TRT && !(cfg->omega_minus > 0.0 && cfg->omega_minus < 2.0)) return LBM_ERR_ARG;

    lbm_ctx* c = new lbm_ctx();
    c->cfg = *cfg;
    c->cfg.q = q;
    {
        const char* e = std::getenv("LBM_Q27_SPLIT");
        // (the pair blocks address a plane with 32-bit byte offsets: N * 2 B
        // plus the plane pad stays below 4 GiB for N < 2^31 - 2^12)
        c->q27_split = q == 27 && (cfg->storage == LBM_FP16S || cfg->storage == LBM_BF16S) && cfg->nx % 2 == 0 &&
                       cfg->nx >= 4 && N < ((int64_t)1 << 31) - 4096 && !(e && e[0] == '0');
    }
    c->q = q;
    c->g.nx = cfg->nx;
    c->g.ny = cfg->ny;
    c->g.nz = cfg->nz;
    c->g.N = N;
    c->g.nxy = (int64_t)cfg->nx * cfg->ny;
    c->g.per = 0;
    for (int a = 0; a < 3; ++a)
        if (cfg->faces[2 * a].kind == LBM_FACE_PERIODIC) c->g.per |= 1 << a;
    for (int f = 0; f < 6; ++f) {
        c->g.face_kind[f] = cfg->faces[f].kind;

    }
    {
        // face-moment slots of the velocity / anti-bounce-back faces
        int64_t slots = 0;
        const int64_t area[3] = {(int64_t)cfg->ny * cfg->nz, (int64_t)cfg->nx * cfg->nz, (int64_t)cfg->nx * cfg->ny};
        for (int f = 0; f < 6; ++f) {
            const int k = cfg->faces[f].kind;
            if (k == LBM_FACE_VELOCITY || k == LBM_FACE_PRESSURE_ABB) {
                c->g.face_base[f] = (int)slots;
                slots += area[f / 2];
            } else {
                c->g.face_base[f] = -1;
            }
        }
        const int cbytes = cfg->storage == LBM_FP64 ? 8 : 4;
        c->side_bytes = (size_t)std::max<int64_t>(slots, 1) * 4 * cbytes;
    }
    {
        // 16-bit storage scale 2^12 (FP16S; BF16S shares FP32's exponent range)
        const bool f16 = cfg->storage == LBM_FP16S;
        c->g.sh.k = f16 ? 4096.0f : 1.0f;
        c->g.sh.ik = f16 ? 1.0f / 4096.0f : 1.0f;
    }
    c->g.wall_point = cfg->wall_point;
    c->g.refill = cfg->refill;
    c->g.rho0 = cfg->rho0;
    c->forced = cfg->g[0] != 0.0 || cfg->g[1] != 0.0 || cfg->g[2] != 0.0;
    c->pd.wp = cfg->omega_plus;
    c->pd.wm = cfg->collision == LBM_TRT ? cfg->omega_minus : cfg->omega_plus;
    c->pf.wp = (float)c->pd.wp;
    c->pf.wm = (float)c->pd.wm;
    for (int k = 0; k < 3; ++k) {
        c->pd.g[k] = cfg->g[k];
        c->pf.g[k] = (float)cfg->g[k];
    }
    const int bytes = cfg->storage == LBM_FP64 ? 8 : (cfg->storage == LBM_FP32 ? 4 : 2);
    // Plane stride: N rounded up to 256 B plus 2304 B, so consecutive direction
    // planes do not start at the same offset modulo the HBM channel interleave
    // (N * 4 B is exactly 2^24 B for a 256x128x128 lattice).
    const int64_t align = 256 / bytes;
    c->g.P = (N + align - 1) / align * align + 2304 / bytes;
    c->fbytes = (size_t)q * c->g.P * bytes;
    if (hipGetDevice(&c->device) != hipSuccess || hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        delete c;
        return LBM_ERR_HIP;
    }
    int rc = scratch_init(c);
    for (int i = 0; i < 2 && rc == LBM_OK; ++i) {
        rc = dev_alloc(c, &c->fbuf_raw[i], c->fbytes + kFrontPad);
        if (rc == LBM_OK) c->fbuf[i] = (char*)c->fbuf_raw[i] + kFrontPad;
    }
    for (int i = 0; i < 2 && rc == LBM_OK; ++i) {
        rc = dev_alloc(c, &c->side[i], c->side_bytes);
        if (rc == LBM_OK && hipMemsetAsync(c->side[i], 0, c->side_bytes, c->stream) != hipSuccess) rc = LBM_ERR_HIP;
    }
    if (rc == LBM_OK) rc = dev_alloc(c, (void**)&c->d_face_u, 18 * sizeof(double));
    if (rc == LBM_OK) {
        double fu[18];
        for (int f = 0; f < 6; ++f)
            for (int k = 0; k < 3; ++k) fu[3 * f + k] = cfg->faces[f].u[k];
        if (hipMemcpyAsync(c->d_face_u, fu, sizeof(fu), hipMemcpyHostToDevice, c->stream) != hipSuccess ||
            hipStreamSynchronize(c->stream) != hipSuccess)
            rc = LBM_ERR_HIP;
        c->g.face_u = c->d_face_u;
    }
    if (rc == LBM_OK) rc = dev_alloc(c, (void**)&c->flags, (size_t)N);
    if (rc == LBM_OK && q == 27) rc = dev_alloc(c, (void**)&c->flags_old, (size_t)N);
    if (rc == LBM_OK) rc = dev_alloc(c, (void**)&c->list_count, 16);
    if (rc == LBM_OK) {
        if (hipMemsetAsync(c->flags, 0, (size_t)N, c->stream) != hipSuccess ||
            hipMemsetAsync(c->list_count, 0, 16, c->stream) != hipSuccess)
            rc = LBM_ERR_HIP;
    }
    for (int i = 0; i < 2 && rc == LBM_OK; ++i)
        if (hipEventCreateWithFlags(&c->rec_ev[i], hipEventDisableTiming) != hipSuccess) rc = LBM_ERR_HIP;
    if (rc == LBM_OK) rc = ensure_records(c);
    if (rc == LBM_OK && hipStreamSynchronize(c->stream) != hipSuccess) rc = LBM_ERR_HIP;
    if (rc != LBM_OK) {
        lbm_destroy(c);
        return rc;
    }
    *out = c;
    return LBM_OK;
}

int lbm_destroy(lbm_ctx* c) {
    if (!c) return LBM_OK;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    for (int i = 0; i < 2; ++i) {
        if (c->fbuf_raw[i]) (void)hipFree(c->fbuf_raw[i]);
        if (c->side[i]) (void)hipFree(c->side[i]);
        if (c->h_rec[i]) (void)hipHostFree(c->h_rec[i]);
        if (c->h_forces[i]) (void)hipHostFree(c->h_forces[i]);
        if (c->rec_ev[i]) (void)hipEventDestroy(c->rec_ev[i]);
    }
    void* ptrs[] = {c->flags, c->flags_old, c->list, c->list_count, c->cnt_a, c->cnt_b, c->d_rec, c->d_partials, c->d_forces,
                    c->d_face_u};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    for (hipEvent_t e : c->ev_pool) (void)hipEventDestroy(e);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return LBM_OK;
}

int lbm_last_error(const lbm_ctx* c, char* buf, size_t n) {
    if (!c || !buf || n == 0) return LBM_ERR_ARG;
    std::snprintf(buf, n, "%s", c->err.c_str());
    return LBM_OK;
}

static int set_body_count(lbm_ctx* c, int n) {
    if (n < 0 || n > LBM_MAX_BODIES) return fail(c, LBM_ERR_ARG, "body count out of range");
    if (n != c->nb) {
        c->nb = n;
        c->geom_valid = false;
        c->list_valid = false;
        c->q27_cur_ok = c->q27_old_ok = false;
        int rc = ensure_records(c);
        if (rc) return rc;
    }
    return LBM_OK;
}

int lbm_set_bodies(lbm_ctx* c, int32_t n, const lbm_body* bodies) {
    if (!c || (n > 0 && !bodies)) return LBM_ERR_ARG;
    for (int i = 0; i < n; ++i) {
        const lbm_body& b = bodies[i];
        const double semi[3] = {b.semi[0], b.semi[1], b.semi[2]};
        double rot[9];
        for (int k = 0; k < 9; ++k) rot[k] = b.rot[k];
        if (const char* e = shape_error(b.shape, semi, rot, b.omega)) return fail(c, LBM_ERR_ARG, e);
        if (b.shape == LBM_NACA4) {
            for (int k = 0; k < 9; ++k) rot[k] = b.rot_half[k];
            if (const char* e = shape_error(b.shape, semi, rot, b.omega)) return fail(c, LBM_ERR_ARG, e);
        }
        if (c->cfg.solid_bc == LBM_SOLID_IBB && (b.shape == LBM_ELLIPSOID || b.shape == LBM_NACA4)) {
            // the wall fraction reads the half-step rotation: it must be one
            for (int i2 = 0; i2 < 3; ++i2)
                for (int j2 = 0; j2 < 3; ++j2) {
                    double dot = 0.0;
                    for (int k = 0; k < 3; ++k) dot += b.rot_half[3 * k + i2] * b.rot_half[3 * k + j2];
                    if (std::fabs(dot - (i2 == j2 ? 1.0 : 0.0)) > 1e-6)
                        return fail(c, LBM_ERR_ARG, "rot_half must be a rotation (interpolated bounce-back)");
                }
        }
    }
    if (c->flags_only) {
        c->flags_only = false;
        c->geom_valid = false;
    }
    int rc = set_body_count(c, n);
    if (rc) return rc;
    c->have_laws = false;
    c->laws.clear();
    c->pending.assign(bodies, bodies + n);
    c->pending_set = true;
    c->reach.clear();
    for (int i = 0; i < n; ++i) {
        const double semi[3] = {bodies[i].semi[0], bodies[i].semi[1], bodies[i].semi[2]};
        double naca[5];
        for (int k = 0; k < 5; ++k) naca[k] = bodies[i].naca[k];
        c->reach.push_back(reach_of(bodies[i].shape, semi, naca, c->g.nz));
    }
    return LBM_OK;
}

int lbm_set_motion(lbm_ctx* c, int32_t n, const lbm_motion* laws) {
    if (!c || (n > 0 && !laws)) return LBM_ERR_ARG;
    for (int i = 0; i < n; ++i)
        if (const char* e = shape_error(laws[i].shape, laws[i].semi, laws[i].rot0, laws[i].omega))
            return fail(c, LBM_ERR_ARG, e);
    if (c->flags_only) {
        c->flags_only = false;
        c->geom_valid = false;
    }
    int rc = set_body_count(c, n);
    if (rc) return rc;
    c->laws.assign(laws, laws + n);
    c->have_laws = n > 0;
    c->pending.clear();
    c->pending_set = false;
    c->reach.clear();
    for (int i = 0; i < n; ++i) c->reach.push_back(reach_of(laws[i].shape, laws[i].semi, laws[i].naca, c->g.nz));
    return LBM_OK;
}

int lbm_set_flags(lbm_ctx* c, const uint8_t* flags) {
    if (!c || !flags) return LBM_ERR_ARG;
    if (c->cfg.solid_bc == LBM_SOLID_IBB)
        return fail(c, LBM_ERR_ARG, "the interpolated bounce-back needs analytic bodies, not explicit flags");
    int maxf = 0;
    int lo[3] = {c->g.nx, c->g.ny, c->g.nz}, hi[3] = {-1, -1, -1};
    for (int64_t i = 0; i < c->g.N; ++i) {
        if (!flags[i]) continue;
        maxf = std::max<int>(maxf, flags[i]);
        const int q[3] = {(int)(i % c->g.nx), (int)((i / c->g.nx) % c->g.ny), (int)(i / c->g.nxy)};
        for (int a = 0; a < 3; ++a) {
            lo[a] = std::min(lo[a], q[a]);
            hi[a] = std::max(hi[a], q[a]);
        }
    }
    if (maxf > LBM_MAX_BODIES) return fail(c, LBM_ERR_ARG, "flag value above the body limit");
    c->have_flags_box = maxf > 0;
    for (int a = 0; a < 3; ++a) {
        c->flags_lo[a] = (float)lo[a] + 0.5f;
        c->flags_hi[a] = (float)hi[a] + 0.5f;
    }
    int rc = set_body_count(c, maxf);
    if (rc) return rc;
    c->have_laws = false;
    c->laws.clear();
    c->pending.clear();
    c->pending_set = false;
    c->flags_only = true;
    // bodies without kinematics: at rest, centre at the origin
    c->reach.assign(maxf, std::array<float, 3>{0.0f, 0.0f, 0.0f});
    DevBody unknown{};
    unknown.reach = -1.0f;  // no analytic shape: every face cell checks its neighbours' flags
    c->step_bodies.assign(maxf, unknown);
    c->mask_bodies = c->step_bodies;
    c->mask_reach = c->reach;
    HIPCHK(c, hipMemcpyAsync(c->flags, flags, (size_t)c->g.N, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    c->geom_valid = true;
    c->last_rec_valid = false;
    c->q27_cur_ok = c->q27_old_ok = false;  // explicit flags, no analytic pose
    if (c->q == 27) return LBM_OK;
    rc = build_list(c, nullptr, nullptr, nullptr);
    if (rc) return rc;
    c->list_valid = true;
    return LBM_OK;
}

int lbm_update_geometry(lbm_ctx* c) {
    if (!c) return LBM_ERR_ARG;
    if (c->flags_only || c->nb == 0) return LBM_OK;
    std::vector<DevBody> b;
    bodies_for_step(c, c->step + 1, b);
    int rc = full_geometry(c, b);
    if (rc) return rc;
    c->step_bodies = b;
    c->last_rec_valid = false;
    return LBM_OK;
}

int lbm_get_flags(lbm_ctx* c, uint8_t* flags) {
    if (!c || !flags) return LBM_ERR_ARG;
    HIPCHK(c, hipMemcpyAsync(flags, c->flags, (size_t)c->g.N, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    for (int64_t i = 0; i < c->g.N; ++i) flags[i] &= kFlagBody;
    return LBM_OK;
}

int lbm_get_links(lbm_ctx* c, int32_t* cell, int8_t* dir, int64_t cap, int64_t* nnz) {
    if (!c) return LBM_ERR_ARG;
    if (c->q == 27) return get_links27(c, cell, dir, cap, nnz);
    // count first
    int64_t need = 0;
    int rc = build_list(c, nullptr, nullptr, &need);
    if (rc) return rc;
    c->list_valid = true;
    if (nnz) *nnz = need;
    if (!cell && !dir) return LBM_OK;
    if (need > cap) return fail(c, LBM_ERR_CAPACITY, "link buffer too small");
    int32_t* d_cell = nullptr;
    int8_t* d_dir = nullptr;
    rc = dev_alloc(c, (void**)&d_cell, (size_t)std::max<int64_t>(need, 1) * 4);
    if (rc == LBM_OK) rc = dev_alloc(c, (void**)&d_dir, (size_t)std::max<int64_t>(need, 1));
    int64_t have = need;
    if (rc == LBM_OK) rc = build_list(c, d_cell, d_dir, &have);
    if (rc == LBM_OK) {
        if (cell && need) HIPKEEP(c, rc, hipMemcpyAsync(cell, d_cell, (size_t)need * 4, hipMemcpyDeviceToHost, c->stream));
        if (dir && need) HIPKEEP(c, rc, hipMemcpyAsync(dir, d_dir, (size_t)need, hipMemcpyDeviceToHost, c->stream));
        HIPKEEP(c, rc, hipStreamSynchronize(c->stream));
    }
    dev_free(c, d_cell, (size_t)std::max<int64_t>(need, 1) * 4);
    dev_free(c, d_dir, (size_t)std::max<int64_t>(need, 1));
    return rc;
}

int lbm_init_equilibrium(lbm_ctx* c, const double* rho, const double* u) {
    if (!c) return LBM_ERR_ARG;
    const int64_t N = c->g.N;
    double* d_rho = nullptr;
    double* d_u = nullptr;
    int rc = LBM_OK;
    if (rho) rc = dev_alloc(c, (void**)&d_rho, (size_t)N * 8);
    if (rc == LBM_OK && u) rc = dev_alloc(c, (void**)&d_u, (size_t)3 * N * 8);
    if (rc == LBM_OK && rho) HIPKEEP(c, rc, hipMemcpyAsync(d_rho, rho, (size_t)N * 8, hipMemcpyHostToDevice, c->stream));
    if (rc == LBM_OK && u) HIPKEEP(c, rc, hipMemcpyAsync(d_u, u, (size_t)3 * N * 8, hipMemcpyHostToDevice, c->stream));
    if (rc == LBM_OK && c->q == 27) {
        const int e = q27_launch_equilibrium(c->cfg.storage, c->g, c->stream, c->fbuf[1 - c->cur], d_rho, d_u,
                                             c->cfg.rho0, c->cfg.u0);
        if (e) rc = fail(c, LBM_ERR_HIP, std::string("equilibrium: ") + hipGetErrorString((hipError_t)e));
    } else if (rc == LBM_OK) {
        LaunchReq r = base_req(c, K_EQUILIBRIUM, 0);
        r.fout = c->fbuf[1 - c->cur];
        r.src_rho = d_rho;
        r.src_u = d_u;
        r.rho0 = c->cfg.rho0;
        for (int k = 0; k < 3; ++k) r.u0[k] = c->cfg.u0[k];
        rc = do_launch(c, r);
    }
    if (rc == LBM_OK) rc = collide_only(c);
    HIPKEEP(c, rc, hipStreamSynchronize(c->stream));
    dev_free(c, d_rho, rho ? (size_t)N * 8 : 0);
    dev_free(c, d_u, u ? (size_t)3 * N * 8 : 0);
    if (rc) return rc;
    c->step = 0;
    c->initialised = true;
    c->raw_state = true;
    c->last_rec_valid = false;
    return LBM_OK;
}

int lbm_set_populations(lbm_ctx* c, const double* f) {
    if (!c || !f) return LBM_ERR_ARG;
    const int64_t N = c->g.N;
    double* d = nullptr;
    const size_t bytes = (size_t)c->q * N * 8;
    int rc = dev_alloc(c, (void**)&d, bytes);
    if (rc) return rc;
    HIPKEEP(c, rc, hipMemcpyAsync(d, f, bytes, hipMemcpyHostToDevice, c->stream));
    if (rc == LBM_OK && c->q == 27) {
        const int e = q27_launch_from_double(c->cfg.storage, c->g, c->stream, c->fbuf[1 - c->cur], d);
        if (e) rc = fail(c, LBM_ERR_HIP, std::string("populations: ") + hipGetErrorString((hipError_t)e));
    } else if (rc == LBM_OK) {
        LaunchReq r = base_req(c, K_FROM_DOUBLE, 0);
        r.fout = c->fbuf[1 - c->cur];
        r.src_f = d;
        rc = do_launch(c, r);
    }
    if (rc == LBM_OK) rc = collide_only(c);
    HIPKEEP(c, rc, hipStreamSynchronize(c->stream));
    dev_free(c, d, bytes);
    if (rc) return rc;
    c->step = 0;
    c->initialised = true;
    c->raw_state = true;
    c->last_rec_valid = false;
    return LBM_OK;
}

int lbm_step(lbm_ctx* c, int32_t nsteps, double* force_out) {
    if (!c || nsteps < 0) return LBM_ERR_ARG;
    if (!c->initialised) return fail(c, LBM_ERR_STATE, "lbm_step before lbm_init_equilibrium");
    int64_t done_total = 0;
    const int n6 = c->nb * 6;
    int64_t chunk_off[2] = {-1, -1};  // first step (within this call) of the chunk held by each parity
    int chunk_cnt[2] = {0, 0};
    int64_t bad_step = -1;  // first step whose force/torque record is not finite
    int bad_body = -1;
    // hands a finished chunk's records to the caller and checks them
    auto take = [&](int par) {
        const double* r = c->h_forces[par];
        for (int64_t i = 0; i < (int64_t)chunk_cnt[par] * n6 && bad_step < 0; ++i)
            if (!std::isfinite(r[i])) {
                bad_step = c->step - done_total + chunk_off[par] + i / n6 + 1;
                bad_body = (int)((i % n6) / 6);
            }
        if (force_out)
            std::memcpy(force_out + (size_t)chunk_off[par] * n6, r, (size_t)chunk_cnt[par] * n6 * sizeof(double));
        chunk_off[par] = -1;
    };
    int rc = [&]() -> int {
        while (done_total < nsteps) {
            const int par = c->parity;
            if (chunk_off[par] >= 0) {
                HIPCHK(c, hipEventSynchronize(c->rec_ev[par]));
                if (n6 > 0) take(par);
                chunk_off[par] = -1;
                if (bad_step >= 0) break;  // the state is no longer finite
            }
            const int k = (int)std::min<int64_t>(kChunk, nsteps - done_total);
            int64_t done = 0;
            int r = run_chunk(c, k, force_out, &done);
            if (r) return r;
            chunk_off[par] = done_total;
            chunk_cnt[par] = (int)done;
            done_total += done;
        }
        return (int)LBM_OK;
    }();
    if (rc) return rc;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (c->profile_stride > 0) {
        // sampled event pairs around the step kernel
        for (size_t i = 0; i + 2 <= c->ev_used; i += 2) {
            float a = 0.f;
            HIPCHK(c, hipEventElapsedTime(&a, c->ev_pool[i], c->ev_pool[i + 1]));
            c->bulk_ms += a;
            ++c->bulk_launches;
        }
        c->ev_used = 0;
    }
    for (int par = 0; par < 2; ++par)
        if (chunk_off[par] >= 0 && n6 > 0) take(par);
    if (bad_step >= 0)
        return fail(c, LBM_ERR_NUMERIC, "non-finite force or torque of body " + std::to_string(bad_body) +
                                            " at step " + std::to_string(bad_step));
    return LBM_OK;
}

int lbm_get_macros(lbm_ctx* c, double* rho, double* u) {
    if (!c) return LBM_ERR_ARG;
    const int64_t N = c->g.N;
    double *d_rho = nullptr, *d_u = nullptr;
    int rc = dev_alloc(c, (void**)&d_rho, (size_t)N * 8);
    if (rc == LBM_OK) rc = dev_alloc(c, (void**)&d_u, (size_t)3 * N * 8);
    if (rc == LBM_OK) rc = get_state(c, nullptr, d_rho, d_u);
    if (rc == LBM_OK) {
        if (rho) HIPKEEP(c, rc, hipMemcpyAsync(rho, d_rho, (size_t)N * 8, hipMemcpyDeviceToHost, c->stream));
        if (u) HIPKEEP(c, rc, hipMemcpyAsync(u, d_u, (size_t)3 * N * 8, hipMemcpyDeviceToHost, c->stream));
        HIPKEEP(c, rc, hipStreamSynchronize(c->stream));
    }
    dev_free(c, d_rho, (size_t)N * 8);
    dev_free(c, d_u, (size_t)3 * N * 8);
    return rc;
}

int lbm_get_populations(lbm_ctx* c, double* f) {
    if (!c || !f) return LBM_ERR_ARG;
    const int64_t N = c->g.N;
    double* d = nullptr;
    const size_t bytes = (size_t)c->q * N * 8;
    int rc = dev_alloc(c, (void**)&d, bytes);
    if (rc == LBM_OK) rc = get_state(c, d, nullptr, nullptr);
    if (rc == LBM_OK) {
        HIPKEEP(c, rc, hipMemcpyAsync(f, d, bytes, hipMemcpyDeviceToHost, c->stream));
        HIPKEEP(c, rc, hipStreamSynchronize(c->stream));
    }
    dev_free(c, d, bytes);
    return rc;
}

int lbm_conservation(lbm_ctx* c, double* mass, double* mom3) {
    if (!c) return LBM_ERR_ARG;
    const int64_t N = c->g.N;
    const int64_t nblk = nblocks_of(N);
    double *d_rho = nullptr, *d_u = nullptr, *d_p = nullptr;
    int rc = dev_alloc(c, (void**)&d_rho, (size_t)N * 8);
    if (rc == LBM_OK) rc = dev_alloc(c, (void**)&d_u, (size_t)3 * N * 8);
    if (rc == LBM_OK) rc = dev_alloc(c, (void**)&d_p, (size_t)nblk * 4 * 8);
    if (rc == LBM_OK) rc = get_state(c, nullptr, d_rho, d_u);
    std::vector<double> h((size_t)nblk * 4);
    if (rc == LBM_OK) {
        HIPKEEP(c, rc, (hipError_t)launch_conservation(d_rho, d_u, c->flags, N, d_p, c->stream));
        if (rc == LBM_OK) HIPKEEP(c, rc, hipMemcpyAsync(h.data(), d_p, h.size() * 8, hipMemcpyDeviceToHost, c->stream));
        HIPKEEP(c, rc, hipStreamSynchronize(c->stream));
    }
    if (rc == LBM_OK) {
        double s[4] = {0, 0, 0, 0};
        for (int64_t b = 0; b < nblk; ++b)
            for (int k = 0; k < 4; ++k) s[k] += h[(size_t)b * 4 + k];
        if (mass) *mass = s[0];
        if (mom3) {
            mom3[0] = s[1];
            mom3[1] = s[2];
            mom3[2] = s[3];
        }
    }
    dev_free(c, d_rho, (size_t)N * 8);
    dev_free(c, d_u, (size_t)3 * N * 8);
    dev_free(c, d_p, (size_t)nblk * 4 * 8);
    return rc;
}

int lbm_conservation_static(lbm_ctx* c, double* mass, double* mom3) {
    if (!c) return LBM_ERR_ARG;
    if (!c->initialised) return fail(c, LBM_ERR_STATE, "not initialised");
    double* sp = scratch_of(c->device);
    if (!sp) return fail(c, LBM_ERR_STATE, "no scratch address for the context's device");
    int cur_dev = -1;
    HIPCHK(c, hipGetDevice(&cur_dev));
    if (cur_dev != c->device) return fail(c, LBM_ERR_STATE, "lbm_conservation_static called off the context's device");
    ConsDirs t{};
    for (int d = 0; d < c->q; ++d) {
        t.cx[d] = (int8_t)(c->q == 27 ? q27::kX[d] : kCx[d]);
        t.cy[d] = (int8_t)(c->q == 27 ? q27::kY[d] : kCy[d]);
        t.cz[d] = (int8_t)(c->q == 27 ? q27::kZ[d] : kCz[d]);
        const int n = (t.cx[d] != 0) + (t.cy[d] != 0) + (t.cz[d] != 0);
        // the weights exactly as the storage formats' loads form them, in FP32
        if (c->q == 27)
            t.w[d] = n == 0 ? 8.0f / 27.0f : (n == 1 ? 2.0f / 27.0f : (n == 2 ? 1.0f / 54.0f : 1.0f / 216.0f));
        else
            t.w[d] = weight<float>(d);
    }
    const void* f = c->fbuf[c->cur];
    const dim3 grid(kConsBlocks), block(kConsThreads);
    switch (c->cfg.storage) {
        case LBM_FP64:
            hipLaunchKernelGGL(k_cons_partials<LBM_FP64>, grid, block, 0, c->stream, f, c->g.P, c->flags, c->g.N, c->q,
                               t, c->g.sh.ik, sp);
            break;
        case LBM_FP32:
            hipLaunchKernelGGL(k_cons_partials<LBM_FP32>, grid, block, 0, c->stream, f, c->g.P, c->flags, c->g.N, c->q,
                               t, c->g.sh.ik, sp);
            break;
        case LBM_FP16S:
            hipLaunchKernelGGL(k_cons_partials<LBM_FP16S>, grid, block, 0, c->stream, f, c->g.P, c->flags, c->g.N,
                               c->q, t, c->g.sh.ik, sp);
            break;
        default:
            hipLaunchKernelGGL(k_cons_partials<LBM_BF16S>, grid, block, 0, c->stream, f, c->g.P, c->flags, c->g.N,
                               c->q, t, c->g.sh.ik, sp);
            break;
    }
    HIPCHK(c, hipGetLastError());
    double* res = sp + (size_t)kConsBlocks * 4;
    hipLaunchKernelGGL(k_cons_final, dim3(1), dim3(64), 0, c->stream, sp, kConsBlocks, c->pd.g[0], c->pd.g[1],
                       c->pd.g[2], res);
    HIPCHK(c, hipGetLastError());
    double h[4];
    HIPCHK(c, hipMemcpyAsync(h, res, sizeof(h), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (mass) *mass = h[0];
    if (mom3) {
        mom3[0] = h[1];
        mom3[1] = h[2];
        mom3[2] = h[3];
    }
    return LBM_OK;
}

int lbm_info(lbm_ctx* c, int64_t* ncells, int64_t* nbb, int64_t* nlinks, int64_t* steps, int64_t* device_bytes) {
    if (!c) return LBM_ERR_ARG;
    if (ncells) *ncells = c->g.N;
    if (nbb) *nbb = c->list_valid ? c->list_n : -1;
    if (nlinks) *nlinks = c->list_valid ? c->nlinks : -1;
    if (steps) *steps = c->step;
    if (device_bytes) *device_bytes = c->device_bytes;
    return LBM_OK;
}

int lbm_profile(lbm_ctx* c, int32_t enable) {
    if (!c) return LBM_ERR_ARG;
    c->profile_stride = enable > 0 ? enable : 0;
    c->bulk_ms = c->link_ms = 0.0;
    c->bulk_launches = c->link_launches = 0;
    return LBM_OK;
}

int lbm_profile_read(lbm_ctx* c, double* bulk_ms, int64_t* bulk_launches, double* link_ms, int64_t* link_launches) {
    if (!c) return LBM_ERR_ARG;
    if (bulk_ms) *bulk_ms = c->bulk_ms;
    if (bulk_launches) *bulk_launches = c->bulk_launches;
    if (link_ms) *link_ms = c->link_ms;
    if (link_launches) *link_launches = c->link_launches;
    return LBM_OK;
}

int lbm_sync(lbm_ctx* c) {
    if (!c) return LBM_ERR_ARG;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return LBM_OK;
}

}  // extern "C"
