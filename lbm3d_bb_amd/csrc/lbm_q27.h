// lbm_q27.h -- the D3Q27 kernels (lbm_q27.hip) as the host code launches them.
#pragma once

#include "lbm_types.h"

namespace lbm {
namespace q27 {

constexpr int kQ = 27;
// ex, ey, ez of src/Utilities.jl:962-964: six faces, twelve edges, eight
// corners, rest last; moving directions pair as (2k, 2k + 1)
constexpr int kX[kQ] = {1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 0, 0, 1, -1, 1, -1, 0, 0, 1, -1, 1, -1, 1, -1, -1, 1, 0};
constexpr int kY[kQ] = {0, 0, 1, -1, 0, 0, 1, -1, 0, 0, 1, -1, -1, 1, 0, 0, 1, -1, 1, -1, 1, -1, -1, 1, 1, -1, 0};
constexpr int kZ[kQ] = {0, 0, 0, 0, 1, -1, 0, 0, 1, -1, 1, -1, 0, 0, -1, 1, -1, 1, 1, -1, -1, 1, 1, -1, 1, -1, 0};

// Cells of a split 16-bit step that the per-cell blocks take besides the rows
// on non-periodic y/z faces: per body, a box around its cells within reach in
// the old and the new pose (host-computed from the step's bodies, a cell wider
// on each side than the device's near test, x bounds on even / odd cells so
// that no pair of the pair blocks is cut).  y0/z0 in [0, L) (they wrap on
// periodic axes); x0 + ex <= nx (a box that would wrap in x is two boxes);
// cell (x, y, z) lies in box k when x - x0[k] < ex[k], (y - y0[k]) mod ny < ey[k]
// and (z - z0[k]) mod nz < ez[k].
// The per-cell blocks of a split step are enumerated from these values alone:
// first the rows of each box (bpx[k] slots per row: the blocks its x range
// can span; base[k] is the first slot of box k, base[n] the boxes' total),
// then the rows on the non-periodic y/z faces (nface rows, bpr block slots per
// row: the blocks a row of nx cells can span); total counts all slots.  A block reached
// twice (rows that share a block, overlapping boxes) computes the same values
// twice, which is harmless.
constexpr int kSplitBoxes = 32;
struct SplitRows {
    int n;
    int y0[kSplitBoxes], ey[kSplitBoxes], z0[kSplitBoxes], ez[kSplitBoxes];
    int x0[kSplitBoxes], ex[kSplitBoxes], bpx[kSplitBoxes];
    int bpr;
    int nface;
    int base[kSplitBoxes + 1];
    int64_t total;
};

// One launch of k27_step: a time step (MODE_STEP), the getters' replay of the
// last step (MODE_GET) or of the start state (MODE_GET_RAW), or the collision
// of the start state (MODE_COLLIDE_ONLY).
struct StepArgs {
    int mode;
    int storage;  // lbm_storage
    int trt;
    int forced;
    int ibb;  // links: 1 interpolated bounce-back, 0 simple
    Grid g;
    double wp, wm;
    double grav[3];
    hipStream_t stream;
    const void* fin;
    void* fout;
    double* gout;     // MODE_GET*: f_old, 27 planes of N
    double* rho_out;  // MODE_GET*: N
    double* u_out;    // MODE_GET*: 3 x N
    const uint8_t* fl_new;
    const uint8_t* fl_old;  // flags before this step's pose change (== fl_new when the pose did not change)
    const DevBody* bnew;    // nb bodies of the step's pose
    const DevBody* bold;    // nb bodies of the previous pose, or null when the pose did not change
    int nb;
    double* partials;  // MODE_STEP: [blocks][nb][6], written by the blocks within reach of a body
    int refill;
    const double* face_u;  // [6][3]
    // 16-bit storage, even nx >= 4: MODE_STEP runs k27_split (pair blocks over
    // the plain rows, per-cell blocks over the rest, `rows`); every mode that
    // stores f* writes the x-face slots
    int split;
    SplitRows rows;
};

}  // namespace q27

int q27_launch_step(const q27::StepArgs& a);
int q27_launch_step_ibb(const q27::StepArgs& a);  // lbm_q27_ibb.hip
// 16-bit storage (shifted FP16 / BF16 with the D3Q27 weights): lbm_q27_16.hip, lbm_q27_16_ibb.hip
int q27_launch_step16(const q27::StepArgs& a);
int q27_launch_step16_ibb(const q27::StepArgs& a);
int q27_launch_equilibrium16(int storage, const Grid& g, hipStream_t st, void* f, const double* rho, const double* u,
                             double rho0, const double u0[3]);
int q27_launch_from_double16(int storage, const Grid& g, hipStream_t st, void* f, const double* src);
// out[s][k] = sum over the blocks b within reach of a body of
// partials[s * stride + b * n6 + k], s < nsteps; recs + s * rec_bytes holds
// step s's bodies (nb of the new pose, then nb of the old one)
int q27_launch_force_sum(const double* partials, int64_t nblocks, int64_t stride, int n6, int nsteps,
                         const char* recs, size_t rec_bytes, const Grid& g, double* out, hipStream_t st);
int q27_launch_equilibrium(int storage, const Grid& g, hipStream_t st, void* f, const double* rho, const double* u,
                           double rho0, const double u0[3]);
int q27_launch_from_double(int storage, const Grid& g, hipStream_t st, void* f, const double* src);
// flags[cell] = solid flag of the bodies for every cell of the boxes
// boxes[b] = {lo x, lo y, lo z, ext x, ext y, ext z} (lo wraps on periodic axes), b < nb
int q27_launch_flags_boxes(uint8_t* flags, const DevBody* bodies, int nb, const Grid& g, const int32_t* boxes,
                           int max_volume, hipStream_t st);
int q27_launch_link_count(const uint8_t* flags, const Grid& g, int32_t* counts, hipStream_t st);
int q27_launch_link_fill(const uint8_t* flags, const Grid& g, const int32_t* block_off, int32_t* cells, int8_t* dirs,
                         hipStream_t st);

}  // namespace lbm
