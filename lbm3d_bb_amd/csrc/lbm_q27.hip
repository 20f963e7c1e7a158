// lbm_q27.hip -- the D3Q27 lattice: the step with simple bounce-back, the
// force sum, start state, moved-body flags and link list kernels, and the
// launchers the host code calls.
#include "lbm_q27_kernel.h"

namespace lbm {
namespace q27 {

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

}  // namespace q27

int q27_launch_step(const q27::StepArgs& a) {
    if (a.storage == LBM_FP16S || a.storage == LBM_BF16S) return a.ibb ? q27_launch_step16_ibb(a) : q27_launch_step16(a);
    if (a.ibb) return q27_launch_step_ibb(a);
    if (a.storage == LBM_FP64) q27::launch_storage<StoreF64, false>(a);
    else q27::launch_storage<StoreF32, false>(a);
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
    if (storage == LBM_FP16S || storage == LBM_BF16S) return q27_launch_equilibrium16(storage, g, st, f, rho, u, rho0, u0);
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
    if (storage == LBM_FP16S || storage == LBM_BF16S) return q27_launch_from_double16(storage, g, st, f, src);
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
