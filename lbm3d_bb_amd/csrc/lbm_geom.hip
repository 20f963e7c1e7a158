// lbm_geom.hip -- storage-independent kernels: solid flags, the ordered
// boundary-node / link compaction (compile_bb_data!, src/Geometry.jl:268-311),
// force finalisation and the FP64 conservation sums.
#include "lbm_device.h"

namespace lbm {

// Sums the block partials of each step of a chunk: one block per (component,
// step), a strided pass over the blocks then a fixed-shape tree, so the result
// is deterministic.  out[s][i] = sum_{b < grid(s)} partials[s * step_stride + b * n6 + i],
// with grid(s) read from the step's record.
__global__ __launch_bounds__(kBlock) void k_force_finalize(const double* __restrict__ partials, int64_t step_stride,
                                                           const char* __restrict__ recs, size_t rec_bytes,
                                                           int n6, double* __restrict__ out) {
    __shared__ double s_sum[kBlock];
    const int i = blockIdx.x;
    const int s = blockIdx.y;
    const StepGeom* sg = (const StepGeom*)(recs + (size_t)s * rec_bytes);
    const int nblocks = sg->nfix_blocks + sg->list_blocks;
    const double* p = partials + (int64_t)s * step_stride;
    double acc = 0.0;
    for (int b = threadIdx.x; b < nblocks; b += kBlock) acc += p[(int64_t)b * n6 + i];
    s_sum[threadIdx.x] = acc;
    __syncthreads();
    for (int off = kBlock / 2; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) s_sum[threadIdx.x] += s_sum[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[(int64_t)s * n6 + i] = s_sum[0];
}

// ---------------------------------------------------------------------------
// Geometry: full-domain flags, then an ordered compaction of the fluid cells
// that have a solid neighbour (count -> scan -> fill, the structure of
// compile_bb_data!, src/Geometry.jl:268-311).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void k_flags_full(uint8_t* __restrict__ flags, const DevBody* __restrict__ bodies,
                                                       int nb, Grid g) {
    const int64_t cell = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (cell >= g.N) return;
    int x, y, z;
    cell_xyz(cell, g, x, y, z);
    flags[cell] = (uint8_t)flag_of(bodies, nb, x, y, z, g);
}

__device__ inline bool is_bb_node(const uint8_t* __restrict__ flags, int64_t cell, const Grid& g) {
    if ((flags[cell] & kFlagBody) != 0) return false;
    int x, y, z;
    cell_xyz(cell, g, x, y, z);
    const int L[3] = {g.nx, g.ny, g.nz};
    for (int d = 0; d < 18; ++d) {
        int q[3] = {x + kCx[d], y + kCy[d], z + kCz[d]};
        bool ok = true;
        for (int a = 0; a < 3; ++a) ok = ok && wrap_coord(q[a], L[a], (g.per >> a) & 1);
        if (ok && (flags[(int64_t)q[0] + (int64_t)g.nx * q[1] + g.nxy * q[2]] & kFlagBody) != 0) return true;
    }
    return false;
}

__device__ inline int link_count(const uint8_t* __restrict__ flags, int64_t cell, const Grid& g) {
    if ((flags[cell] & kFlagBody) != 0) return 0;
    int x, y, z;
    cell_xyz(cell, g, x, y, z);
    const int L[3] = {g.nx, g.ny, g.nz};
    int n = 0;
    for (int d = 0; d < 18; ++d) {
        int q[3] = {x + kCx[d], y + kCy[d], z + kCz[d]};
        bool ok = true;
        for (int a = 0; a < 3; ++a) ok = ok && wrap_coord(q[a], L[a], (g.per >> a) & 1);
        if (ok && (flags[(int64_t)q[0] + (int64_t)g.nx * q[1] + g.nxy * q[2]] & kFlagBody) != 0) ++n;
    }
    return n;
}

// Per-block counts of boundary nodes and of links.
__global__ __launch_bounds__(kBlock) void k_bb_count(const uint8_t* __restrict__ flags, Grid g,
                                                     int32_t* __restrict__ node_counts,
                                                     int32_t* __restrict__ link_counts) {
    __shared__ int s_n[kWaves], s_l[kWaves];
    const int64_t cell = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    int nl = 0;
    if (cell < g.N) nl = link_count(flags, cell, g);
    const uint64_t bal = __ballot(nl > 0);
    int wl = nl;
    for (int off = 32; off > 0; off >>= 1) wl += __shfl_xor(wl, off, 64);
    if ((threadIdx.x & 63) == 0) {
        s_n[threadIdx.x >> 6] = __popcll(bal);
        s_l[threadIdx.x >> 6] = wl;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        int a = 0, b = 0;
        for (int w = 0; w < kWaves; ++w) {
            a += s_n[w];
            b += s_l[w];
        }
        node_counts[blockIdx.x] = a;
        link_counts[blockIdx.x] = b;
    }
}

// Exclusive scan of both count arrays by one block; totals written at [n].
__global__ __launch_bounds__(1024) void k_scan2(int32_t* __restrict__ a, int32_t* __restrict__ b, int n) {
    __shared__ int64_t s_a[1024], s_b[1024];
    __shared__ int64_t carry_a, carry_b;
    if (threadIdx.x == 0) carry_a = carry_b = 0;
    __syncthreads();
    for (int base = 0; base < n; base += 1024) {
        const int i = base + threadIdx.x;
        const int64_t va = i < n ? a[i] : 0, vb = i < n ? b[i] : 0;
        s_a[threadIdx.x] = va;
        s_b[threadIdx.x] = vb;
        __syncthreads();
        for (int off = 1; off < 1024; off <<= 1) {
            int64_t ta = 0, tb = 0;
            if ((int)threadIdx.x >= off) {
                ta = s_a[threadIdx.x - off];
                tb = s_b[threadIdx.x - off];
            }
            __syncthreads();
            s_a[threadIdx.x] += ta;
            s_b[threadIdx.x] += tb;
            __syncthreads();
        }
        if (i < n) {
            a[i] = (int32_t)(carry_a + s_a[threadIdx.x] - va);
            b[i] = (int32_t)(carry_b + s_b[threadIdx.x] - vb);
        }
        __syncthreads();
        if (threadIdx.x == 1023) {
            carry_a += s_a[1023];
            carry_b += s_b[1023];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        a[n] = (int32_t)carry_a;
        b[n] = (int32_t)carry_b;
    }
}

// Fills the ordered node list and, optionally, the (cell, dir) link list in
// ascending cell then direction order (kernel_fill_bb_csr!, src/Geometry.jl:249-257).
__global__ __launch_bounds__(kBlock) void k_bb_fill(const uint8_t* __restrict__ flags, Grid g,
                                                    const int32_t* __restrict__ node_off,
                                                    const int32_t* __restrict__ link_off, int32_t* __restrict__ nodes,
                                                    int32_t* __restrict__ link_cell, int8_t* __restrict__ link_dir) {
    __shared__ int s_n[kWaves], s_l[kWaves];
    const int64_t cell = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    int nl = 0;
    if (cell < g.N) nl = link_count(flags, cell, g);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const uint64_t bal = __ballot(nl > 0);
    const int nbefore = __popcll(bal & ((1ull << lane) - 1ull));
    // inclusive scan of nl within the wave
    int incl = nl;
    for (int off = 1; off < 64; off <<= 1) {
        int t = __shfl_up(incl, off, 64);
        if (lane >= off) incl += t;
    }
    if (lane == 63) {
        s_n[wave] = __popcll(bal);
        s_l[wave] = incl;
    }
    __syncthreads();
    int wn = 0, wl = 0;
    for (int w = 0; w < wave; ++w) {
        wn += s_n[w];
        wl += s_l[w];
    }
    if (nl > 0) {
        nodes[node_off[blockIdx.x] + wn + nbefore] = (int32_t)cell;
        if (link_cell) {
            int pos = link_off[blockIdx.x] + wl + incl - nl;
            int x, y, z;
            cell_xyz(cell, g, x, y, z);
            const int L[3] = {g.nx, g.ny, g.nz};
            for (int d = 0; d < 18; ++d) {
                int q[3] = {x + kCx[d], y + kCy[d], z + kCz[d]};
                bool ok = true;
                for (int a = 0; a < 3; ++a) ok = ok && wrap_coord(q[a], L[a], (g.per >> a) & 1);
                if (ok && (flags[(int64_t)q[0] + (int64_t)g.nx * q[1] + g.nxy * q[2]] & kFlagBody) != 0) {
                    link_cell[pos] = (int32_t)cell;
                    link_dir[pos] = (int8_t)d;
                    ++pos;
                }
            }
        }
    }
}

// FP64 sums of rho and rho*u over fluid cells: per-block partials [nblocks][4].
__global__ __launch_bounds__(kBlock) void k_conservation(const double* __restrict__ rho, const double* __restrict__ u,
                                                         const uint8_t* __restrict__ flags, int64_t N,
                                                         double* __restrict__ partials) {
    __shared__ double s[kWaves][4];
    const int64_t cell = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    double v[4] = {0, 0, 0, 0};
    if (cell < N && (flags[cell] & kFlagBody) == 0) {
        const double r = rho[cell];
        v[0] = r;
        v[1] = r * u[cell];
        v[2] = r * u[N + cell];
        v[3] = r * u[2 * N + cell];
    }
    for (int c = 0; c < 4; ++c) v[c] = wave_sum(v[c]);
    if ((threadIdx.x & 63) == 0)
        for (int c = 0; c < 4; ++c) s[threadIdx.x >> 6][c] = v[c];
    __syncthreads();
    if (threadIdx.x < 4) {
        double a = 0.0;
        for (int w = 0; w < kWaves; ++w) a += s[w][threadIdx.x];
        partials[(int64_t)blockIdx.x * 4 + threadIdx.x] = a;
    }
}


// Sets (set = 1) the fix-up bit on the listed cells, or clears it everywhere (set = 0).
__global__ __launch_bounds__(kBlock) void k_mark_list(uint8_t* __restrict__ flags, const int32_t* __restrict__ list,
                                                      int64_t n, int set, int64_t N) {
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (set) {
        if (i < n) flags[list[i]] |= kFlagFixup;
    } else if (i < N) {
        flags[i] &= kFlagBody;
    }
}

#define LAUNCH_CHECK()                          \
    do {                                        \
        hipError_t e_ = hipGetLastError();      \
        return e_ == hipSuccess ? 0 : (int)e_;  \
    } while (0)

int launch_flags_full(uint8_t* flags, const DevBody* bodies, int nb, const Grid& g, hipStream_t st) {
    hipLaunchKernelGGL(k_flags_full, dim3((unsigned)nblocks_of(g.N)), dim3(kBlock), 0, st, flags, bodies, nb, g);
    LAUNCH_CHECK();
}

int launch_bb_count(const uint8_t* flags, const Grid& g, int32_t* node_counts, int32_t* link_counts, hipStream_t st) {
    hipLaunchKernelGGL(k_bb_count, dim3((unsigned)nblocks_of(g.N)), dim3(kBlock), 0, st, flags, g, node_counts,
                       link_counts);
    LAUNCH_CHECK();
}

int launch_scan2(int32_t* a, int32_t* b, int n, hipStream_t st) {
    hipLaunchKernelGGL(k_scan2, dim3(1), dim3(1024), 0, st, a, b, n);
    LAUNCH_CHECK();
}

int launch_bb_fill(const uint8_t* flags, const Grid& g, const int32_t* node_off, const int32_t* link_off,
                   int32_t* nodes, int32_t* link_cell, int8_t* link_dir, hipStream_t st) {
    hipLaunchKernelGGL(k_bb_fill, dim3((unsigned)nblocks_of(g.N)), dim3(kBlock), 0, st, flags, g, node_off, link_off,
                       nodes, link_cell, link_dir);
    LAUNCH_CHECK();
}

int launch_force_finalize(const double* partials, int64_t step_stride, const char* recs, size_t rec_bytes,
                          int nsteps, int n6, double* out, hipStream_t st) {
    hipLaunchKernelGGL(k_force_finalize, dim3((unsigned)n6, (unsigned)nsteps), dim3(kBlock), 0, st,
                       partials, step_stride, recs, rec_bytes, n6, out);
    LAUNCH_CHECK();
}

int launch_mark_list(uint8_t* flags, const int32_t* list, int64_t n, int set, int64_t N, hipStream_t st) {
    const int64_t m = set ? n : N;
    if (m <= 0) return 0;
    hipLaunchKernelGGL(k_mark_list, dim3((unsigned)nblocks_of(m)), dim3(kBlock), 0, st, flags, list, n, set, N);
    LAUNCH_CHECK();
}

int launch_conservation(const double* rho, const double* u, const uint8_t* flags, int64_t N, double* partials,
                        hipStream_t st) {
    hipLaunchKernelGGL(k_conservation, dim3((unsigned)nblocks_of(N)), dim3(kBlock), 0, st, rho, u, flags, N,
                       partials);
    LAUNCH_CHECK();
}

}  // namespace lbm
