// unet_kernels.h -- gfx950 kernels of the warm-start U-Net.
//
// Activations are BF16, channels last: element (voxel v, channel c) of a tensor
// with channel stride cs and channel offset co sits at p[v * cs + co + c];
// v = ((n * D + z) * H + y) * W + x.  Convolutions are implicit GEMMs on
// v_mfma_f32_16x16x32_bf16: rows = voxels, columns = output channels, the
// reduction runs over (tap, input channel) in groups of 8 channels so that one
// lane loads 16 contiguous bytes of one voxel per group.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace unet {

typedef __bf16 bf16;
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int kThreads = 256;

// Geometry of a convolution: which input voxel feeds output row r through tap t.
enum { G_CONV3 = 0, G_DOWN2 = 1, G_UP2 = 2, G_PW = 3 };

struct Dims {
    int n, d, h, w;
};

__device__ inline int64_t vox(const Dims& s, int n, int z, int y, int x) {
    return (((int64_t)n * s.d + z) * s.h + y) * s.w + x;
}

// Voxel v of a tensor as (n, z, y, x).  Every tensor of a network holds
// fewer than 2^31 voxels (checked when the network is created), so the
// divisions are 32-bit ones: a 64-bit division has no hardware instruction.
__device__ inline void unvox(const Dims& s, int64_t v, int& n, int& z, int& y, int& x) {
    const uint32_t u = (uint32_t)v;
    const uint32_t t1 = u / (uint32_t)s.w;
    x = (int)(u - t1 * (uint32_t)s.w);
    const uint32_t t2 = t1 / (uint32_t)s.h;
    y = (int)(t1 - t2 * (uint32_t)s.h);
    const uint32_t t3 = t2 / (uint32_t)s.d;
    z = (int)(t2 - t3 * (uint32_t)s.d);
    n = (int)t3;
}

struct ConvArgs {
    const bf16* x;
    int xcs, xco, cin;      // input: channel stride, offset, channels (multiple of 8)
    const bf16* wt;         // [npar][coutp][kpad]
    int coutp, kpad, ntaps;
    const float* bias;      // [coutp] or null
    bf16* y;                // BF16 output or null
    int ycs, yco, cstore;   // output channel stride, offset, channels stored
    float* yf;              // FP32 output (channel stride cstore) or null
    const bf16* aux;        // if set: output *= (aux > 0) (ReLU backward)
    int acs, aco;
    int relu;
    Dims rows;              // dims the rows enumerate (output; coarse input for UP2)
    Dims in;                // input dims
    Dims out;               // output dims
    int64_t m;              // number of rows
    int wrap;               // 3x3x3 taps wrap around along x (bit 0), y (1), z (2) instead of reading zeros
};

// Coordinate c + o of an axis of size n under the padding rule: wrapped when
// the axis is periodic, -1 (zero padding) when it falls outside otherwise.
__device__ inline int pad_coord(int c, int n, bool wrap) {
    if (c >= 0 && c < n) return c;
    if (!wrap) return -1;
    return c < 0 ? c + n : c - n;
}

// Input voxel of row (n,z,y,x) through tap t; false for zero padding.
template <int GEOM>
__device__ inline bool in_voxel(const ConvArgs& a, int n, int z, int y, int x, int t, int64_t& v) {
    if (GEOM == G_CONV3) {
        const int zz = pad_coord(z + t / 9 - 1, a.in.d, a.wrap & 4), yy = pad_coord(y + (t / 3) % 3 - 1, a.in.h, a.wrap & 2),
                  xx = pad_coord(x + t % 3 - 1, a.in.w, a.wrap & 1);
        if (zz < 0 || yy < 0 || xx < 0) return false;
        v = vox(a.in, n, zz, yy, xx);
    } else if (GEOM == G_DOWN2) {
        v = vox(a.in, n, 2 * z + (t >> 2), 2 * y + ((t >> 1) & 1), 2 * x + (t & 1));
    } else {
        v = vox(a.in, n, z, y, x);
    }
    return true;
}

__device__ inline bf16x8 load8(const bf16* p) { return *reinterpret_cast<const bf16x8*>(p); }

__device__ inline bf16x8 zero8() {
    bf16x8 z;
#pragma unroll
    for (int j = 0; j < 8; ++j) z[j] = (bf16)0.0f;
    return z;
}

// Input-voxel offset of tap t from the row's base voxel (see k_conv).
template <int GEOM>
__device__ inline int tap_offset(const Dims& in, int t) {
    if (GEOM == G_CONV3) return (t / 9 - 1) * in.h * in.w + ((t / 3) % 3 - 1) * in.w + (t % 3 - 1);
    if (GEOM == G_DOWN2) return (t >> 2) * in.h * in.w + ((t >> 1) & 1) * in.w + (t & 1);
    return 0;
}

// One wave computes 64 rows x 16*NT output channels; a block is 4 waves
// (256 rows).  grid = (ceil(m/256), coutp/(16*NT), npar).  Each row keeps its
// base input voxel and a bit mask of the taps that fall inside the input
// (the 27 taps of a same-size 3x3x3 convolution, whose rows and input share
// their dims; every tap otherwise), so the reduction loop does no division
// and no bounds test: tap t of a row reads voxel base + tap_offset(t) when bit
// t is set and zero padding when it is not.  Voxel indices are 32-bit (the
// host checks that every tensor has fewer than 2^31 voxels).
template <int GEOM, int NT>
__global__ __launch_bounds__(kThreads) void k_conv(ConvArgs a) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int g = lane >> 4, r16 = lane & 15;
    const int par = blockIdx.z;
    const int co_base = blockIdx.y * 16 * NT;
    const int64_t row0 = (int64_t)blockIdx.x * kThreads + wave * 64;
    const bf16* wt = a.wt + (int64_t)par * a.coutp * a.kpad;

    int vb[4];
    uint32_t tmask[4];
#pragma unroll
    for (int mt = 0; mt < 4; ++mt) {
        const int64_t r = row0 + mt * 16 + r16;
        const bool ok = r < a.m;
        int n, z, y, x;
        unvox(a.rows, ok ? r : 0, n, z, y, x);
        if (GEOM == G_CONV3) {
            vb[mt] = (int)vox(a.in, n, z, y, x);
            const uint32_t xb = (x > 0 ? 1u : 0u) | 2u | (x + 1 < a.in.w ? 4u : 0u);
            const uint32_t yb = (y > 0 ? 1u : 0u) | 2u | (y + 1 < a.in.h ? 4u : 0u);
            const uint32_t zb = (z > 0 ? 1u : 0u) | 2u | (z + 1 < a.in.d ? 4u : 0u);
            const uint32_t xy = ((yb & 1u) ? xb : 0u) | ((yb & 2u) ? xb << 3 : 0u) | ((yb & 4u) ? xb << 6 : 0u);
            const uint32_t m = ((zb & 1u) ? xy : 0u) | ((zb & 2u) ? xy << 9 : 0u) | ((zb & 4u) ? xy << 18 : 0u);
            tmask[mt] = ok ? m : 0u;
        } else if (GEOM == G_DOWN2) {
            vb[mt] = (int)vox(a.in, n, 2 * z, 2 * y, 2 * x);
            tmask[mt] = ok ? 0xffu : 0u;
        } else {
            vb[mt] = (int)vox(a.in, n, z, y, x);
            tmask[mt] = ok ? 1u : 0u;
        }
    }
    f32x4 acc[4][NT];
#pragma unroll
    for (int mt = 0; mt < 4; ++mt)
#pragma unroll
        for (int nt = 0; nt < NT; ++nt) acc[mt][nt] = f32x4{0.f, 0.f, 0.f, 0.f};

    // this lane's 8-channel group G = k0/8 + g as (tap t, group cg within the tap)
    const int gpt = a.cin >> 3;
    int t = g / gpt, cg = g - (g / gpt) * gpt;
    const bf16* xb = a.x + a.xco;
    for (int k0 = 0; k0 < a.kpad; k0 += 32) {
        bf16x8 A[4];
        if (t < a.ntaps) {
            const int off = tap_offset<GEOM>(a.in, t);
            const int c0 = cg * 8;
#pragma unroll
            for (int mt = 0; mt < 4; ++mt) {
                if ((tmask[mt] >> t) & 1u)
                    A[mt] = load8(xb + (int64_t)(vb[mt] + off) * a.xcs + c0);
                else
                    A[mt] = zero8();
            }
        } else {
#pragma unroll
            for (int mt = 0; mt < 4; ++mt) A[mt] = zero8();
        }
        bf16x8 B[NT];
#pragma unroll
        for (int nt = 0; nt < NT; ++nt) B[nt] = load8(wt + (int64_t)(co_base + nt * 16 + r16) * a.kpad + k0 + g * 8);
#pragma unroll
        for (int mt = 0; mt < 4; ++mt)
#pragma unroll
            for (int nt = 0; nt < NT; ++nt)
                acc[mt][nt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A[mt], B[nt], acc[mt][nt], 0, 0, 0);
        cg += 4;
        while (cg >= gpt) {
            cg -= gpt;
            ++t;
        }
    }
    // epilogue: lane holds rows 4g..4g+3 of column r16 of each tile
#pragma unroll
    for (int mt = 0; mt < 4; ++mt) {
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int64_t r = row0 + mt * 16 + 4 * g + i;
            if (r >= a.m) continue;
            // rows enumerate the output grid except for UP2, whose rows are coarse voxels
            int64_t ov = r;
            if (GEOM == G_UP2) {
                int n, z, y, x;
                unvox(a.rows, r, n, z, y, x);
                ov = vox(a.out, n, 2 * z + (par >> 2), 2 * y + ((par >> 1) & 1), 2 * x + (par & 1));
            }
#pragma unroll
            for (int nt = 0; nt < NT; ++nt) {
                const int co = co_base + nt * 16 + r16;
                if (co >= a.cstore) continue;
                float v = acc[mt][nt][i];
                if (a.bias) v += a.bias[co];
                if (a.relu) v = v > 0.f ? v : 0.f;
                if (a.aux && !((float)a.aux[ov * a.acs + a.aco + co] > 0.f)) v = 0.f;
                if (a.yf) a.yf[ov * a.cstore + co] = v;
                else a.y[ov * a.ycs + a.yco + co] = (bf16)v;
            }
        }
    }
}

// Transposed 2x2x2 stride-2 convolution with all 8 output parities in one
// block (k_conv takes one parity per block, blockIdx.z, and so reads every
// input row 8 times from memory).  Rows are coarse voxels; a wave's 64 rows
// load their KS 32-channel k-steps of input once into registers and feed the
// MFMAs of every parity par (weights wt[par]); output row r of parity par is
// fine voxel (2z + par/4, 2y + (par/2) % 2, 2x + par % 2) of the coarse voxel
// (z, y, x) of r.  The epilogue is k_conv's.
template <int NT, int KS>
__global__ __launch_bounds__(kThreads) void k_up2(ConvArgs a) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int g = lane >> 4, r16 = lane & 15;
    const int co_base = blockIdx.y * 16 * NT;
    const int64_t row0 = (int64_t)blockIdx.x * kThreads + wave * 64;
    bf16x8 A[KS][4];
#pragma unroll
    for (int mt = 0; mt < 4; ++mt) {
        const int64_t r = row0 + mt * 16 + r16;
        const bool ok = r < a.m;
        int n, z, y, x;
        unvox(a.rows, ok ? r : 0, n, z, y, x);
        const int64_t v = vox(a.in, n, z, y, x);
#pragma unroll
        for (int ks = 0; ks < KS; ++ks) {
            const int c0 = ks * 32 + g * 8;
            A[ks][mt] = (ok && c0 < a.cin) ? load8(a.x + v * a.xcs + a.xco + c0) : zero8();
        }
    }
    // the fine voxel of parity 0 of each output row this lane holds (rows 4g + i of each m-tile)
    int ovb[4][4];
#pragma unroll
    for (int mt = 0; mt < 4; ++mt)
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int64_t r = row0 + mt * 16 + 4 * g + i;
            int n, z, y, x;
            unvox(a.rows, r < a.m ? r : 0, n, z, y, x);
            ovb[mt][i] = (int)vox(a.out, n, 2 * z, 2 * y, 2 * x);
        }
    const int plane = a.out.h * a.out.w;
    for (int par = 0; par < 8; ++par) {
        const bf16* wt = a.wt + (int64_t)par * a.coutp * a.kpad;
        f32x4 acc[4][NT];
#pragma unroll
        for (int mt = 0; mt < 4; ++mt)
#pragma unroll
            for (int nt = 0; nt < NT; ++nt) acc[mt][nt] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int ks = 0; ks < KS; ++ks) {
            bf16x8 B[NT];
#pragma unroll
            for (int nt = 0; nt < NT; ++nt) B[nt] = load8(wt + (int64_t)(co_base + nt * 16 + r16) * a.kpad + ks * 32 + g * 8);
#pragma unroll
            for (int mt = 0; mt < 4; ++mt)
#pragma unroll
                for (int nt = 0; nt < NT; ++nt)
                    acc[mt][nt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A[ks][mt], B[nt], acc[mt][nt], 0, 0, 0);
        }
        const int poff = (par >> 2) * plane + ((par >> 1) & 1) * a.out.w + (par & 1);
#pragma unroll
        for (int mt = 0; mt < 4; ++mt) {
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int64_t r = row0 + mt * 16 + 4 * g + i;
                if (r >= a.m) continue;
                const int64_t ov = (int64_t)ovb[mt][i] + poff;
#pragma unroll
                for (int nt = 0; nt < NT; ++nt) {
                    const int co = co_base + nt * 16 + r16;
                    if (co >= a.cstore) continue;
                    float v = acc[mt][nt][i];
                    if (a.bias) v += a.bias[co];
                    if (a.relu) v = v > 0.f ? v : 0.f;
                    if (a.aux && !((float)a.aux[ov * a.acs + a.aco + co] > 0.f)) v = 0.f;
                    if (a.yf) a.yf[ov * a.cstore + co] = v;
                    else a.y[ov * a.ycs + a.yco + co] = (bf16)v;
                }
            }
        }
    }
}

// Pointwise (1x1x1) convolution of a narrow layer (cin <= 32 input channels,
// at most 16 stored outputs, used for the network's output head): one thread
// per voxel, its CIN input channels in 16-byte loads, the
// weights and bias as FP32 in LDS (every lane of a wave reads the same word),
// FP32 sums and k_conv's epilogue.  k_conv's MFMA tiles would pad the 4 head
// outputs to 16 columns and run 32-wide k-steps over 16 channels.  Rows, input
// and output share one grid (row r is voxel r of both).
template <int CIN>
__global__ __launch_bounds__(kThreads) void k_pw_narrow(ConvArgs a) {
    __shared__ float s_w[16][CIN];
    __shared__ float s_b[16];
    for (int i = threadIdx.x; i < 16 * CIN; i += kThreads) {
        const int co = i / CIN, ci = i - co * CIN;
        s_w[co][ci] = (co < a.cstore && ci < a.cin) ? (float)a.wt[(int64_t)co * a.kpad + ci] : 0.f;
    }
    if (threadIdx.x < 16) s_b[threadIdx.x] = (a.bias && (int)threadIdx.x < a.cstore) ? a.bias[threadIdx.x] : 0.f;
    __syncthreads();
    const int64_t r = (int64_t)blockIdx.x * kThreads + threadIdx.x;
    if (r >= a.m) return;
    float xv[CIN];
    const bf16* xp = a.x + r * a.xcs + a.xco;
#pragma unroll
    for (int c8 = 0; c8 < CIN / 8; ++c8) {
        const bf16x8 v = load8(xp + c8 * 8);
#pragma unroll
        for (int j = 0; j < 8; ++j) xv[c8 * 8 + j] = (float)v[j];
    }
    if (!a.yf && (a.cstore & 7) == 0 && (a.ycs & 7) == 0 && (a.yco & 7) == 0 &&
        (!a.aux || ((a.acs & 7) == 0 && (a.aco & 7) == 0))) {
        // BF16 outputs 8 channels at a time: the ReLU mask and the output in
        // 16-byte pieces (one voxel per lane: scalar 2-byte accesses would be
        // strided across the wave)
#pragma unroll
        for (int c8 = 0; c8 < 2; ++c8) {
            if (c8 * 8 < a.cstore) {
                const bf16x8 m8 = a.aux ? load8(a.aux + r * a.acs + a.aco + c8 * 8) : zero8();
                bf16x8 o;
#pragma unroll
                for (int j = 0; j < 8; ++j) {
                    const int co = c8 * 8 + j;
                    float v = 0.f;
#pragma unroll
                    for (int ci = 0; ci < CIN; ++ci) v = fmaf(s_w[co][ci], xv[ci], v);
                    v += s_b[co];
                    if (a.relu) v = v > 0.f ? v : 0.f;
                    if (a.aux && !((float)m8[j] > 0.f)) v = 0.f;
                    o[j] = (bf16)v;
                }
                *reinterpret_cast<bf16x8*>(a.y + r * a.ycs + a.yco + c8 * 8) = o;
            }
        }
        return;
    }
#pragma unroll
    for (int co = 0; co < 16; ++co) {
        if (co < a.cstore) {
            float v = 0.f;
#pragma unroll
            for (int ci = 0; ci < CIN; ++ci) v = fmaf(s_w[co][ci], xv[ci], v);
            v += s_b[co];
            if (a.relu) v = v > 0.f ? v : 0.f;
            if (a.aux && !((float)a.aux[r * a.acs + a.aco + co] > 0.f)) v = 0.f;
            if (a.yf) a.yf[r * a.cstore + co] = v;
            else a.y[r * a.ycs + a.yco + co] = (bf16)v;
        }
    }
}

// Same-size 3x3x3 convolution with the input staged in LDS.  A block computes
// a brick of 4 (z) x 4 (y) x 16 (x) output voxels of one sample, 256 rows, for
// 16*NT output channels: wave w takes the z-plane w of the brick, its m-tile mt
// the y-row mt, lane row r16 the voxel x0 + r16.  The input brick with its
// one-voxel halo (6 x 6 x 18 voxels) is loaded once per chunk of CK input
// channels (16-byte pieces, zeros outside the grid) and read by every tap from
// LDS: the A fragment of tap (dz, dy, dx) for row (w, mt, r16) is brick voxel
// (w + dz, mt + dy, r16 + dx); halo voxels outside the grid are zeros, or the
// wrapped-around voxels along the axes a.wrap marks periodic.  The reduction
// index is k = t * cin + ci as in the weight layout, t = 9 dz + 3 dy + dx: a
// 32-wide k-step holds one tap column of 32 channels (CK = 32, cin a
// multiple of 32, one chunk of 32 channels at a time), two columns of 16
// channels (CK = 16) or three of 8 (CK = 8, the fourth lane group reading
// zeros), and the three dy steps of one group share their A fragments (the
// loop below).  B fragments come from global memory (the weights stay in L2;
// staging them in LDS beside the brick measured slower: the lower occupancy
// cost more than the L2 reads it saved).
constexpr int kBz = 4, kBy = 4, kBx = 16;
constexpr int kHz = kBz + 2, kHy = kBy + 2, kHx = kBx + 2;
constexpr int kHvox = kHz * kHy * kHx;

template <int CK>
struct BrickLds {
    // voxel stride in LDS, in elements (a multiple of 8, so every 16-byte read
    // is aligned), chosen so that the 16 row lanes of an A read (16 consecutive
    // voxels, one 16-byte piece each) cover the 64 banks without conflict:
    // CK = 32: 40 (32 channels and 8 of padding); CK = 8: 8, unpadded (a
    // 16-byte stride covers the banks once); CK = 16: 16, unpadded, with the
    // voxel's two 16-byte halves swapped on every other run of 8 voxels, so that
    // voxels v and v + 8 use different halves of the same 32 banks.  Unpadded,
    // the 8- and 16-channel bricks take 10 and 21 KB of LDS instead of 31 KB,
    // and the occupancy is no longer set by the LDS (5 blocks per CU before).
    static constexpr int S = CK == 32 ? 40 : CK;
    // element offset of the 8-channel group q of brick voxel v
    __device__ static inline int at(int v, int q) {
        if constexpr (CK == 16) return v * S + ((q ^ ((v >> 3) & 1)) << 3);
        else return v * S + q * 8;
    }
};

// Staging of a brick chunk: thread t holds pieces p = t + i * kThreads
// (piece = voxel v of the brick with its halo, 8-channel group q) in
// registers, so all of its global loads are in flight at once, and the next
// chunk's loads overlap the current chunk's MFMAs.
template <int CK>
struct BrickStage {
    static constexpr int PPV = CK / 8;  // 16-byte pieces per voxel
    static constexpr int NP = kHvox * PPV;
    static constexpr int NIT = (NP + kThreads - 1) / kThreads;
};

template <int CK>
__device__ inline void brick_load(const ConvArgs& a, int n, int x0, int y0, int z0, int j,
                                  bf16x8 (&v)[BrickStage<CK>::NIT]) {
    constexpr int PPV = BrickStage<CK>::PPV, NP = BrickStage<CK>::NP;
    const bf16* xb = a.x + a.xco + j * 32;
#pragma unroll
    for (int i = 0; i < BrickStage<CK>::NIT; ++i) {
        const int p = threadIdx.x + i * kThreads;
        v[i] = zero8();
        if (p < NP) {
            const int vx = p / PPV, q = p - vx * PPV;
            const int hz = vx / (kHy * kHx), rem = vx - hz * (kHy * kHx);
            const int hy = rem / kHx, hx = rem - hy * kHx;
            const int zz = pad_coord(z0 - 1 + hz, a.in.d, a.wrap & 4), yy = pad_coord(y0 - 1 + hy, a.in.h, a.wrap & 2),
                      xx = pad_coord(x0 - 1 + hx, a.in.w, a.wrap & 1);
            // (a brick that overhangs the grid's far end by more than the halo reads only zeros there)
            if (zz >= 0 && yy >= 0 && xx >= 0 && zz < a.in.d && yy < a.in.h && xx < a.in.w)
                v[i] = load8(xb + (int64_t)vox(a.in, n, zz, yy, xx) * a.xcs + q * 8);
        }
    }
}

template <int CK>
__device__ inline void brick_store(bf16* s_in, const bf16x8 (&v)[BrickStage<CK>::NIT]) {
    constexpr int PPV = BrickStage<CK>::PPV, NP = BrickStage<CK>::NP;
#pragma unroll
    for (int i = 0; i < BrickStage<CK>::NIT; ++i) {
        const int p = threadIdx.x + i * kThreads;
        if (p < NP) {
            const int vx = p / PPV, q = p - vx * PPV;
            *reinterpret_cast<bf16x8*>(&s_in[BrickLds<CK>::at(vx, q)]) = v[i];
        }
    }
}

template <int CK, int NT>
__global__ __launch_bounds__(kThreads) void k_conv3_lds(ConvArgs a, int nbx, int nby, int nbz) {
    constexpr int S = BrickLds<CK>::S;
    __shared__ __attribute__((aligned(16))) bf16 s_in[kHvox * S];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int g = lane >> 4, r16 = lane & 15;
    const int co_base = blockIdx.y * 16 * NT;
    // brick coordinates: x fastest, then y, z, sample (numbering the bricks
    // per XCD, so that bricks sharing halo voxels share an L2, measured no
    // faster)
    uint32_t b = blockIdx.x;
    const int bx = (int)(b % (uint32_t)nbx);
    b /= (uint32_t)nbx;
    const int by = (int)(b % (uint32_t)nby);
    b /= (uint32_t)nby;
    const int bz = (int)(b % (uint32_t)nbz);
    const int n = (int)(b / (uint32_t)nbz);
    const int x0 = bx * kBx, y0 = by * kBy, z0 = bz * kBz;
    f32x4 acc[4][NT];
#pragma unroll
    for (int mt = 0; mt < 4; ++mt)
#pragma unroll
        for (int nt = 0; nt < NT; ++nt) acc[mt][nt] = f32x4{0.f, 0.f, 0.f, 0.f};

    const int nchunks = CK == 32 ? a.cin / 32 : 1;
    const int rowbase = wave * (kHy * kHx) + r16;  // brick voxel of (w, mt = 0, r16), tap (0, 0, 0)
    const bf16* wrow[NT];
#pragma unroll
    for (int nt = 0; nt < NT; ++nt) wrow[nt] = a.wt + (int64_t)(co_base + nt * 16 + r16) * a.kpad;

    bf16x8 stage[BrickStage<CK>::NIT];
    brick_load<CK>(a, n, x0, y0, z0, 0, stage);
    brick_store<CK>(s_in, stage);
    __syncthreads();
    for (int j = 0; j < nchunks; ++j) {
        // the next chunk's loads are in flight during this chunk's MFMAs
        if (j + 1 < nchunks) brick_load<CK>(a, n, x0, y0, z0, j + 1, stage);
        // k-steps in groups: in group gi the lanes of group g take one tap
        // column (dz, dx) of the brick and channels c of it.  CK = 32: nine
        // groups (dz, dx), all 32 channels.  CK = 16: groups 0-2 take
        // dx = g/2 at dz = gi; groups 3 and 4 take the dx = 2 columns of
        // dz = 0, 1 and of dz = 2 (the other half of group 4 reads zeros), so
        // that 5 groups cover the 9 columns (16 channels each, c = 8 (g & 1)).
        // CK = 8: group dz, dx = g (g = 3 reads zeros).  The six brick rows
        // hy = 0..5 of the column serve the three taps dy = 0..2 of all four
        // m-tiles (row mt + dy), so each A fragment feeds two MFMAs on average.
        // A lane's B fragment holds the weights of its own tap (9 dz + 3 dy +
        // dx) and channels, so lanes of one k-step may take different taps.
        constexpr int NG = CK == 32 ? 9 : (CK == 16 ? 5 : 3);
        const int cgl = CK == 32 ? g * 8 : (CK == 16 ? (g & 1) * 8 : 0);
#pragma unroll
        for (int gi = 0; gi < NG; ++gi) {
            int dz, dx;
            if (CK == 32) {
                dz = gi / 3;
                dx = gi % 3;
            } else if (CK == 16) {
                dz = gi < 3 ? gi : 2 * (gi - 3) + (g >> 1);
                dx = gi < 3 ? (g >> 1) : 2;
            } else {
                dz = gi;
                dx = g;
            }
            const bool valid = dz < 3 && dx < 3;
            bf16x8 B[3][NT];
            if (!valid) {
#pragma unroll
                for (int dy = 0; dy < 3; ++dy)
#pragma unroll
                    for (int nt = 0; nt < NT; ++nt) B[dy][nt] = zero8();
            } else {
                // weight row offset k = t * cin + ci of tap (dz, dy = 0, dx) and this lane's channels
                const int kw = (dz * 9 + dx) * a.cin + (CK == 32 ? j * 32 : 0) + cgl;
#pragma unroll
                for (int dy = 0; dy < 3; ++dy)
#pragma unroll
                    for (int nt = 0; nt < NT; ++nt) B[dy][nt] = load8(wrow[nt] + kw + dy * 3 * a.cin);
            }
            bf16x8 F[kHy];
            if (valid) {
                const int v0 = rowbase + dz * (kHy * kHx) + dx;
#pragma unroll
                for (int hy = 0; hy < kHy; ++hy) F[hy] = load8(&s_in[BrickLds<CK>::at(v0 + hy * kHx, cgl >> 3)]);
            } else {
#pragma unroll
                for (int hy = 0; hy < kHy; ++hy) F[hy] = zero8();
            }
#pragma unroll
            for (int dy = 0; dy < 3; ++dy)
#pragma unroll
                for (int mt = 0; mt < 4; ++mt)
#pragma unroll
                    for (int nt = 0; nt < NT; ++nt)
                        acc[mt][nt] =
                            __builtin_amdgcn_mfma_f32_16x16x32_bf16(F[mt + dy], B[dy][nt], acc[mt][nt], 0, 0, 0);
        }
        if (j + 1 < nchunks) {
            __syncthreads();
            brick_store<CK>(s_in, stage);
            __syncthreads();
        }
    }
    // epilogue: lane holds x = x0 + 4g + i of row (z0 + w, y0 + mt), column r16 of each tile
    const int z = z0 + wave;
    if (z >= a.out.d) return;
#pragma unroll
    for (int mt = 0; mt < 4; ++mt) {
        const int y = y0 + mt;
        if (y >= a.out.h) continue;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int x = x0 + 4 * g + i;
            if (x >= a.out.w) continue;
            const int64_t ov = vox(a.out, n, z, y, x);
#pragma unroll
            for (int nt = 0; nt < NT; ++nt) {
                const int co = co_base + nt * 16 + r16;
                if (co >= a.cstore) continue;
                float v = acc[mt][nt][i];
                if (a.bias) v += a.bias[co];
                if (a.relu) v = v > 0.f ? v : 0.f;
                if (a.aux && !((float)a.aux[ov * a.acs + a.aco + co] > 0.f)) v = 0.f;
                if (a.yf) a.yf[ov * a.cstore + co] = v;
                else a.y[ov * a.ycs + a.yco + co] = (bf16)v;
            }
        }
    }
}

// Weight gradient: dw[par][co][t*cin + ci] = sum_rows dy[out(row)][co] * x[in(row,t)][ci],
// split over chunks of rows: block (chunk, tap, tile) writes its sum to
// part[chunk][par][co][t*cin + ci] and k_wgrad_sum adds the chunks in order
// (a fixed-order split-K sum, so gradients are bitwise reproducible).
// Block: one tap (blockIdx.y), a TC x TC channel tile with TC = 16*N
// (blockIdx.z), a chunk of rows (blockIdx.x).  The 4 waves are N ci sub-tiles
// times 4/N row slabs of 32 rows; each step stages R = 128/N rows of TC dy and
// TC x channels in LDS transposed (channel-major, one 16-byte load per thread
// and tensor), so a lane reads 8 consecutive rows of a channel with one
// 16-byte LDS load.  N is the smallest of 1, 2, 4 with 16*N >= the channel
// counts, so the 16-channel layers do not run a mostly empty 64x64 tile.
struct WgradArgs {
    const bf16* x;
    int xcs, xco, cin;     // cin = padded input channels (multiple of 8)
    const bf16* dy;
    int dcs, dco, coutp;
    float* part;           // [chunks][npar][coutp][kpad] block sums
    int npar;              // 8 for a transposed 2x2x2 convolution (blockIdx.y = parity), else 1
    int kpad, ntaps;
    Dims rows, in, out;
    int64_t m;
    int64_t chunk;         // rows per block
    int nci_tiles;         // ceil(cin / TC)
    int wrap;              // periodic axes of a 3x3x3 convolution (as ConvArgs::wrap)
    int xcd_order;         // 1: a 1-D grid numbered per XCD (see k_wgrad), 0: grid (chunks, taps, tiles)
    int nchunks, gy, ntiles;
    int ci_off, co_off;    // k_wgrad3_lds: the first input and output channels of this call
};

template <int GEOM, int N>
__global__ __launch_bounds__(kThreads) void k_wgrad(WgradArgs a) {
    constexpr int TC = 16 * N, R = 128 / N, RP = R + 8;
    __shared__ __attribute__((aligned(16))) bf16 s_dy[TC][RP];
    __shared__ __attribute__((aligned(16))) bf16 s_x[TC][RP];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int g = lane >> 4, r16 = lane & 15;
    const int cis = wave % N, slab = wave / N;
    // Block -> (chunk, tap or parity, channel tile).  With a.xcd_order the
    // grid is 1-D and numbered so that every block of one chunk runs on the
    // same XCD (the hardware deals block ids round-robin to the 8 XCDs) and
    // close together in time: the chunk's dy rows and the x rows of its 27
    // taps are then read from that XCD's L2 instead of from memory by each
    // tap's block.  Grid: ceil(chunks / 8) * 8 * taps * tiles blocks.
    int chunk, ty, tile;
    if (a.xcd_order) {
        const uint32_t b = blockIdx.x, inner = (uint32_t)(a.gy * a.ntiles);
        const uint32_t q = b >> 3, j = q % inner;
        chunk = (int)((q / inner) * 8u + (b & 7u));
        ty = (int)(j % (uint32_t)a.gy);
        tile = (int)(j / (uint32_t)a.gy);
        if (chunk >= a.nchunks) return;
    } else {
        chunk = blockIdx.x;
        ty = blockIdx.y;
        tile = blockIdx.z;
    }
    const int t = GEOM == G_UP2 ? 0 : ty;
    const int par = GEOM == G_UP2 ? ty : 0;
    const int co_tile = tile / a.nci_tiles, ci_tile = tile % a.nci_tiles;
    const int co0 = co_tile * TC, ci0 = ci_tile * TC;
    const int64_t rbeg = (int64_t)chunk * a.chunk;
    const int64_t rend = rbeg + a.chunk < a.m ? rbeg + a.chunk : a.m;

    f32x4 acc[N];
#pragma unroll
    for (int i = 0; i < N; ++i) acc[i] = f32x4{0.f, 0.f, 0.f, 0.f};

    ConvArgs ca;
    ca.in = a.in;
    ca.wrap = a.wrap;
    // staging role: row rr of the step, channels cc..cc+7 of the tile
    const int rr = threadIdx.x % R, cc = (threadIdx.x / R) * 8;
    for (int64_t r0 = rbeg; r0 < rend; r0 += R) {
        {
            const int64_t r = r0 + rr;
            bf16x8 vdy = zero8(), vx = zero8();
            if (r < rend) {
                int n, z, y, x;
                unvox(a.rows, r, n, z, y, x);
                int64_t ov;
                if (GEOM == G_UP2) ov = vox(a.out, n, 2 * z + (par >> 2), 2 * y + ((par >> 1) & 1), 2 * x + (par & 1));
                else ov = vox(a.out, n, z, y, x);
                if (co0 + cc < a.coutp) vdy = load8(a.dy + ov * a.dcs + a.dco + co0 + cc);
                int64_t iv;
                if (ci0 + cc < a.cin && in_voxel<GEOM>(ca, n, z, y, x, t, iv)) vx = load8(a.x + iv * a.xcs + a.xco + ci0 + cc);
            }
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                s_dy[cc + j][rr] = vdy[j];
                s_x[cc + j][rr] = vx[j];
            }
        }
        __syncthreads();
        // sub-tiles past the channel counts (a 16-channel side of a 32-channel
        // tile) hold zeros: their waves and MFMAs skip them (wave-uniform)
        if (ci0 + cis * 16 < a.cin) {
            const bf16x8 B = load8(&s_x[cis * 16 + r16][slab * 32 + g * 8]);
#pragma unroll
            for (int ct = 0; ct < N; ++ct) {
                if (co0 + ct * 16 >= a.coutp) continue;
                const bf16x8 A = load8(&s_dy[ct * 16 + r16][slab * 32 + g * 8]);
                acc[ct] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A, B, acc[ct], 0, 0, 0);
            }
        }
        __syncthreads();
    }
    // the row slabs of one ci sub-tile add their sums in LDS (fixed order), so
    // only the first slab's wave writes the block sums
    if (N < 4) {
        __shared__ float s_acc[4][4 * N][64];
#pragma unroll
        for (int ct = 0; ct < N; ++ct)
#pragma unroll
            for (int i = 0; i < 4; ++i) s_acc[wave][ct * 4 + i][lane] = acc[ct][i];
        __syncthreads();
        if (slab != 0) return;
#pragma unroll
        for (int sl = 1; sl < 4 / N; ++sl)
#pragma unroll
            for (int ct = 0; ct < N; ++ct)
#pragma unroll
                for (int i = 0; i < 4; ++i) acc[ct][i] += s_acc[sl * N + cis][ct * 4 + i][lane];
    }
    // acc[ct]: rows co = co0 + ct*16 + 4g + i, column ci = ci0 + cis*16 + r16
    const int ci = ci0 + cis * 16 + r16;
    if (ci >= a.cin) return;
    float* dw = a.part + ((int64_t)chunk * a.npar + par) * a.coutp * a.kpad;
#pragma unroll
    for (int ct = 0; ct < N; ++ct)
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int co = co0 + ct * 16 + 4 * g + i;
            if (co < a.coutp) dw[(int64_t)co * a.kpad + t * a.cin + ci] = acc[ct][i];
        }
}

// Weight gradient of a same-size 3x3x3 convolution, 16 output channels and
// CK = 8 or 16 input channels per call (wider layers run in 16-channel slices
// of each side, one call per pair: co_off, ci_off = 0, 16, ...), all 27 taps
// in one block.  A block walks a
// fixed range of bricks of 4 (z) x 4 (y) x 16 (x) output voxels; per brick it
// stages dy (channel-major, so a lane reads 8 consecutive rows of one output
// channel as the A fragment of dy^T) and the input brick with its one-voxel
// halo, as three copies shifted by dx = 0, 1, 2 in x (so the B fragment of
// every tap is an aligned 16-byte read), once, and runs the 27 taps from LDS:
// wave w takes output plane w (64 rows, two 32-row k-steps), tap t adds
// dy^T x_t into acc[t] (a 16 x 16 tile, co x ci).  At the end the four waves'
// tiles are added in order and the block writes them as its partial
// part[block][co][t * cin + ci] (k_wgrad_sum adds the blocks in order).  The
// row-chunk kernel (k_wgrad) stages its rows once per tap instead.
template <int CK>
struct WgradBrick {
    static constexpr int CIS = kHz * kHy * 16 + 8;  // x copy: ci stride (elements); the 16 ci lanes of a B read on disjoint banks
    static constexpr int COPY = CK * CIS;
    static constexpr int COS = kBz * 64 + 8;       // dy: co stride (elements), likewise
};

template <int CK>
__global__ __launch_bounds__(kThreads) void k_wgrad3_lds(WgradArgs a, int nbx, int nby, int nbz, int nbricks) {
    using WB = WgradBrick<CK>;
    __shared__ __attribute__((aligned(16))) bf16 s_x[3 * WB::COPY];
    __shared__ __attribute__((aligned(16))) bf16 s_dy[16 * WB::COS];
    __shared__ float s_red[kThreads / 64][256];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int g = lane >> 4, r16 = lane & 15;
    const int bb0 = (int)((int64_t)blockIdx.x * nbricks / gridDim.x);
    const int bb1 = (int)((int64_t)(blockIdx.x + 1) * nbricks / gridDim.x);
    const Dims in = a.in;
    f32x4 acc[27];
#pragma unroll
    for (int t = 0; t < 27; ++t) acc[t] = f32x4{0.f, 0.f, 0.f, 0.f};
    for (int br = bb0; br < bb1; ++br) {
        uint32_t b = (uint32_t)br;
        const int bx = (int)(b % (uint32_t)nbx);
        b /= (uint32_t)nbx;
        const int by = (int)(b % (uint32_t)nby);
        b /= (uint32_t)nby;
        const int bz = (int)(b % (uint32_t)nbz);
        const int n = (int)(b / (uint32_t)nbz);
        const int x0 = bx * kBx, y0 = by * kBy, z0 = bz * kBz;
        __syncthreads();  // the previous brick's reads are done
        // dy: piece p = (voxel v of the brick, 8-channel group q)
        for (int p = threadIdx.x; p < kBz * 64 * 2; p += kThreads) {
            const int v = p >> 1, q = p & 1;
            const int w = v >> 6, r = v & 63;
            const int zz = z0 + w, yy = y0 + (r >> 4), xx = x0 + (r & 15);
            bf16x8 val = zero8();
            if (zz < a.out.d && yy < a.out.h && xx < a.out.w)
                val = load8(a.dy + vox(a.out, n, zz, yy, xx) * a.dcs + a.dco + a.co_off + q * 8);
#pragma unroll
            for (int j = 0; j < 8; ++j) s_dy[(q * 8 + j) * WB::COS + v] = val[j];
        }
        // x with its halo: piece p = (halo voxel, 8-channel group q), into the copies it belongs to
        for (int p = threadIdx.x; p < kHvox * (CK / 8); p += kThreads) {
            const int v = p / (CK / 8), q = p - v * (CK / 8);
            const int hz = v / (kHy * kHx), rem = v - hz * (kHy * kHx);
            const int hy = rem / kHx, hx = rem - hy * kHx;
            const int zz = pad_coord(z0 - 1 + hz, in.d, a.wrap & 4), yy = pad_coord(y0 - 1 + hy, in.h, a.wrap & 2),
                      xx = pad_coord(x0 - 1 + hx, in.w, a.wrap & 1);
            bf16x8 val = zero8();
            if (zz >= 0 && yy >= 0 && xx >= 0 && zz < in.d && yy < in.h && xx < in.w)
                val = load8(a.x + vox(in, n, zz, yy, xx) * a.xcs + a.xco + a.ci_off + q * 8);
            const int row = (hz * kHy + hy) * 16;
#pragma unroll
            for (int dx = 0; dx < 3; ++dx) {
                const int e = hx - dx;
                if (e < 0 || e >= 16) continue;
#pragma unroll
                for (int j = 0; j < 8; ++j) s_x[dx * WB::COPY + (q * 8 + j) * WB::CIS + row + e] = val[j];
            }
        }
        __syncthreads();
        // k-step s of the wave: rows s*32 + g*8 .. + 7 = y-row 2s + g/2, x (g & 1) * 8 .. + 7
        const bf16x8 A0 = load8(&s_dy[r16 * WB::COS + wave * 64 + g * 8]);
        const bf16x8 A1 = load8(&s_dy[r16 * WB::COS + wave * 64 + 32 + g * 8]);
#pragma unroll
        for (int t = 0; t < 27; ++t) {
            const int dz = t / 9, dyy = (t / 3) % 3, dx = t % 3;
#pragma unroll
            for (int st = 0; st < 2; ++st) {
                const int hz = wave + dz, hy = 2 * st + (g >> 1) + dyy;
                const bf16x8 B = r16 < CK ? load8(&s_x[dx * WB::COPY + r16 * WB::CIS + (hz * kHy + hy) * 16 + (g & 1) * 8])
                                          : zero8();
                acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(st ? A1 : A0, B, acc[t], 0, 0, 0);
            }
        }
    }
    // the waves' tiles, tap by tap, added in wave order; acc[t] lane (g, r16)
    // element i: co = 4g + i, ci = r16
    float* part = a.part + (int64_t)blockIdx.x * a.coutp * a.kpad;
#pragma unroll
    for (int t = 0; t < 27; ++t) {
#pragma unroll
        for (int i = 0; i < 4; ++i) s_red[wave][i * 64 + lane] = acc[t][i];
        __syncthreads();
        if (wave == 0 && r16 < CK && a.ci_off + r16 < a.cin) {
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                float v = s_red[0][i * 64 + lane];
#pragma unroll
                for (int w = 1; w < kThreads / 64; ++w) v += s_red[w][i * 64 + lane];
                part[(int64_t)(a.co_off + 4 * g + i) * a.kpad + t * a.cin + a.ci_off + r16] = v;
            }
        }
        __syncthreads();
    }
}

// dw[e] = the sum over chunks c of part[c][e] (e = (par, co, k) of an
// [npar][coutp][kpad] tensor) in a fixed order: a block takes 32 elements;
// its 8 lane groups g each add the chunks c = g, g + 8, ... in turn, then
// group 0 adds the 8 group sums in order.  k at or past kval (the taps times
// the padded input channels) holds no block sum and gets zero.
constexpr int kWsumEl = 32, kWsumGroups = kThreads / kWsumEl;

__global__ __launch_bounds__(kThreads) void k_wgrad_sum(const float* __restrict__ part, int nchunks, int64_t stride,
                                                         int kpad, int kval, float* __restrict__ dw) {
    __shared__ float s_red[kWsumGroups][kWsumEl];
    const int el = threadIdx.x % kWsumEl, grp = threadIdx.x / kWsumEl;
    const int64_t e = (int64_t)blockIdx.x * kWsumEl + el;
    const bool live = e < stride && (int)(e % kpad) < kval;
    float s = 0.f;
    if (live) {
        const float* p = part + e;
        for (int c = grp; c < nchunks; c += kWsumGroups) s += p[(int64_t)c * stride];
    }
    s_red[grp][el] = s;
    __syncthreads();
    if (grp == 0 && e < stride) {
        float t = 0.f;
#pragma unroll
        for (int q = 0; q < kWsumGroups; ++q) t += s_red[q][el];
        dw[e] = t;
    }
}

// The elementwise kernels below take one thread per voxel (or per voxel and
// 8-channel group, blockIdx.y) with 32-bit voxel indices (every tensor holds
// fewer than 2^31 voxels) and 16-byte accesses to channels-last rows: an
// index decomposition per element in 64 bits costs more than the access.
__device__ inline void store8(bf16* p, const bf16x8& v) { *reinterpret_cast<bf16x8*>(p) = v; }

// 2x2x2 average pooling: out[v] = mean of the 8 children of v.  Thread t
// takes the 8-channel group t % cg8 of coarse voxel t / cg8 (groups fastest,
// so that neighbouring lanes read neighbouring bytes); t < nv * cg8 < 2^32.
__global__ __launch_bounds__(kThreads) void k_avgpool(const bf16* __restrict__ x, int xcs, int xco, bf16* __restrict__ y,
                                                       int ycs, int yco, Dims out, Dims in, int cg8, uint32_t total) {
    const uint32_t t = blockIdx.x * kThreads + threadIdx.x;
    if (t >= total) return;
    const uint32_t v = t / (uint32_t)cg8;
    const int c0 = (int)(t - v * (uint32_t)cg8) * 8;
    int n, z, yy, xx;
    unvox(out, v, n, z, yy, xx);
    float s[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) s[j] = 0.f;
    for (int p = 0; p < 8; ++p) {
        const int64_t iv = vox(in, n, 2 * z + (p >> 2), 2 * yy + ((p >> 1) & 1), 2 * xx + (p & 1));
        const bf16x8 a = load8(x + iv * xcs + xco + c0);
#pragma unroll
        for (int j = 0; j < 8; ++j) s[j] += (float)a[j];
    }
    bf16x8 o;
#pragma unroll
    for (int j = 0; j < 8; ++j) o[j] = (bf16)(0.125f * s[j]);
    store8(y + (int64_t)v * ycs + yco + c0, o);
}

// Backward of the pooling, added into dx and masked by (fwd > 0); thread t
// takes group t % cg8 of fine voxel t / cg8.
__global__ __launch_bounds__(kThreads) void k_avgpool_bwd(const bf16* __restrict__ dy, int dcs, int dco,
                                                           bf16* __restrict__ dx, int xcs, int xco,
                                                           const bf16* __restrict__ fwd, int fcs, int fco,
                                                           Dims in, Dims out, int cg8, uint32_t total) {
    const uint32_t t = blockIdx.x * kThreads + threadIdx.x;
    if (t >= total) return;
    const uint32_t v = t / (uint32_t)cg8;
    const int c0 = (int)(t - v * (uint32_t)cg8) * 8;
    int n, z, yy, xx;
    unvox(in, v, n, z, yy, xx);
    const int64_t ov = vox(out, n, z >> 1, yy >> 1, xx >> 1);
    const bf16x8 g8 = load8(dy + ov * dcs + dco + c0);
    const bf16x8 f8 = load8(fwd + (int64_t)v * fcs + fco + c0);
    bf16* p = dx + (int64_t)v * xcs + xco + c0;
    bf16x8 d8 = load8(p);
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        float g = 0.125f * (float)g8[j];
        if (!((float)f8[j] > 0.f)) g = 0.f;
        d8[j] = (bf16)((float)d8[j] + g);
    }
    store8(p, d8);
}

// FP32 host-layout input [b][c][N] -> BF16 channels-last with cp >= c
// channels (a multiple of 8), one thread per voxel of the batch.
__global__ __launch_bounds__(kThreads) void k_pack_input(const float* __restrict__ src, int c, int cp, int nvox,
                                                          int batch_vox, bf16* __restrict__ dst) {
    const int v = (int)(blockIdx.x * kThreads + threadIdx.x);
    if (v >= batch_vox) return;
    const int b = v / nvox, s = v - b * nvox;
    for (int c0 = 0; c0 < cp; c0 += 8) {
        bf16x8 o;
#pragma unroll
        for (int j = 0; j < 8; ++j)
            o[j] = (bf16)(c0 + j < c ? src[((int64_t)b * c + c0 + j) * nvox + s] : 0.f);
        store8(dst + (int64_t)v * cp + c0, o);
    }
}

// FP32 channels-last [v][c] -> host layout [b][c][N], one thread per voxel.
__global__ __launch_bounds__(kThreads) void k_unpack_output(const float* __restrict__ src, int c, int nvox,
                                                             int batch_vox, float* __restrict__ dst) {
    const int v = (int)(blockIdx.x * kThreads + threadIdx.x);
    if (v >= batch_vox) return;
    const int b = v / nvox, s = v - b * nvox;
    for (int ch = 0; ch < c; ++ch) dst[((int64_t)b * c + ch) * nvox + s] = src[(int64_t)v * c + ch];
}

// Mean-squared error, one thread per voxel: dy (BF16, channel stride cp = 16,
// zero padded) = 2 (y - t) / count; per-block partial sums of (y - t)^2 in
// FP32 to partial[blockIdx.x].
__global__ __launch_bounds__(kThreads) void k_mse(const float* __restrict__ y, const float* __restrict__ target, int c,
                                                   int nvox, int batch_vox, float inv_count,
                                                   bf16* __restrict__ dy, float* __restrict__ partial) {
    __shared__ float s[kThreads];
    const int v = (int)(blockIdx.x * kThreads + threadIdx.x);
    float e2 = 0.f;
    if (v < batch_vox) {
        const int b = v / nvox, sidx = v - b * nvox;
        bf16x8 g[2];
#pragma unroll
        for (int ch = 0; ch < 16; ++ch) {
            float gv = 0.f;
            if (ch < c) {
                const float d = y[(int64_t)v * c + ch] - target[((int64_t)b * c + ch) * nvox + sidx];
                e2 += d * d;
                gv = 2.f * d * inv_count;
            }
            g[ch >> 3][ch & 7] = (bf16)gv;
        }
        store8(dy + (int64_t)v * 16, g[0]);
        store8(dy + (int64_t)v * 16 + 8, g[1]);
    }
    s[threadIdx.x] = e2;
    __syncthreads();
    for (int off = kThreads / 2; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) s[threadIdx.x] += s[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = s[0];
}

// Master layout [npar][cout][taps][cin] (+ bias) <-> padded layouts.
struct LayerDesc {
    int geom;      // G_CONV3, G_UP2, G_PW
    int npar, ntaps;
    int cin, cinp, cout, coutp;
    int kpad;      // forward: roundup(ntaps * cinp, 32)
    int kpad_d;    // input-gradient conv: roundup(ntaps * coutp, 32)
    int64_t w_off, b_off;
};

// forward weights wt[par][co][t*cinp + ci] (BF16, zero padded) and bias (FP32, padded)
__global__ __launch_bounds__(kThreads) void k_pack_fwd(const float* __restrict__ master, LayerDesc L,
                                                        bf16* __restrict__ wt, float* __restrict__ bias) {
    const int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x;
    const int64_t total = (int64_t)L.npar * L.coutp * L.kpad;
    if (i < L.coutp && bias) bias[i] = i < L.cout ? master[L.b_off + i] : 0.f;
    if (i >= total) return;
    const int k = (int)(i % L.kpad);
    const int64_t r = i / L.kpad;
    const int co = (int)(r % L.coutp), par = (int)(r / L.coutp);
    const int t = k / L.cinp, ci = k % L.cinp;
    float v = 0.f;
    if (t < L.ntaps && ci < L.cin && co < L.cout)
        v = master[L.w_off + (((int64_t)par * L.cout + co) * L.ntaps + t) * L.cin + ci];
    wt[i] = (bf16)v;
}

// input-gradient weights: the conv from dy (coutp channels) to dx (cinp channels).
//   conv3: wd[ci][t'*coutp + co] = w[co][26 - t'][ci]
//   up2:   wd[ci][p*coutp + co]  = w[p][co][ci]      (a stride-2 2x2x2 conv)
//   pw:    wd[ci][co]            = w[co][ci]
__global__ __launch_bounds__(kThreads) void k_pack_dgrad(const float* __restrict__ master, LayerDesc L,
                                                          bf16* __restrict__ wd) {
    const int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x;
    const int64_t total = (int64_t)L.cinp * L.kpad_d;
    if (i >= total) return;
    const int k = (int)(i % L.kpad_d);
    const int ci = (int)(i / L.kpad_d);
    const int tt = k / L.coutp, co = k % L.coutp;
    const int taps_d = L.geom == G_UP2 ? L.npar : L.ntaps;
    float v = 0.f;
    if (tt < taps_d && ci < L.cin && co < L.cout) {
        if (L.geom == G_CONV3) v = master[L.w_off + ((int64_t)co * 27 + (26 - tt)) * L.cin + ci];
        else if (L.geom == G_UP2) v = master[L.w_off + ((int64_t)tt * L.cout + co) * L.cin + ci];
        else v = master[L.w_off + (int64_t)co * L.cin + ci];
    }
    wd[i] = (bf16)v;
}

// padded FP32 weight gradient -> master layout, plus the bias gradient computed
// separately (k_bias_part, k_bias_sum).
__global__ __launch_bounds__(kThreads) void k_unpack_grad(const float* __restrict__ dwp, LayerDesc L,
                                                           float* __restrict__ grad) {
    const int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x;
    const int64_t total = (int64_t)L.npar * L.cout * L.ntaps * L.cin;
    if (i >= total) return;
    const int ci = (int)(i % L.cin);
    int64_t r = i / L.cin;
    const int t = (int)(r % L.ntaps);
    r /= L.ntaps;
    const int co = (int)(r % L.cout), par = (int)(r / L.cout);
    grad[L.w_off + i] = dwp[((int64_t)par * L.coutp + co) * L.kpad + t * L.cinp + ci];
}

// Bias gradient, sum over voxels of dy[v][co], in two fixed-order passes.
// Pass 1: block (b, cb) sums its range of `chunk` voxels for channels
// cb*cp .. cb*cp + cp - 1 (cp a power of two <= kThreads): thread = (8-channel
// group tid % ng, phase tid / ng), ng = ceil(cp / 8), so that a lane loads 16
// bytes and a wave reads whole channel runs of consecutive voxels; the phases
// are then summed by a tree in LDS.  The channel stride and offset of dy are
// multiples of 8 and the rows hold at least 8 channels (zero padded).
constexpr int kBiasParts = 512;

__global__ __launch_bounds__(kThreads) void k_bias_part(const bf16* __restrict__ dy, int dcs, int dco, int cout,
                                                         int cp, int64_t nvox, int64_t chunk,
                                                         float* __restrict__ part) {
    __shared__ float s[8][kThreads];
    const int ng = (cp + 7) >> 3, nph = kThreads / ng, ph = threadIdx.x / ng;
    const int c0 = blockIdx.y * cp + (int)(threadIdx.x % ng) * 8;
    const int64_t v0 = (int64_t)blockIdx.x * chunk;
    const int64_t v1 = v0 + chunk < nvox ? v0 + chunk : nvox;
    float a[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) a[j] = 0.f;
    if (c0 < cout)
        for (int64_t v = v0 + ph; v < v1; v += nph) {
            const bf16x8 x = load8(dy + v * dcs + dco + c0);
#pragma unroll
            for (int j = 0; j < 8; ++j) a[j] += (float)x[j];
        }
#pragma unroll
    for (int j = 0; j < 8; ++j) s[j][threadIdx.x] = a[j];
    __syncthreads();
    for (int off = nph / 2; off > 0; off >>= 1) {
        if (ph < off) {
#pragma unroll
            for (int j = 0; j < 8; ++j) s[j][threadIdx.x] += s[j][threadIdx.x + off * ng];
        }
        __syncthreads();
    }
    if (ph == 0) {
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            const int c = c0 + j;
            if (j < cp && c < cout) part[(int64_t)blockIdx.x * cout + c] = s[j][threadIdx.x];
        }
    }
}

// Pass 2: grad[b_off + co] = sum of the nb partials of channel co (one block per channel).
__global__ __launch_bounds__(kThreads) void k_bias_sum(const float* __restrict__ part, int nb, int cout,
                                                        int64_t b_off, float* __restrict__ grad) {
    __shared__ float s[kThreads];
    const int co = blockIdx.x;
    float a = 0.f;
    for (int b = threadIdx.x; b < nb; b += kThreads) a += part[(int64_t)b * cout + co];
    s[threadIdx.x] = a;
    __syncthreads();
    for (int off = kThreads / 2; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) s[threadIdx.x] += s[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0) grad[b_off + co] = s[0];
}

__global__ __launch_bounds__(kThreads) void k_adam(float* __restrict__ p, const float* __restrict__ g,
                                                    float* __restrict__ m, float* __restrict__ v, int64_t n, float lr,
                                                    float b1, float b2, float eps, float c1, float c2) {
    const int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

__global__ __launch_bounds__(kThreads) void k_fill_bf16(bf16* __restrict__ p, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x;
    if (i < n) p[i] = (bf16)0.f;
}

__global__ __launch_bounds__(kThreads) void k_fill_f32(float* __restrict__ p, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x;
    if (i < n) p[i] = 0.f;
}

}  // namespace unet
