// unet_capi.hip -- the warm-start U-Net runtime behind include/lbm_unet.h:
// layer table, activation/gradient buffers, forward, backward and Adam, all
// on this library's kernels (unet_kernels.h).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "../../include/lbm_unet.h"
#include "unet_kernels.h"

using namespace unet;

namespace {

struct Buf {
    bf16* p = nullptr;
    int cs = 0;  // channel stride
};

int64_t rup(int64_t a, int64_t b) { return (a + b - 1) / b * b; }
unsigned nblk(int64_t n) { return (unsigned)((n + kThreads - 1) / kThreads); }

}  // namespace

struct lbm_unet {
    lbm_unet_config cfg{};
    int L = 0;
    std::vector<int> C;       // channels per level
    std::vector<Dims> dims;   // per level, batch = cfg.batch
    std::vector<int64_t> nvox;  // voxels per sample per level
    std::vector<LayerDesc> layers;
    // layer indices
    std::vector<int> enc1, enc2, up, dec1, dec2;
    int head = -1;
    int64_t nparams = 0;
    float* d_params = nullptr;
    float* d_grad = nullptr;
    float* d_m = nullptr;
    float* d_v = nullptr;
    std::vector<bf16*> wt, wd;      // per layer packed BF16 weights
    std::vector<float*> bias, dwp;  // per layer padded bias and FP32 weight gradient
    // activations (forward) and gradients
    Buf X0;
    std::vector<Buf> E1, Cat, P, F1, D;
    std::vector<Buf> dE1, dCat, dP, dF1, dD;
    float* Out = nullptr;   // FP32 [B*N0][cout]
    Buf dOut;               // BF16 [B*N0][16]
    float* d_in = nullptr;  // FP32 staging of host x
    float* d_tgt = nullptr;
    float* d_part = nullptr;
    float* d_bpart = nullptr;  // bias-gradient partials [kBiasParts][largest cout]
    float* d_wpart = nullptr;  // weight-gradient block sums [chunks][npar][coutp][kpad] (grown on demand)
    int64_t wpart_n = 0;
    // weight-gradient blocks numbered per XCD (LBM_UNET_WGRAD_XCD=0: the plain
    // (chunk, tap, tile) grid, for comparison); the all-taps brick kernel for
    // the 16-channel layers (LBM_UNET_WGRAD_BRICK=0: the row-chunk kernel)
    bool wgrad_xcd = true;
    bool wgrad_brick = true;
    // narrow pointwise layers one voxel per thread (LBM_UNET_PW_NARROW=0: k_conv)
    bool pw_narrow = true;
    int64_t part_n = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    int64_t adam_t = 0;
    double last_ms = 0.0, last_flops = 0.0;
    std::vector<void*> allocs;
    std::string err;
};

namespace {

int fail(lbm_unet* n, int code, const std::string& m) {
    if (n) n->err = m;
    return code;
}

#define UCHK(net, call)                                                                              \
    do {                                                                                             \
        hipError_t e_ = (call);                                                                      \
        if (e_ != hipSuccess) return fail((net), -2, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

template <class T>
int alloc(lbm_unet* n, T** p, int64_t count) {
    hipError_t e = hipMalloc((void**)p, (size_t)std::max<int64_t>(count, 1) * sizeof(T));
    if (e != hipSuccess) return fail(n, -2, std::string("hipMalloc: ") + hipGetErrorString(e));
    n->allocs.push_back(*p);
    return 0;
}

int alloc_act(lbm_unet* n, Buf& b, int level, int cs) {
    b.cs = cs;
    return alloc(n, &b.p, (int64_t)n->cfg.batch * n->nvox[level] * cs);
}

Dims with_batch(Dims d, int b) {
    d.n = b;
    return d;
}

int add_layer(lbm_unet* n, int geom, int cin, int cinp, int cout, int coutp) {
    LayerDesc L;
    std::memset(&L, 0, sizeof(L));
    L.geom = geom;
    L.npar = geom == G_UP2 ? 8 : 1;
    L.ntaps = geom == G_CONV3 ? 27 : 1;
    L.cin = cin;
    L.cinp = cinp;
    L.cout = cout;
    L.coutp = coutp;
    L.kpad = (int)rup((int64_t)L.ntaps * cinp, 32);
    const int taps_d = geom == G_UP2 ? 8 : L.ntaps;
    L.kpad_d = (int)rup((int64_t)taps_d * coutp, 32);
    L.w_off = n->nparams;
    n->nparams += (int64_t)L.npar * cout * L.ntaps * cin;
    L.b_off = n->nparams;
    n->nparams += cout;
    n->layers.push_back(L);
    return (int)n->layers.size() - 1;
}

// ---- launches --------------------------------------------------------------
int conv(lbm_unet* n, int geom, const Buf& x, int xco, int cin, const bf16* w, int coutp, int kpad, int ntaps,
         int npar, const float* bias, bf16* y, int ycs, int yco, int cstore, float* yf, const Buf* aux, int aco,
         int relu, Dims rows, Dims in, Dims out) {
    ConvArgs a;
    std::memset(&a, 0, sizeof(a));
    a.x = x.p;
    a.xcs = x.cs;
    a.xco = xco;
    a.cin = cin;
    a.wt = w;
    a.coutp = coutp;
    a.kpad = kpad;
    a.ntaps = ntaps;
    a.bias = bias;
    a.y = y;
    a.ycs = ycs;
    a.yco = yco;
    a.cstore = cstore;
    a.yf = yf;
    a.aux = aux ? aux->p : nullptr;
    a.acs = aux ? aux->cs : 0;
    a.aco = aco;
    a.relu = relu;
    a.rows = rows;
    a.in = in;
    a.out = out;
    a.m = (int64_t)rows.n * rows.d * rows.h * rows.w;
    a.wrap = geom == G_CONV3 ? (n->cfg.periodic & 7) : 0;
    // k_conv indexes voxels in 32 bits and takes a same-size 3x3x3 convolution's rows on the input grid
    const int64_t nin = (int64_t)in.n * in.d * in.h * in.w;
    if (nin >= (int64_t)1 << 31 || a.m >= (int64_t)1 << 31) return fail(n, -2, "convolution tensor of 2^31 voxels or more");
    if (geom == G_CONV3 && (rows.n != in.n || rows.d != in.d || rows.h != in.h || rows.w != in.w))
        return fail(n, -2, "3x3x3 convolution rows and input differ in size");
    if (geom != G_UP2 && (rows.n != out.n || rows.d != out.d || rows.h != out.h || rows.w != out.w))
        return fail(n, -2, "convolution rows and output differ in size");
    // 16*nt output channels per wave; narrower where a wide tile leaves the chip short of blocks
    // (the coarse levels have few rows)
    // nt in {1, 2, 4} divides coutp / 16, so the grid's y extent covers every channel
    int nt = 4;
    while (nt > 1 && (coutp / 16) % nt != 0) nt >>= 1;
    // same-size 3x3x3 convolutions on grids at least one brick wide take the
    // LDS-staged kernel when the input channels fit its chunking (8, 16 or a
    // multiple of 32)
    const int ck = cin == 8 ? 8 : (cin == 16 ? 16 : (cin % 32 == 0 ? 32 : 0));
    if (geom == G_CONV3 && ck && in.w >= kBx && in.h >= kBy && in.d >= kBz) {
        const int nbx = (in.w + kBx - 1) / kBx, nby = (in.h + kBy - 1) / kBy, nbz = (in.d + kBz - 1) / kBz;
        const int64_t bricks = (int64_t)nbx * nby * nbz * in.n;
        if (bricks >= (int64_t)1 << 31) return fail(n, -2, "convolution grid too large");
        while (nt > 1 && bricks * (coutp / (16 * nt)) < 512) nt >>= 1;
        const dim3 grid3((unsigned)bricks, (unsigned)(coutp / (16 * nt)), 1), block3(kThreads);
#define LAUNCH_C3(CK, NTV) hipLaunchKernelGGL((k_conv3_lds<CK, NTV>), grid3, block3, 0, n->stream, a, nbx, nby, nbz)
#define C3_BY_NT(CK)                     \
    if (nt == 1) LAUNCH_C3(CK, 1);       \
    else if (nt == 2) LAUNCH_C3(CK, 2);  \
    else LAUNCH_C3(CK, 4);
        if (ck == 8) {
            C3_BY_NT(8);
        } else if (ck == 16) {
            C3_BY_NT(16);
        } else {
            C3_BY_NT(32);
        }
#undef C3_BY_NT
#undef LAUNCH_C3
        UCHK(n, hipGetLastError());
        return 0;
    }
    if (a.wrap) return fail(n, -2, "periodic padding needs 3x3x3 convolutions of at least 16 x 4 x 4 voxels");
    if (geom == G_UP2 && kpad % 32 == 0 && kpad / 32 <= 4 && (int64_t)nblk(a.m) * (coutp / 16) >= 256) {
        // all 8 parities per block, the input rows read once (k_up2); a grid
        // that would leave most CUs idle keeps one parity per block (k_conv)
        while (nt > 1 && (int64_t)nblk(a.m) * (coutp / (16 * nt)) < 512) nt >>= 1;
        const dim3 gridu(nblk(a.m), (unsigned)(coutp / (16 * nt)), 1), blocku(kThreads);
        const int ks = kpad / 32;
#define LAUNCH_UP2(NTV, KSV) hipLaunchKernelGGL((k_up2<NTV, KSV>), gridu, blocku, 0, n->stream, a)
#define UP2_BY_KS(NTV)                     \
    if (ks == 1) LAUNCH_UP2(NTV, 1);       \
    else if (ks == 2) LAUNCH_UP2(NTV, 2);  \
    else if (ks == 3) LAUNCH_UP2(NTV, 3);  \
    else LAUNCH_UP2(NTV, 4);
        if (nt == 1) {
            UP2_BY_KS(1);
        } else if (nt == 2) {
            UP2_BY_KS(2);
        } else {
            UP2_BY_KS(4);
        }
#undef UP2_BY_KS
#undef LAUNCH_UP2
        UCHK(n, hipGetLastError());
        return 0;
    }
    // BF16 outputs of 16 channels go in 16-byte pieces (k_pw_narrow's vector
    // path); with scalar 2-byte stores strided across the wave the head's input
    // gradient took 329 us one voxel per thread against 108 us on k_conv
    const bool pw_vec = !yf && cstore % 8 == 0 && ycs % 8 == 0 && yco % 8 == 0 &&
                        (!aux || (aux->cs % 8 == 0 && aco % 8 == 0));
    if (geom == G_PW && n->pw_narrow && npar == 1 && cin % 8 == 0 && cin <= 32 &&
        (cstore <= 8 || (cstore <= 16 && pw_vec)) && rows.n == in.n && rows.d == in.d && rows.h == in.h &&
        rows.w == in.w) {
        // narrow pointwise layers (the output head and its input gradient):
        // one thread per voxel (k_pw_narrow)
        const dim3 gp((unsigned)((a.m + kThreads - 1) / kThreads)), bp(kThreads);
        if (cin == 8) hipLaunchKernelGGL((k_pw_narrow<8>), gp, bp, 0, n->stream, a);
        else if (cin == 16) hipLaunchKernelGGL((k_pw_narrow<16>), gp, bp, 0, n->stream, a);
        else if (cin == 24) hipLaunchKernelGGL((k_pw_narrow<24>), gp, bp, 0, n->stream, a);
        else hipLaunchKernelGGL((k_pw_narrow<32>), gp, bp, 0, n->stream, a);
        UCHK(n, hipGetLastError());
        return 0;
    }
    while (nt > 1 && (int64_t)nblk(a.m) * (coutp / (16 * nt)) * npar < 512) nt >>= 1;
    const dim3 grid(nblk(a.m), (unsigned)(coutp / (16 * nt)), (unsigned)npar), block(kThreads);
#define LAUNCH_CONV(G, NTV) hipLaunchKernelGGL((k_conv<G, NTV>), grid, block, 0, n->stream, a)
#define BY_NT(G)                       \
    if (nt == 1) LAUNCH_CONV(G, 1);    \
    else if (nt == 2) LAUNCH_CONV(G, 2); \
    else LAUNCH_CONV(G, 4);
    switch (geom) {
        case G_CONV3: BY_NT(G_CONV3); break;
        case G_DOWN2: BY_NT(G_DOWN2); break;
        case G_UP2: BY_NT(G_UP2); break;
        default: BY_NT(G_PW); break;
    }
#undef BY_NT
#undef LAUNCH_CONV
    UCHK(n, hipGetLastError());
    return 0;
}

int wgrad(lbm_unet* n, int li, const Buf& x, int xco, const Buf& dy, int dco, Dims rows, Dims in, Dims out) {
    const LayerDesc& L = n->layers[li];
    WgradArgs a;
    std::memset(&a, 0, sizeof(a));
    a.x = x.p;
    a.xcs = x.cs;
    a.xco = xco;
    a.cin = L.cinp;
    a.dy = dy.p;
    a.dcs = dy.cs;
    a.dco = dco;
    a.coutp = L.coutp;
    a.npar = L.geom == G_UP2 ? 8 : 1;
    a.kpad = L.kpad;
    a.ntaps = L.ntaps;
    a.rows = rows;
    a.in = in;
    a.out = out;
    a.m = (int64_t)rows.n * rows.d * rows.h * rows.w;
    a.chunk = std::max<int64_t>(512, rup(a.m / 512 + 1, 128));
    a.wrap = L.geom == G_CONV3 ? (n->cfg.periodic & 7) : 0;
    // channel tile 16*N: the smallest that holds both channel counts, at most 64
    const int cmax = std::max(L.cinp, L.coutp);
    const int N = cmax <= 16 ? 1 : (cmax <= 32 ? 2 : 4);
    const int TC = 16 * N;
    a.nci_tiles = (L.cinp + TC - 1) / TC;
    const int nco_tiles = (L.coutp + TC - 1) / TC;
    const unsigned gy = L.geom == G_UP2 ? 8 : (unsigned)L.ntaps;
    // 16-output-channel 3x3x3 layers with 8 or 16 input channels: all 27 taps
    // per block from LDS bricks (k_wgrad3_lds), one partial per block
    const int nbx = (in.w + kBx - 1) / kBx, nby = (in.h + kBy - 1) / kBy, nbz = (in.d + kBz - 1) / kBz;
    const int64_t nbricks = (int64_t)nbx * nby * nbz * in.n;
    // (one call per 16 x 16 channel slice pair, on grids of at least 512 bricks)
    const bool brick = n->wgrad_brick && L.geom == G_CONV3 && L.coutp % 16 == 0 &&
                       (L.cinp == 8 || L.cinp % 16 == 0) && (L.coutp == 16 || nbricks >= 512);
    const int nchunks = brick ? (int)std::min<int64_t>(nbricks, 512) : (int)((a.m + a.chunk - 1) / a.chunk);
    const int64_t stride = (int64_t)a.npar * L.coutp * L.kpad;
    if ((int64_t)nchunks * stride > n->wpart_n) {
        // the block sums of the largest layer seen so far (the first training step grows it)
        UCHK(n, hipStreamSynchronize(n->stream));
        float* old = n->d_wpart;
        n->d_wpart = nullptr;
        n->wpart_n = 0;
        if (old) UCHK(n, hipFree(old));
        UCHK(n, hipMalloc(&n->d_wpart, (size_t)nchunks * stride * sizeof(float)));
        n->wpart_n = (int64_t)nchunks * stride;
    }
    a.part = n->d_wpart;
    if (brick) {
        // one call per (16 output, 16 input) channel slice; each writes its
        // channels of every block partial
        for (int o0 = 0; o0 < L.coutp; o0 += 16) {
            a.co_off = o0;
            if (L.cinp == 8) {
                a.ci_off = 0;
                hipLaunchKernelGGL((k_wgrad3_lds<8>), dim3((unsigned)nchunks), dim3(kThreads), 0, n->stream, a, nbx,
                                   nby, nbz, (int)nbricks);
                continue;
            }
            for (int c0 = 0; c0 < L.cinp; c0 += 16) {
                a.ci_off = c0;
                hipLaunchKernelGGL((k_wgrad3_lds<16>), dim3((unsigned)nchunks), dim3(kThreads), 0, n->stream, a, nbx,
                                   nby, nbz, (int)nbricks);
            }
        }
        a.ci_off = 0;
        a.co_off = 0;
    }
    a.xcd_order = n->wgrad_xcd ? 1 : 0;
    a.nchunks = nchunks;
    a.gy = (int)gy;
    a.ntiles = nco_tiles * a.nci_tiles;
    const dim3 grid = a.xcd_order ? dim3((unsigned)(((nchunks + 7) / 8) * 8 * (int)gy * a.ntiles))
                                  : dim3((unsigned)nchunks, gy, (unsigned)a.ntiles);
#define LAUNCH_WGRAD(G)                                                                          \
    if (N == 1) hipLaunchKernelGGL((k_wgrad<G, 1>), grid, dim3(kThreads), 0, n->stream, a);    \
    else if (N == 2) hipLaunchKernelGGL((k_wgrad<G, 2>), grid, dim3(kThreads), 0, n->stream, a); \
    else hipLaunchKernelGGL((k_wgrad<G, 4>), grid, dim3(kThreads), 0, n->stream, a);
    if (!brick) {
        switch (L.geom) {
            case G_CONV3: LAUNCH_WGRAD(G_CONV3); break;
            case G_UP2: LAUNCH_WGRAD(G_UP2); break;
            default: LAUNCH_WGRAD(G_PW); break;
        }
    }
#undef LAUNCH_WGRAD
    UCHK(n, hipGetLastError());
    hipLaunchKernelGGL(k_wgrad_sum, dim3((unsigned)((stride + kWsumEl - 1) / kWsumEl)), dim3(kThreads), 0, n->stream,
                       n->d_wpart, nchunks, stride, L.kpad, L.ntaps * L.cinp, n->dwp[li]);
    UCHK(n, hipGetLastError());
    const int64_t ng = (int64_t)L.npar * L.cout * L.ntaps * L.cin;
    hipLaunchKernelGGL(k_unpack_grad, dim3(nblk(ng)), dim3(kThreads), 0, n->stream, n->dwp[li], L, n->d_grad);
    const int64_t nv = (int64_t)out.n * out.d * out.h * out.w;
    int cp = 1;
    while (cp < L.cout && cp < kThreads) cp <<= 1;
    const int64_t nb = std::min<int64_t>(kBiasParts, (nv + kThreads - 1) / kThreads);
    const int64_t chunk = (nv + nb - 1) / nb;
    hipLaunchKernelGGL(k_bias_part, dim3((unsigned)nb, (unsigned)((L.cout + cp - 1) / cp)), dim3(kThreads), 0,
                       n->stream, dy.p, dy.cs, dco, L.cout, cp, nv, chunk, n->d_bpart);
    hipLaunchKernelGGL(k_bias_sum, dim3((unsigned)L.cout), dim3(kThreads), 0, n->stream, n->d_bpart, (int)nb, L.cout,
                       L.b_off, n->d_grad);
    UCHK(n, hipGetLastError());
    return 0;
}

// forward of layer li from x (channel offset xco) into y
int layer_fwd(lbm_unet* n, int li, const Buf& x, int xco, Buf* y, int yco, float* yf, int relu, Dims rows, Dims in,
              Dims out) {
    const LayerDesc& L = n->layers[li];
    return conv(n, L.geom, x, xco, L.cinp, n->wt[li], L.coutp, L.kpad, L.ntaps, L.npar, n->bias[li],
                y ? y->p : nullptr, y ? y->cs : 0, yco, yf ? L.cout : L.cout, yf, nullptr, 0, relu, rows, in, out);
}

// input gradient of layer li: dx (channel offset dxo) from dy (offset dco), masked by aux
int layer_dgrad(lbm_unet* n, int li, const Buf& dy, int dco, Buf& dx, int dxo, const Buf* aux, int aco, Dims out_dims,
                Dims in_dims) {
    const LayerDesc& L = n->layers[li];
    // the input-gradient conv reads dy (out_dims) and writes dx (in_dims)
    int geom = L.geom == G_UP2 ? G_DOWN2 : L.geom;
    const int taps_d = L.geom == G_UP2 ? 8 : L.ntaps;
    return conv(n, geom, dy, dco, L.coutp, n->wd[li], L.cinp, L.kpad_d, taps_d, 1, nullptr, dx.p, dx.cs, dxo, L.cin,
                nullptr, aux, aco, 0, in_dims, out_dims, in_dims);
}

int pack_weights(lbm_unet* n) {
    for (size_t li = 0; li < n->layers.size(); ++li) {
        const LayerDesc& L = n->layers[li];
        const int64_t nf = (int64_t)L.npar * L.coutp * L.kpad;
        hipLaunchKernelGGL(k_pack_fwd, dim3(nblk(std::max<int64_t>(nf, L.coutp))), dim3(kThreads), 0, n->stream,
                           n->d_params, L, n->wt[li], n->bias[li]);
        const int64_t nd = (int64_t)L.cinp * L.kpad_d;
        hipLaunchKernelGGL(k_pack_dgrad, dim3(nblk(nd)), dim3(kThreads), 0, n->stream, n->d_params, L, n->wd[li]);
    }
    UCHK(n, hipGetLastError());
    return 0;
}

double conv_flops(const LayerDesc& L, int64_t rows) {
    // rows = output voxels (conv3, pw) or coarse input voxels (up2: 8 outputs each)
    return 2.0 * (double)rows * L.npar * L.ntaps * L.cin * L.cout;
}

int forward(lbm_unet* n, int B) {
    const int Lv = n->L;
    std::vector<Dims> d(Lv);
    for (int l = 0; l < Lv; ++l) d[l] = with_batch(n->dims[l], B);
    double fl = 0.0;
    int rc;
    for (int l = 0; l < Lv; ++l) {
        const Buf& in = l == 0 ? n->X0 : n->P[l];
        rc = layer_fwd(n, n->enc1[l], in, 0, &n->E1[l], 0, nullptr, 1, d[l], d[l], d[l]);
        if (rc) return rc;
        fl += conv_flops(n->layers[n->enc1[l]], (int64_t)B * n->nvox[l]);
        if (l < Lv - 1) {
            rc = layer_fwd(n, n->enc2[l], n->E1[l], 0, &n->Cat[l], n->C[l], nullptr, 1, d[l], d[l], d[l]);
            if (rc) return rc;
            const int64_t tot = (int64_t)B * n->nvox[l + 1] * (n->C[l] / 8);
            hipLaunchKernelGGL(k_avgpool, dim3(nblk(tot)), dim3(kThreads), 0, n->stream, n->Cat[l].p, n->Cat[l].cs,
                               n->C[l], n->P[l + 1].p, n->P[l + 1].cs, 0, d[l + 1], d[l], n->C[l] / 8, (uint32_t)tot);
        } else {
            rc = layer_fwd(n, n->enc2[l], n->E1[l], 0, &n->D[l], 0, nullptr, 1, d[l], d[l], d[l]);
            if (rc) return rc;
        }
        fl += conv_flops(n->layers[n->enc2[l]], (int64_t)B * n->nvox[l]);
    }
    for (int l = Lv - 2; l >= 0; --l) {
        rc = layer_fwd(n, n->up[l], n->D[l + 1], 0, &n->Cat[l], 0, nullptr, 1, d[l + 1], d[l + 1], d[l]);
        if (rc) return rc;
        fl += conv_flops(n->layers[n->up[l]], (int64_t)B * n->nvox[l + 1]);
        rc = layer_fwd(n, n->dec1[l], n->Cat[l], 0, &n->F1[l], 0, nullptr, 1, d[l], d[l], d[l]);
        if (rc) return rc;
        rc = layer_fwd(n, n->dec2[l], n->F1[l], 0, &n->D[l], 0, nullptr, 1, d[l], d[l], d[l]);
        if (rc) return rc;
        fl += conv_flops(n->layers[n->dec1[l]], (int64_t)B * n->nvox[l]) +
              conv_flops(n->layers[n->dec2[l]], (int64_t)B * n->nvox[l]);
    }
    rc = layer_fwd(n, n->head, n->D[0], 0, nullptr, 0, n->Out, 0, d[0], d[0], d[0]);
    fl += conv_flops(n->layers[n->head], (int64_t)B * n->nvox[0]);
    n->last_flops = fl;
    return rc;
}

int zero_bf16(lbm_unet* n, const Buf& b, int64_t vox) {
    const int64_t cnt = vox * b.cs;
    hipLaunchKernelGGL(k_fill_bf16, dim3(nblk(cnt)), dim3(kThreads), 0, n->stream, b.p, cnt);
    UCHK(n, hipGetLastError());
    return 0;
}

int backward(lbm_unet* n, int B) {
    const int Lv = n->L;
    std::vector<Dims> d(Lv);
    for (int l = 0; l < Lv; ++l) d[l] = with_batch(n->dims[l], B);
    int rc;
    // head
    rc = wgrad(n, n->head, n->D[0], 0, n->dOut, 0, d[0], d[0], d[0]);
    if (rc) return rc;
    rc = layer_dgrad(n, n->head, n->dOut, 0, n->dD[0], 0, &n->D[0], 0, d[0], d[0]);
    if (rc) return rc;
    // decoder, fine to coarse
    for (int l = 0; l <= Lv - 2; ++l) {
        rc = wgrad(n, n->dec2[l], n->F1[l], 0, n->dD[l], 0, d[l], d[l], d[l]);
        if (rc) return rc;
        rc = layer_dgrad(n, n->dec2[l], n->dD[l], 0, n->dF1[l], 0, &n->F1[l], 0, d[l], d[l]);
        if (rc) return rc;
        rc = wgrad(n, n->dec1[l], n->Cat[l], 0, n->dF1[l], 0, d[l], d[l], d[l]);
        if (rc) return rc;
        rc = layer_dgrad(n, n->dec1[l], n->dF1[l], 0, n->dCat[l], 0, &n->Cat[l], 0, d[l], d[l]);
        if (rc) return rc;
        // up-conv from level l+1: its output is dCat[l] channels [0, C_l)
        rc = wgrad(n, n->up[l], n->D[l + 1], 0, n->dCat[l], 0, d[l + 1], d[l + 1], d[l]);
        if (rc) return rc;
        rc = layer_dgrad(n, n->up[l], n->dCat[l], 0, n->dD[l + 1], 0, &n->D[l + 1], 0, d[l], d[l + 1]);
        if (rc) return rc;
    }
    // encoder, coarse to fine
    for (int l = Lv - 1; l >= 0; --l) {
        const Buf& dy2 = l == Lv - 1 ? n->dD[l] : n->dCat[l];
        const int dco2 = l == Lv - 1 ? 0 : n->C[l];
        rc = wgrad(n, n->enc2[l], n->E1[l], 0, dy2, dco2, d[l], d[l], d[l]);
        if (rc) return rc;
        rc = layer_dgrad(n, n->enc2[l], dy2, dco2, n->dE1[l], 0, &n->E1[l], 0, d[l], d[l]);
        if (rc) return rc;
        const Buf& in = l == 0 ? n->X0 : n->P[l];
        rc = wgrad(n, n->enc1[l], in, 0, n->dE1[l], 0, d[l], d[l], d[l]);
        if (rc) return rc;
        if (l > 0) {
            rc = layer_dgrad(n, n->enc1[l], n->dE1[l], 0, n->dP[l], 0, nullptr, 0, d[l], d[l]);
            if (rc) return rc;
            const int64_t tot = (int64_t)B * n->nvox[l - 1] * (n->C[l - 1] / 8);
            hipLaunchKernelGGL(k_avgpool_bwd, dim3(nblk(tot)), dim3(kThreads), 0, n->stream, n->dP[l].p,
                               n->dP[l].cs, 0, n->dCat[l - 1].p, n->dCat[l - 1].cs, n->C[l - 1], n->Cat[l - 1].p,
                               n->Cat[l - 1].cs, n->C[l - 1], d[l - 1], d[l], n->C[l - 1] / 8, (uint32_t)tot);
            UCHK(n, hipGetLastError());
        }
    }
    return 0;
}

int upload_x(lbm_unet* n, const float* x, int B) {
    const int64_t N0 = n->nvox[0];
    UCHK(n, hipMemcpyAsync(n->d_in, x, (size_t)B * n->cfg.cin * N0 * sizeof(float), hipMemcpyHostToDevice, n->stream));
    hipLaunchKernelGGL(k_pack_input, dim3(nblk((int64_t)B * N0)), dim3(kThreads), 0, n->stream, n->d_in, n->cfg.cin,
                       8, (int)N0, (int)((int64_t)B * N0), n->X0.p);
    UCHK(n, hipGetLastError());
    return 0;
}

int loss_grad(lbm_unet* n, const float* target, int B, float* loss) {
    const int64_t N0 = n->nvox[0];
    UCHK(n, hipMemcpyAsync(n->d_tgt, target, (size_t)B * n->cfg.cout * N0 * sizeof(float), hipMemcpyHostToDevice,
                           n->stream));
    const int64_t nb = nblk((int64_t)B * N0);
    const float inv = 1.0f / (float)((double)B * N0 * n->cfg.cout);
    hipLaunchKernelGGL(k_mse, dim3((unsigned)nb), dim3(kThreads), 0, n->stream, n->Out, n->d_tgt, n->cfg.cout, (int)N0,
                       (int)((int64_t)B * N0), inv, n->dOut.p, n->d_part);
    UCHK(n, hipGetLastError());
    std::vector<float> h((size_t)nb);
    UCHK(n, hipMemcpyAsync(h.data(), n->d_part, (size_t)nb * sizeof(float), hipMemcpyDeviceToHost, n->stream));
    UCHK(n, hipStreamSynchronize(n->stream));
    double s = 0.0;
    for (float v : h) s += v;
    if (loss) *loss = (float)(s * inv);
    return 0;
}

}  // namespace

extern "C" {

int lbm_unet_create(const lbm_unet_config* cfg, lbm_unet** out) {
    if (!cfg || !out) return -1;
    *out = nullptr;
    if (cfg->levels < 2 || cfg->levels > 5 || cfg->cin < 1 || cfg->cin > 8 || cfg->cout < 1 || cfg->cout > 16 ||
        cfg->base < 16 || cfg->base % 16 || cfg->batch < 1 || (cfg->periodic & ~7))
        return -1;
    const int f = 1 << (cfg->levels - 1);
    if (cfg->nx % f || cfg->ny % f || cfg->nz % f || cfg->nx < f || cfg->ny < f || cfg->nz < f) return -1;
    // wrapped convolutions run only in the brick kernel, which needs 16 x 4 x 4 voxels at the coarsest level
    if (cfg->periodic && (cfg->nx / f < 16 || cfg->ny / f < 4 || cfg->nz / f < 4)) return -1;
    // the kernels index voxels (not elements) in 32 bits, and the pooling
    // kernels (voxel, 8-channel group) pairs of level 0 in 32 unsigned bits
    if ((int64_t)cfg->batch * cfg->nx * cfg->ny * cfg->nz >= (int64_t)1 << 31) return -1;
    if ((int64_t)cfg->batch * cfg->nx * cfg->ny * cfg->nz * (cfg->base / 8) >= (int64_t)1 << 32) return -1;
    lbm_unet* n = new lbm_unet();
    n->cfg = *cfg;
    if (const char* e = std::getenv("LBM_UNET_WGRAD_XCD")) n->wgrad_xcd = std::atoi(e) != 0;
    if (const char* e = std::getenv("LBM_UNET_WGRAD_BRICK")) n->wgrad_brick = std::atoi(e) != 0;
    if (const char* e = std::getenv("LBM_UNET_PW_NARROW")) n->pw_narrow = std::atoi(e) != 0;
    n->L = cfg->levels;
    for (int l = 0; l < n->L; ++l) {
        n->C.push_back(cfg->base << l);
        Dims dm{cfg->batch, cfg->nz >> l, cfg->ny >> l, cfg->nx >> l};
        n->dims.push_back(dm);
        n->nvox.push_back((int64_t)dm.d * dm.h * dm.w);
    }
    const int Lv = n->L;
    n->enc1.resize(Lv);
    n->enc2.resize(Lv);
    n->up.resize(Lv);
    n->dec1.resize(Lv);
    n->dec2.resize(Lv);
    for (int l = 0; l < Lv; ++l) {
        const int cin = l == 0 ? cfg->cin : n->C[l - 1];
        const int cinp = l == 0 ? 8 : n->C[l - 1];
        n->enc1[l] = add_layer(n, G_CONV3, cin, cinp, n->C[l], n->C[l]);
        n->enc2[l] = add_layer(n, G_CONV3, n->C[l], n->C[l], n->C[l], n->C[l]);
    }
    for (int l = Lv - 2; l >= 0; --l) {
        n->up[l] = add_layer(n, G_UP2, n->C[l + 1], n->C[l + 1], n->C[l], n->C[l]);
        n->dec1[l] = add_layer(n, G_CONV3, 2 * n->C[l], 2 * n->C[l], n->C[l], n->C[l]);
        n->dec2[l] = add_layer(n, G_CONV3, n->C[l], n->C[l], n->C[l], n->C[l]);
    }
    n->head = add_layer(n, G_PW, n->C[0], n->C[0], cfg->cout, 16);
    int rc = 0;
    if (hipStreamCreateWithFlags(&n->stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&n->ev0) != hipSuccess || hipEventCreate(&n->ev1) != hipSuccess) {
        delete n;
        return -2;
    }
    rc = rc ? rc : alloc(n, &n->d_params, n->nparams);
    rc = rc ? rc : alloc(n, &n->d_grad, n->nparams);
    rc = rc ? rc : alloc(n, &n->d_m, n->nparams);
    rc = rc ? rc : alloc(n, &n->d_v, n->nparams);
    n->wt.resize(n->layers.size());
    n->wd.resize(n->layers.size());
    n->bias.resize(n->layers.size());
    n->dwp.resize(n->layers.size());
    for (size_t li = 0; li < n->layers.size() && !rc; ++li) {
        const LayerDesc& L = n->layers[li];
        rc = alloc(n, &n->wt[li], (int64_t)L.npar * L.coutp * L.kpad);
        rc = rc ? rc : alloc(n, &n->wd[li], (int64_t)L.cinp * L.kpad_d);
        rc = rc ? rc : alloc(n, &n->bias[li], L.coutp);
        rc = rc ? rc : alloc(n, &n->dwp[li], (int64_t)L.npar * L.coutp * L.kpad);
    }
    n->E1.resize(Lv);
    n->Cat.resize(Lv);
    n->P.resize(Lv);
    n->F1.resize(Lv);
    n->D.resize(Lv);
    n->dE1.resize(Lv);
    n->dCat.resize(Lv);
    n->dP.resize(Lv);
    n->dF1.resize(Lv);
    n->dD.resize(Lv);
    rc = rc ? rc : alloc_act(n, n->X0, 0, 8);
    for (int l = 0; l < Lv && !rc; ++l) {
        rc = alloc_act(n, n->E1[l], l, n->C[l]);
        rc = rc ? rc : alloc_act(n, n->dE1[l], l, n->C[l]);
        rc = rc ? rc : alloc_act(n, n->D[l], l, n->C[l]);
        rc = rc ? rc : alloc_act(n, n->dD[l], l, n->C[l]);
        if (l < Lv - 1) {
            rc = rc ? rc : alloc_act(n, n->Cat[l], l, 2 * n->C[l]);
            rc = rc ? rc : alloc_act(n, n->dCat[l], l, 2 * n->C[l]);
            rc = rc ? rc : alloc_act(n, n->F1[l], l, n->C[l]);
            rc = rc ? rc : alloc_act(n, n->dF1[l], l, n->C[l]);
        }
        if (l > 0) {
            rc = rc ? rc : alloc_act(n, n->P[l], l, n->C[l - 1]);
            rc = rc ? rc : alloc_act(n, n->dP[l], l, n->C[l - 1]);
        }
    }
    const int64_t N0b = (int64_t)cfg->batch * n->nvox[0];
    rc = rc ? rc : alloc(n, &n->Out, N0b * cfg->cout);
    rc = rc ? rc : alloc_act(n, n->dOut, 0, 16);
    rc = rc ? rc : alloc(n, &n->d_in, N0b * cfg->cin);
    rc = rc ? rc : alloc(n, &n->d_tgt, N0b * cfg->cout);
    n->part_n = (N0b * 16 + kThreads - 1) / kThreads;
    rc = rc ? rc : alloc(n, &n->d_part, n->part_n);
    int maxc = 0;
    for (const LayerDesc& L : n->layers) maxc = std::max(maxc, L.cout);
    rc = rc ? rc : alloc(n, &n->d_bpart, (int64_t)kBiasParts * maxc);
    if (rc) {
        lbm_unet_destroy(n);
        return rc;
    }
    // He initialisation (normal, std sqrt(2 / fan_in)), zero bias, zero Adam state
    std::vector<float> h((size_t)n->nparams, 0.0f);
    std::mt19937_64 rng(cfg->seed);
    for (const LayerDesc& L : n->layers) {
        std::normal_distribution<float> nd(0.0f, std::sqrt(2.0f / (float)(L.ntaps * L.cin)));
        const int64_t nw = (int64_t)L.npar * L.cout * L.ntaps * L.cin;
        for (int64_t i = 0; i < nw; ++i) h[(size_t)(L.w_off + i)] = nd(rng);
    }
    rc = lbm_unet_set_params(n, h.data());
    std::vector<float> z((size_t)n->nparams, 0.0f);
    if (!rc && (hipMemcpy(n->d_m, z.data(), z.size() * 4, hipMemcpyHostToDevice) != hipSuccess ||
                hipMemcpy(n->d_v, z.data(), z.size() * 4, hipMemcpyHostToDevice) != hipSuccess))
        rc = -2;
    if (rc) {
        lbm_unet_destroy(n);
        return rc;
    }
    *out = n;
    return 0;
}

int lbm_unet_destroy(lbm_unet* n) {
    if (!n) return 0;
    if (n->stream) (void)hipStreamSynchronize(n->stream);
    for (void* p : n->allocs) (void)hipFree(p);
    if (n->d_wpart) (void)hipFree(n->d_wpart);
    if (n->ev0) (void)hipEventDestroy(n->ev0);
    if (n->ev1) (void)hipEventDestroy(n->ev1);
    if (n->stream) (void)hipStreamDestroy(n->stream);
    delete n;
    return 0;
}

int lbm_unet_last_error(const lbm_unet* n, char* buf, size_t len) {
    if (!n || !buf || !len) return -1;
    std::snprintf(buf, len, "%s", n->err.c_str());
    return 0;
}

int lbm_unet_num_params(lbm_unet* n, int64_t* np) {
    if (!n || !np) return -1;
    *np = n->nparams;
    return 0;
}

int lbm_unet_get_params(lbm_unet* n, float* out) {
    if (!n || !out) return -1;
    UCHK(n, hipMemcpyAsync(out, n->d_params, (size_t)n->nparams * 4, hipMemcpyDeviceToHost, n->stream));
    UCHK(n, hipStreamSynchronize(n->stream));
    return 0;
}

int lbm_unet_set_params(lbm_unet* n, const float* in) {
    if (!n || !in) return -1;
    UCHK(n, hipMemcpyAsync(n->d_params, in, (size_t)n->nparams * 4, hipMemcpyHostToDevice, n->stream));
    int rc = pack_weights(n);
    if (rc) return rc;
    UCHK(n, hipStreamSynchronize(n->stream));
    return 0;
}

int lbm_unet_param_layout(lbm_unet* n, int32_t* nlayers, int32_t* kind, int32_t* cin, int32_t* cout, int64_t* w_off,
                          int64_t* b_off) {
    if (!n || !nlayers) return -1;
    *nlayers = (int32_t)n->layers.size();
    for (size_t i = 0; i < n->layers.size(); ++i) {
        const LayerDesc& L = n->layers[i];
        if (kind) kind[i] = L.geom == G_CONV3 ? 0 : (L.geom == G_UP2 ? 1 : 2);
        if (cin) cin[i] = L.cin;
        if (cout) cout[i] = L.cout;
        if (w_off) w_off[i] = L.w_off;
        if (b_off) b_off[i] = L.b_off;
    }
    return 0;
}

int lbm_unet_forward(lbm_unet* n, const float* x, int32_t B, float* y) {
    if (!n || !x || !y || B < 1 || B > n->cfg.batch) return -1;
    int rc = upload_x(n, x, B);
    if (rc) return rc;
    UCHK(n, hipEventRecord(n->ev0, n->stream));
    rc = forward(n, B);
    if (rc) return rc;
    UCHK(n, hipEventRecord(n->ev1, n->stream));
    const int64_t N0 = n->nvox[0];
    float* d_y = n->d_tgt;  // staging for the host layout
    const int64_t tot = (int64_t)B * N0 * n->cfg.cout;
    hipLaunchKernelGGL(k_unpack_output, dim3(nblk((int64_t)B * N0)), dim3(kThreads), 0, n->stream, n->Out,
                       n->cfg.cout, (int)N0, (int)((int64_t)B * N0), d_y);
    UCHK(n, hipGetLastError());
    UCHK(n, hipMemcpyAsync(y, d_y, (size_t)tot * 4, hipMemcpyDeviceToHost, n->stream));
    UCHK(n, hipStreamSynchronize(n->stream));
    float ms = 0.f;
    UCHK(n, hipEventElapsedTime(&ms, n->ev0, n->ev1));
    n->last_ms = ms;
    return 0;
}

static int grad_at(lbm_unet* n, const float* x, const float* target, int B, float* loss) {
    int rc = upload_x(n, x, B);
    if (rc) return rc;
    UCHK(n, hipEventRecord(n->ev0, n->stream));
    rc = forward(n, B);
    if (rc) return rc;
    const double ffl = n->last_flops;
    rc = loss_grad(n, target, B, loss);  // synchronises
    if (rc) return rc;
    rc = backward(n, B);
    if (rc) return rc;
    n->last_flops = 3.0 * ffl;
    return 0;
}

int lbm_unet_gradient(lbm_unet* n, const float* x, const float* target, int32_t B, float* loss, float* grad) {
    if (!n || !x || !target || B < 1 || B > n->cfg.batch) return -1;
    int rc = grad_at(n, x, target, B, loss);
    if (rc) return rc;
    if (grad) UCHK(n, hipMemcpyAsync(grad, n->d_grad, (size_t)n->nparams * 4, hipMemcpyDeviceToHost, n->stream));
    UCHK(n, hipStreamSynchronize(n->stream));
    return 0;
}

int lbm_unet_train_step(lbm_unet* n, const float* x, const float* target, int32_t B, float* loss) {
    if (!n || !x || !target || B < 1 || B > n->cfg.batch) return -1;
    int rc = grad_at(n, x, target, B, loss);
    if (rc) return rc;
    ++n->adam_t;
    const float c1 = 1.0f - std::pow(n->cfg.beta1, (float)n->adam_t);
    const float c2 = 1.0f - std::pow(n->cfg.beta2, (float)n->adam_t);
    hipLaunchKernelGGL(k_adam, dim3(nblk(n->nparams)), dim3(kThreads), 0, n->stream, n->d_params, n->d_grad, n->d_m,
                       n->d_v, n->nparams, n->cfg.lr, n->cfg.beta1, n->cfg.beta2, n->cfg.eps, c1, c2);
    UCHK(n, hipGetLastError());
    rc = pack_weights(n);
    if (rc) return rc;
    UCHK(n, hipEventRecord(n->ev1, n->stream));
    UCHK(n, hipStreamSynchronize(n->stream));
    float ms = 0.f;
    UCHK(n, hipEventElapsedTime(&ms, n->ev0, n->ev1));
    n->last_ms = ms;
    return 0;
}

int lbm_unet_set_lr(lbm_unet* n, float lr) {
    if (!n || !(lr > 0.0f)) return -1;
    n->cfg.lr = lr;
    return 0;
}

int lbm_unet_last_timing(lbm_unet* n, double* ms, double* flops) {
    if (!n) return -1;
    if (ms) *ms = n->last_ms;
    if (flops) *flops = n->last_flops;
    return 0;
}

}  // extern "C"
