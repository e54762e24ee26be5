// k_unet.hip — kernels and host driver of the warm-start U-Net (unet.h):
// 3x3x3 convolutions as implicit GEMMs on the BF16 MFMA units (FP32
// accumulation), their weight and input gradients, pooling, upsampling, the
// 1x1 head with the mean-squared-error loss, and Adam. No reference
// counterpart (the reference has no learned component).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <new>
#include <random>
#include <vector>

#include "dev_math.h"
#include "unet.h"

namespace blz {

namespace {

typedef short s4v __attribute__((ext_vector_type(4)));
typedef short s8v __attribute__((ext_vector_type(8)));
typedef __bf16 bf8v __attribute__((ext_vector_type(8)));
typedef float f4m __attribute__((ext_vector_type(4)));

// BF16 <-> FP32: round to nearest even, NaN kept quiet
__device__ __forceinline__ float bf2f(bf16 h) { return __uint_as_float((unsigned)h << 16); }
__device__ __forceinline__ bf16 f2bf(float f) {
    const unsigned u = __float_as_uint(f);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (bf16)((u >> 16) | 0x40u);
    return (bf16)((u + 0x7fffu + ((u >> 16) & 1u)) >> 16);
}

struct Geo {
    int B, D, H, W;
};

__device__ __forceinline__ void vox_coords(long long v, const Geo& g, int& b, int& z, int& y, int& x) {
    x = (int)(v % g.W);
    long long t = v / g.W;
    y = (int)(t % g.H);
    t /= g.H;
    z = (int)(t % g.D);
    b = (int)(t / g.D);
}

// the same for grids of fewer than 2^31 voxels (32-bit divisions)
__device__ __forceinline__ void vox_coords32(unsigned v, const Geo& g, int& b, int& z, int& y, int& x) {
    x = (int)(v % (unsigned)g.W);
    unsigned t = v / (unsigned)g.W;
    y = (int)(t % (unsigned)g.H);
    t /= (unsigned)g.H;
    z = (int)(t % (unsigned)g.D);
    b = (int)(t / (unsigned)g.D);
}

// index in the grid twice as fine of the first child (2z, 2y, 2x) of voxel v
// (v < 2^31)
__device__ __forceinline__ long long up_child0(unsigned v, const Geo& g) {
    int b, z, y, x;
    vox_coords32(v, g, b, z, y, x);
    return (((long long)b * 2 * g.D + 2 * z) * (2 * (long long)g.H) + 2 * y) * (2 * (long long)g.W) + 2 * x;
}

// 3x3x3 convolution, zero padding: Y[v][co] = sum_tap sum_ci X[v+tap][ci] W[tap][co][ci],
// as an implicit GEMM on gfx950's v_mfma_f32_16x16x32_bf16 (A 16 x 32,
// B 32 x 16, D 16 x 16 in FP32). The reduction runs over K = 27 * CIN
// flattened as k = tap * CIN + ci, in chunks of 32: in a chunk, lane group
// kg = lane / 16 supplies k = 32 c + 8 kg .. +7, i.e. eight consecutive
// input channels of one tap (CIN is a multiple of 16, so a group never
// straddles two taps), for its voxel (A, row lane % 16) and for its output
// channel (B, column lane % 16); both operands take the same k for the same
// (lane group, element), so the product is the convolution whatever order
// the unit sums them in. A padding group past the last tap (27 * CIN not a
// multiple of 32) supplies zeros. One wave: 16 consecutive voxels x one
// 16-channel output tile (blockIdx.y); every load is 16 bytes and
// unconditional (an out-of-grid tap reads voxel 0 and is replaced by
// zeros), and the chunk loop is unrolled, so the loads are issued ahead of
// the MFMAs. D: lane l holds output channel l % 16 of voxels 4 (l / 16)
// .. +3. Optional epilogue: leaky ReLU, or the leaky ReLU's derivative at
// the activation M (1 where M > 0, NET_LEAK elsewhere).
//
// Split reduction (KS = 2 or 4) for the small grids of the lower levels,
// where one wave per tile leaves most SIMDs idle and each wave's chain of
// dependent MFMAs and loads is long: the KS waves of a tile group take
// consecutive ranges of the chunks, and the first adds the others' FP32
// tiles from LDS in split order (fixed, so every run gives the same bits).
// A workgroup holds 4 / KS tiles.
//
// UP: U is a view of the grid twice as fine in each axis that also
// receives every output value at the eight children of its voxel (the
// nearest 2x upsampling of the decoder, fused into the producing layer); a
// compile-time choice, so the other layers carry none of its code.
//
// The body of one workgroup (tile group bx along the voxels, output tile
// blockIdx.y), shared by the single-grid kernel and the multi-view one.
// ONE: each wave takes a whole tile and runs the KS splits' MFMA chains
// itself, each from zero over the same chunk range, then adds them in split
// order: the same sums as KS waves and the LDS reduction, for launches with
// tiles enough to fill the GPU without splitting them.
template <int CIN, int KS, bool UP, bool ONE = false>
__device__ __forceinline__ void conv3_tile(const BLZ_GLOBAL bf16* __restrict__ X, int Cx, int xoff,
                                           const BLZ_GLOBAL bf16* __restrict__ Wt, int Cout,
                                           BLZ_GLOBAL bf16* __restrict__ Y, int Cy, int yoff,
                                           const BLZ_GLOBAL bf16* __restrict__ Mk, int Cm, int moff, int relu,
                                           const Geo& g, long long nvox, BLZ_GLOBAL bf16* __restrict__ U, int Cu,
                                           int uoff, unsigned bx) {
    constexpr int NCH = (27 * CIN + 31) / 32;
    constexpr int NPER = (NCH + KS - 1) / KS; // chunks per split
    const int c0 = blockIdx.y * 16;
    const int lane = threadIdx.x & 63;
    // wave index and split as scalars: the chunk-range tests below are then
    // scalar branches, so every MFMA runs with the whole wave active (an
    // MFMA does not honour a partial or empty EXEC mask)
    const int wv = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6)), split = ONE ? 0 : wv % KS;
    const long long v0 = ONE ? ((long long)bx * 4 + wv) * 16 : ((long long)bx * (4 / KS) + wv / KS) * 16;
    // no early return: the waves of a workgroup meet at the split reduction
    // (a tile past the end computes zeros from clamped addresses and stores
    // nothing)
    const int li = lane & 15, kg = lane >> 4;
    const long long va = v0 + li;
    int b, z, y, x;
    // voxel indices fit in 32 bits (checked by the launcher)
    vox_coords32((unsigned)(va < nvox ? va : nvox - 1), g, b, z, y, x);
    const bool vok = va < nvox;
    // element offsets fit in 32 bits (checked by the launcher)
    const int HW = g.H * g.W;
    const int vai = (int)(va < nvox ? va : 0);
    const s8v zero = {0, 0, 0, 0, 0, 0, 0, 0};
    const BLZ_GLOBAL bf16* wrow = Wt + (size_t)(c0 + li) * CIN; // tap 0 of this lane's output channel
    // chunks in batches of CB: all loads of a batch are issued before its
    // first MFMA, so CB chunks' loads are in flight together
    constexpr int CB = 8;
    constexpr int NS = ONE ? KS : 1; // splits this wave runs
    f4m accs[NS];
#pragma unroll
    for (int sw = 0; sw < NS; ++sw) {
        const int sp = ONE ? sw : split;
        f4m acc = f4m{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int cb = 0; cb < NPER; cb += CB) {
            s8v av[CB], bv[CB];
#pragma unroll
            for (int u = 0; u < CB; ++u) {
                const int c = sp * NPER + cb + u;
                if (cb + u >= NPER || c >= NCH) continue; // wave-uniform
                const int k0 = 32 * c + 8 * kg;
                const int tap = k0 / CIN, ci = k0 - tap * CIN;
                const bool real = tap < 27;
                const int tp = real ? tap : 26;
                const int dz = tp / 9 - 1, dy = (tp / 3) % 3 - 1, dx = tp % 3 - 1;
                const int zz = z + dz, yy = y + dy, xx = x + dx;
                const bool inb = real && vok && zz >= 0 && zz < g.D && yy >= 0 && yy < g.H && xx >= 0 && xx < g.W;
                const int vv = vai + dz * HW + dy * g.W + dx;
                const s8v ld = *(const BLZ_GLOBAL s8v*)(X + (inb ? vv : 0) * Cx + xoff + (real ? ci : 0));
                const s8v wl = *(const BLZ_GLOBAL s8v*)(wrow + tp * Cout * CIN + (real ? ci : 0));
                av[u] = inb ? ld : zero;
                bv[u] = real ? wl : zero;
            }
#pragma unroll
            for (int u = 0; u < CB; ++u) {
                const int c = sp * NPER + cb + u;
                if (cb + u >= NPER || c >= NCH) continue;
                acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf8v, av[u]),
                                                               __builtin_bit_cast(bf8v, bv[u]), acc, 0, 0, 0);
            }
        }
        accs[sw] = acc;
    }
    f4m acc = accs[0];
    if constexpr (ONE) {
#pragma unroll
        for (int q = 1; q < KS; ++q) acc += accs[q];
    } else if constexpr (KS > 1) {
        __shared__ f4m part[4][64];
        if (split != 0) {
            // through VGPRs (v_accvgpr_read) rather than an LDS write straight
            // from the accumulation registers the MFMAs just wrote
            f4m t = acc;
            asm volatile("" : "+v"(t));
            part[wv][lane] = t;
        }
        __syncthreads();
        if (split != 0) return;
#pragma unroll
        for (int q = 1; q < KS; ++q) acc += part[wv + q][lane];
    }
    if (v0 >= nvox) return;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const long long vo = v0 + kg * 4 + q;
        if (vo >= nvox) continue;
        float r = acc[q];
        const int co = c0 + li;
        if (relu) r = r > 0.0f ? r : NET_LEAK * r;
        if (Mk) r = bf2f(Mk[vo * Cm + moff + co]) > 0.0f ? r : NET_LEAK * r;
        const bf16 out = f2bf(r);
        Y[vo * Cy + yoff + co] = out;
        if constexpr (UP) {
            const long long f0 = up_child0((unsigned)vo, g);
            const long long FW = 2 * (long long)g.W, FHW = 4 * (long long)g.H * g.W;
#pragma unroll
            for (int d = 0; d < 8; ++d) {
                const long long vf = f0 + (d >> 2) * FHW + ((d >> 1) & 1) * FW + (d & 1);
                U[vf * Cu + uoff + co] = out;
            }
        }
    }
}

template <int CIN, int KS, bool UP>
__global__ void __launch_bounds__(256) k_conv3_k32(const bf16* __restrict__ X, int Cx, int xoff,
                                                   const bf16* __restrict__ Wt, int Cout, bf16* __restrict__ Y,
                                                   int Cy, int yoff, const bf16* __restrict__ Mk, int Cm, int moff,
                                                   int relu, Geo g, long long nvox, bf16* __restrict__ U, int Cu,
                                                   int uoff) {
    conv3_tile<CIN, KS, UP>(as_global(X), Cx, xoff, as_global(Wt), Cout, as_global(Y), Cy, yoff, as_global(Mk), Cm,
                            moff, relu, g, nvox, as_global(U), Cu, uoff, blockIdx.x);
}

// One view's part of a multi-view launch: the forwards of several views of
// one network (same weights, each on its own grid with its own activation
// buffers) run a layer as one launch, blockIdx.z selecting the view. The
// channel layout (counts and offsets) is the layer's and common to all.
struct MView {
    const bf16* X; // layer input (pooling: the fine activation)
    bf16* Y;       // layer output (head: unused)
    bf16* U;       // upsampled copy of the output (convolutions with UP)
    float* out;    // head output
    Geo g;         // grid of the input
    long long n;   // voxels of the output (pooling: coarse voxels)
    unsigned nblk; // workgroups of this view (the grid's x extent is the largest)
    const int* stop; // nonzero: the view's output is not needed, its workgroups return
};

__device__ __forceinline__ bool mview_idle(const MView& m) {
    return blockIdx.x >= m.nblk || (m.stop && *m.stop);
}

// the forward convolution (leaky ReLU epilogue) of every view in `tab`; all
// views of one launch share the reduction split KS (ONE: each wave runs a
// whole tile's splits, for launches whose views give tiles enough)
template <int CIN, int KS, bool UP, bool ONE>
__global__ void __launch_bounds__(256) k_conv3_multi(const MView* __restrict__ tab, int Cx, int xoff,
                                                     const bf16* __restrict__ Wt, int Cout, int Cy, int yoff, int Cu,
                                                     int uoff) {
    const MView m = tab[blockIdx.z];
    if (mview_idle(m)) return; // the whole workgroup: no wave waits at the split reduction
    conv3_tile<CIN, KS, UP, ONE>(as_global(m.X), Cx, xoff, as_global(Wt), Cout, as_global(m.Y), Cy, yoff, nullptr, 0,
                                 0, 1,
                            m.g, m.n, as_global(m.U), Cu, uoff,
                            blockIdx.x);
}

// weight gradient: dW[tap][co][ci] += sum_v G[v][co] X[v+tap][ci]. A wave owns
// one 16x16 (co, ci) tile and a run of voxels; per 16 voxels the G fragment
// (A = 16 co x 16 voxels) is loaded once and used for all 27 taps (B = 16
// voxels x 16 ci, shifted by the tap). FP32 partial tiles are added with
// atomics.
__global__ void __launch_bounds__(256) k_conv3_dw(const bf16* __restrict__ X, int Cx, int xoff, int Cin,
                                                  const bf16* __restrict__ Gd, int Cg, int goff, int Cout,
                                                  float* __restrict__ dW, Geo g, long long nvox, long long run) {
    const int lane = threadIdx.x & 63;
    const long long wv = (long long)blockIdx.x * 4 + (threadIdx.x >> 6);
    const int nct = Cout >> 4, ncit = Cin >> 4;
    const int tiles = nct * ncit;
    const long long chunk = wv / tiles;
    const int tile = (int)(wv % tiles);
    const int cot = tile / ncit, cit = tile % ncit;
    const long long vbeg = chunk * run;
    if (vbeg >= nvox) return;
    const long long vend = min(nvox, vbeg + run);
    const int li = lane & 15, kb = lane >> 4;
    const long long HW = (long long)g.H * g.W;
    f4m acc[27];
#pragma unroll
    for (int t = 0; t < 27; ++t) acc[t] = f4m{0.f, 0.f, 0.f, 0.f};
    for (long long vb = vbeg; vb < vend; vb += 16) {
        s4v a;
        int bq[4], zq[4], yq[4], xq[4];
        bool okq[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const long long v = vb + kb * 4 + q;
            okq[q] = v < vend;
            vox_coords(okq[q] ? v : vbeg, g, bq[q], zq[q], yq[q], xq[q]);
            a[q] = okq[q] ? (short)Gd[v * Cg + goff + cot * 16 + li] : (short)0;
        }
#pragma unroll
        for (int tap = 0; tap < 27; ++tap) {
            const int dz = tap / 9 - 1, dy = (tap / 3) % 3 - 1, dx = tap % 3 - 1;
            s4v bx;
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int zz = zq[q] + dz, yy = yq[q] + dy, xx = xq[q] + dx;
                const bool inb = okq[q] && zz >= 0 && zz < g.D && yy >= 0 && yy < g.H && xx >= 0 && xx < g.W;
                const long long vv = vb + kb * 4 + q + dz * HW + dy * g.W + dx;
                bx[q] = inb ? (short)X[vv * Cx + xoff + cit * 16 + li] : (short)0;
            }
            acc[tap] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, bx, acc[tap], 0, 0, 0);
        }
    }
    // D[i][j]: i = co (4*kb + q), j = ci (li)
#pragma unroll
    for (int tap = 0; tap < 27; ++tap)
#pragma unroll
        for (int q = 0; q < 4; ++q)
            atomicAdd(&dW[((size_t)tap * Cout + cot * 16 + kb * 4 + q) * Cin + cit * 16 + li], acc[tap][q]);
}

// Wt[tap'][ci][co] = W[26 - tap'][co][ci]: the input gradient is the forward
// convolution of the output gradient with these weights
__global__ void k_wflip(const bf16* __restrict__ W, bf16* __restrict__ Wt, int Cout, int Cin) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int n = 27 * Cout * Cin;
    if (t >= n) return;
    const int ci = t % Cin, co = (t / Cin) % Cout, tap = t / (Cin * Cout);
    Wt[((size_t)(26 - tap) * Cin + ci) * Cout + co] = W[t];
}

__device__ __forceinline__ long long parent_of(long long v, const Geo& fine, int& b, int& z, int& y, int& x) {
    vox_coords(v, fine, b, z, y, x);
    const int D2 = fine.D / 2, H2 = fine.H / 2, W2 = fine.W / 2;
    return (((long long)b * D2 + z / 2) * H2 + y / 2) * W2 + x / 2;
}

// 2x2x2 average pooling (coarse voxel per thread-channel)
__device__ __forceinline__ void pool_one(const BLZ_GLOBAL bf16* __restrict__ S, int Cs, int soff,
                                         BLZ_GLOBAL bf16* __restrict__ Dd, int Cd,
                                         int doff, int C, const Geo& fine, long long ncoarse, long long t) {
    if (t >= ncoarse * C) return;
    // ncoarse C < 2^31 (the pooled view's 32-bit offsets, checked where
    // its convolution is launched): 32-bit divisions
    const int c = (int)((unsigned)t % (unsigned)C);
    const long long vc = (long long)((unsigned)t / (unsigned)C);
    const Geo cg{fine.B, fine.D / 2, fine.H / 2, fine.W / 2};
    int b, z, y, x;
    vox_coords32((unsigned)vc, cg, b, z, y, x);
    float s = 0.0f;
    for (int q = 0; q < 8; ++q) {
        const long long vf = (((long long)b * fine.D + 2 * z + (q >> 2)) * fine.H + 2 * y + ((q >> 1) & 1)) * fine.W +
                             2 * x + (q & 1);
        s += bf2f(S[vf * Cs + soff + c]);
    }
    Dd[vc * Cd + doff + c] = f2bf(s * 0.125f);
}

__global__ void k_pool_fwd(const bf16* __restrict__ S, int Cs, int soff, bf16* __restrict__ Dd, int Cd, int doff,
                           int C, Geo fine, long long ncoarse) {
    pool_one(as_global(S), Cs, soff, as_global(Dd), Cd, doff, C, fine, ncoarse,
             (long long)blockIdx.x * blockDim.x + threadIdx.x);
}

// the pooling of every view in `tab` (X fine, Y coarse)
__global__ void k_pool_multi(const MView* __restrict__ tab, int Cs, int soff, int Cd, int doff, int C) {
    const MView m = tab[blockIdx.z];
    if (mview_idle(m)) return;
    pool_one(as_global(m.X), Cs, soff, as_global(m.Y), Cd, doff, C, m.g, m.n,
             (long long)blockIdx.x * blockDim.x + threadIdx.x);
}

// pooling backward added to the gradient of the pooled activation S, then
// ReLU mask of S: gS = (gS + gP[parent] / 8) * (S > 0)
__global__ void k_pool_bwd(const bf16* __restrict__ gP, int Cp, int poff, bf16* __restrict__ gS, int Cgs, int gsoff,
                           const bf16* __restrict__ S, int Cs, int soff, int C, Geo fine, long long nfine) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nfine * C) return;
    const int c = (int)(t % C);
    const long long v = t / C;
    int b, z, y, x;
    const long long vp = parent_of(v, fine, b, z, y, x);
    const float r = bf2f(gS[v * Cgs + gsoff + c]) + bf2f(gP[vp * Cp + poff + c]) * 0.125f;
    gS[v * Cgs + gsoff + c] = f2bf(bf2f(S[v * Cs + soff + c]) > 0.0f ? r : NET_LEAK * r);
}

// upsampling backward: gB[parent] = sum of the 8 children's gradients, then
// ReLU mask of the coarse activation B
__global__ void k_up_bwd(const bf16* __restrict__ gU, int Cu, int uoff, bf16* __restrict__ gB, int Cb, int boff,
                         const bf16* __restrict__ Bm, int Cbm, int bmoff, int C, Geo fine, long long ncoarse) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ncoarse * C) return;
    // ncoarse C < 2^31 (the pooled view's 32-bit offsets, checked where
    // its convolution is launched): 32-bit divisions
    const int c = (int)((unsigned)t % (unsigned)C);
    const long long vc = (long long)((unsigned)t / (unsigned)C);
    const Geo cg{fine.B, fine.D / 2, fine.H / 2, fine.W / 2};
    int b, z, y, x;
    vox_coords32((unsigned)vc, cg, b, z, y, x);
    float s = 0.0f;
    for (int q = 0; q < 8; ++q) {
        const long long vf = (((long long)b * fine.D + 2 * z + (q >> 2)) * fine.H + 2 * y + ((q >> 1) & 1)) * fine.W +
                             2 * x + (q & 1);
        s += bf2f(gU[vf * Cu + uoff + c]);
    }
    gB[vc * Cb + boff + c] = f2bf(bf2f(Bm[vc * Cbm + bmoff + c]) > 0.0f ? s : NET_LEAK * s);
}

// 1x1 head: out[v] = sum_c X[v][c] w[c] (FP32)
__device__ __forceinline__ void head_one(const BLZ_GLOBAL bf16* __restrict__ X, const float* __restrict__ w,
                                         BLZ_GLOBAL float* __restrict__ out, long long n, long long v) {
    if (v >= n) return;
    float s = 0.0f;
#pragma unroll
    for (int c = 0; c < NET_HEAD; ++c) s = fmaf(bf2f(X[v * NET_HEAD + c]), bf2f(f2bf(w[c])), s);
    out[v] = s;
}

__global__ void k_head_fwd(const bf16* __restrict__ X, const float* __restrict__ w, float* __restrict__ out,
                           long long n) {
    head_one(as_global(X), w, as_global(out), n, (long long)blockIdx.x * blockDim.x + threadIdx.x);
}

// the head of every view in `tab`
__global__ void k_head_multi(const MView* __restrict__ tab, const float* __restrict__ w) {
    const MView m = tab[blockIdx.z];
    if (mview_idle(m)) return;
    head_one(as_global(m.X), w, as_global(m.out), m.n, (long long)blockIdx.x * blockDim.x + threadIdx.x);
}

// loss = sum (out - target)^2 / n; gout = 2 (out - target) / n
__global__ void __launch_bounds__(256) k_mse(const float* __restrict__ out, const float* __restrict__ tg,
                                             float* __restrict__ gout, double* loss, long long n) {
    double s = 0.0;
    const float inv = 1.0f / (float)n;
    for (long long v = (long long)blockIdx.x * 256 + threadIdx.x; v < n; v += (long long)gridDim.x * 256) {
        const float d = out[v] - tg[v];
        gout[v] = 2.0f * d * inv;
        s += (double)d * (double)d;
    }
    for (int o = 32; o >= 1; o >>= 1) s += __shfl_xor(s, o, 64);
    __shared__ double sm[4];
    if ((threadIdx.x & 63) == 0) sm[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) atomicAdd(loss, (sm[0] + sm[1] + sm[2] + sm[3]) / (double)n);
}

// head backward: gX[v][c] = gout[v] w[c] (ReLU mask of X), gw[c] += sum_v gout[v] X[v][c]
__global__ void __launch_bounds__(256) k_head_bwd(const float* __restrict__ gout, const bf16* __restrict__ X,
                                                  const float* __restrict__ w, bf16* __restrict__ gX,
                                                  float* __restrict__ gw, long long n) {
    float acc[NET_HEAD];
#pragma unroll
    for (int c = 0; c < NET_HEAD; ++c) acc[c] = 0.0f;
    for (long long v = (long long)blockIdx.x * 256 + threadIdx.x; v < n; v += (long long)gridDim.x * 256) {
        const float go = gout[v];
#pragma unroll
        for (int c = 0; c < NET_HEAD; ++c) {
            const float xv = bf2f(X[v * NET_HEAD + c]);
            acc[c] = fmaf(go, xv, acc[c]);
            const float gx = go * bf2f(f2bf(w[c]));
            gX[v * NET_HEAD + c] = f2bf(xv > 0.0f ? gx : NET_LEAK * gx);
        }
    }
#pragma unroll
    for (int c = 0; c < NET_HEAD; ++c) {
        float s = acc[c];
        for (int o = 32; o >= 1; o >>= 1) s += __shfl_xor(s, o, 64);
        if ((threadIdx.x & 63) == 0) atomicAdd(&gw[c], s);
    }
}

// global gradient-norm clipping: *scale = min(1, maxnorm / ||G||) (one block)
__global__ void __launch_bounds__(1024) k_gclip(const float* __restrict__ G, size_t n, float maxnorm,
                                                float* __restrict__ scale) {
    double a = 0.0;
    for (size_t t = threadIdx.x; t < n; t += 1024) a += (double)G[t] * (double)G[t];
    for (int o = 32; o >= 1; o >>= 1) a += __shfl_xor(a, o, 64);
    __shared__ double sm[16];
    if ((threadIdx.x & 63) == 0) sm[threadIdx.x >> 6] = a;
    __syncthreads();
    if (threadIdx.x == 0) {
        double tot = 0.0;
        for (int q = 0; q < 16; ++q) tot += sm[q];
        const double nrm = sqrt(tot);
        *scale = (nrm > (double)maxnorm && nrm == nrm) ? (float)((double)maxnorm / nrm) : (nrm == nrm ? 1.0f : 0.0f);
    }
}

__global__ void k_adam(float* __restrict__ P, const float* __restrict__ G, float* __restrict__ M,
                       float* __restrict__ V, size_t n, float lr, float b1, float b2, float eps, float c1, float c2,
                       const float* __restrict__ gscale) {
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const float gg = G[t] * *gscale;
    const float m = b1 * M[t] + (1.0f - b1) * gg;
    const float v = b2 * V[t] + (1.0f - b2) * gg * gg;
    M[t] = m;
    V[t] = v;
    P[t] -= lr * (m / c1) / (sqrtf(v / c2) + eps);
}

__global__ void k_to_bf16(const float* __restrict__ P, bf16* __restrict__ Wb, size_t n) {
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) Wb[t] = f2bf(P[t]);
}

__global__ void k_fill_f32(float* __restrict__ p, size_t n, float v) {
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) p[t] = v;
}

__global__ void k_fill_bf16(bf16* __restrict__ p, size_t n) {
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) p[t] = 0;
}

__global__ void k_fill_d(double* p) { *p = 0.0; }

// x0[v] = (res[v], mask[v], 0 ...) in BF16
__global__ void k_pack_input(const float* __restrict__ res, const float* __restrict__ mask, bf16* __restrict__ x0,
                             long long n) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n * NET_IN_CH) return;
    const int c = (int)(t % NET_IN_CH);
    const long long v = t / NET_IN_CH;
    const float val = c == 0 ? res[v] : (c == 1 ? mask[v] : 0.0f);
    x0[t] = f2bf(val);
}

inline unsigned nblk(long long n, int bs) { return (unsigned)((n + bs - 1) / bs); }

} // namespace

#define NET_HIP(x)                                                   \
    do {                                                             \
        hipError_t e_ = (x);                                         \
        if (e_ != hipSuccess) {                                      \
            blz::set_error(std::string("HIP: ") + hipGetErrorString(e_)); \
            return BLZ_E_HIP;                                        \
        }                                                            \
    } while (0)

// reduction split of a forward convolution: over KS waves where the 16-voxel
// tiles alone would give few waves (the 16^3 and 8^3 levels of a 32^3 input)
static int conv_split(long long nv, int cin, int cout) {
    const long long waves = (nv + 15) / 16 * (cout / 16);
    const int nch = (27 * cin + 31) / 32;
    return (waves <= 1024 && nch >= 12) ? 4 : (waves <= 4096 && nch >= 8) ? 2 : 1;
}

// up (optional): the upsampling target of the output (see k_conv3_k32's U)
static int conv_fwd(Net& n, int layer, TV x, TV y, TV mask, bool relu, const NetDims& d, bool transposed,
                    TV up = TV{nullptr, 0, 0}) {
    const int cin = transposed ? kConv[layer].cout : kConv[layer].cin;
    const int cout = transposed ? kConv[layer].cin : kConv[layer].cout;
    const bf16* w = (transposed ? n.Wt : n.Wb) + n.off[layer];
    const Geo g{d.B, d.D, d.H, d.W};
    const long long nv = d.n();
    // one 16-channel output tile per wave group (grid.y tiles) of 16 voxels
    if (nv * std::max(std::max(x.C, y.C), mask.p ? mask.C : 0) >= (1ll << 31)) {
        set_error("network activations too large for 32-bit offsets");
        return BLZ_E_INVALID;
    }
    const long long tiles = (nv + 15) / 16;
    const int ks = conv_split(nv, cin, cout);
    const dim3 gr(nblk(tiles, 4 / ks), (unsigned)(cout / 16));
#define CONV_KS(CIN, KS)                                                                                           \
    do {                                                                                                           \
        if (up.p)                                                                                                  \
            hipLaunchKernelGGL((k_conv3_k32<CIN, KS, true>), gr, dim3(256), 0, n.stream, x.p, x.C, x.off, w, cout, \
                               y.p, y.C, y.off, mask.p, mask.C, mask.off, relu ? 1 : 0, g, nv, up.p, up.C, up.off); \
        else                                                                                                       \
            hipLaunchKernelGGL((k_conv3_k32<CIN, KS, false>), gr, dim3(256), 0, n.stream, x.p, x.C, x.off, w,     \
                               cout, y.p, y.C, y.off, mask.p, mask.C, mask.off, relu ? 1 : 0, g, nv, up.p, up.C,   \
                               up.off);                                                                            \
    } while (0)
#define CONV_CASE(CIN)                                                                                             \
    case CIN:                                                                                                      \
        if (ks == 4) CONV_KS(CIN, 4);                                                                              \
        else if (ks == 2) CONV_KS(CIN, 2);                                                                         \
        else CONV_KS(CIN, 1);                                                                                      \
        break;
    switch (cin) {
        CONV_CASE(16)
        CONV_CASE(32)
        CONV_CASE(48)
        CONV_CASE(64)
        CONV_CASE(96)
    default: set_error("unsupported channel count"); return BLZ_E_INVALID;
    }
#undef CONV_CASE
#undef CONV_KS
    NET_HIP(hipGetLastError());
    return BLZ_OK;
}

static int conv_dw(Net& n, int layer, TV x, TV gy, const NetDims& d) {
    const int cin = kConv[layer].cin, cout = kConv[layer].cout;
    const Geo g{d.B, d.D, d.H, d.W};
    const long long nv = d.n();
    const int tiles = (cout / 16) * (cin / 16);
    // voxel run per wave: at least 512, at most 32 runs per tile
    long long run = std::max<long long>(512, (nv + 31) / 32);
    run = (run + 15) / 16 * 16;
    const long long chunks = (nv + run - 1) / run;
    const long long waves = chunks * tiles;
    hipLaunchKernelGGL(k_conv3_dw, dim3(nblk(waves, 4)), dim3(256), 0, n.stream, x.p, x.C, x.off, cin, gy.p, gy.C,
                       gy.off, cout, n.G + n.off[layer], g, nv, run);
    NET_HIP(hipGetLastError());
    return BLZ_OK;
}

Net::~Net() {
    if (stream) (void)hipStreamSynchronize(stream);
    if (!view) {
        for (void* p : {(void*)P, (void*)G, (void*)M, (void*)V, (void*)Wb, (void*)Wt, (void*)loss})
            if (p) (void)hipFree(p);
        if (loss_host) (void)hipHostFree(loss_host);
        if (stream) (void)hipStreamDestroy(stream);
    }
    if (act) (void)hipFree(act);
}

int Net::init_view(const Net& src, hipStream_t st) {
    view = true;
    stream = st;
    for (int l = 0; l <= L_NCONV; ++l) off[l] = src.off[l];
    nparam = src.nparam;
    P = src.P;
    Wb = src.Wb;
    Wt = src.Wt;
    return BLZ_OK;
}

int Net::init(unsigned long long seed) {
    NET_HIP(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
    size_t o = 0;
    for (int l = 0; l < L_NCONV; ++l) {
        off[l] = o;
        o += (size_t)27 * kConv[l].cout * kConv[l].cin;
    }
    off[L_NCONV] = o;
    nparam = o + NET_HEAD;
    if (hipMalloc(&P, nparam * 4) != hipSuccess || hipMalloc(&G, nparam * 4) != hipSuccess ||
        hipMalloc(&M, nparam * 4) != hipSuccess || hipMalloc(&V, nparam * 4) != hipSuccess ||
        hipMalloc(&Wb, nparam * 2) != hipSuccess || hipMalloc(&Wt, nparam * 2) != hipSuccess ||
        hipMalloc(&loss, sizeof(double)) != hipSuccess) {
        set_error("hipMalloc failed for network parameters");
        return BLZ_E_NOMEM;
    }
    NET_HIP(hipHostMalloc((void**)&loss_host, sizeof(double), hipHostMallocDefault));
    // He-normal initialisation from a seeded generator (fan-in = 27 x used
    // input channels; the padded input channels of the first layer are 0)
    std::vector<float> h(nparam, 0.0f);
    std::mt19937_64 rng(seed);
    for (int l = 0; l < L_NCONV; ++l) {
        const int cin = kConv[l].cin, cout = kConv[l].cout;
        const int used = l == L_E1A ? 2 : cin;
        std::normal_distribution<float> nd(0.0f, std::sqrt(2.0f / (27.0f * (float)used)));
        for (int tap = 0; tap < 27; ++tap)
            for (int co = 0; co < cout; ++co)
                for (int ci = 0; ci < cin; ++ci)
                    h[off[l] + ((size_t)tap * cout + co) * cin + ci] = ci < used ? nd(rng) : 0.0f;
    }
    std::normal_distribution<float> nh(0.0f, std::sqrt(1.0f / NET_HEAD));
    for (int c = 0; c < NET_HEAD; ++c) h[off[L_NCONV] + c] = nh(rng);
    NET_HIP(hipMemcpy(P, h.data(), nparam * 4, hipMemcpyHostToDevice));
    hipLaunchKernelGGL(k_fill_f32, dim3(nblk((long long)nparam, 256)), dim3(256), 0, stream, M, nparam, 0.0f);
    hipLaunchKernelGGL(k_fill_f32, dim3(nblk((long long)nparam, 256)), dim3(256), 0, stream, V, nparam, 0.0f);
    NET_HIP(hipGetLastError());
    adam_t = 0;
    return refresh_weights();
}

int Net::refresh_weights() {
    hipLaunchKernelGGL(k_to_bf16, dim3(nblk((long long)nparam, 256)), dim3(256), 0, stream, P, Wb, nparam);
    for (int l = 0; l < L_NCONV; ++l) {
        const int n = 27 * kConv[l].cout * kConv[l].cin;
        hipLaunchKernelGGL(k_wflip, dim3(nblk(n, 256)), dim3(256), 0, stream, Wb + off[l], Wt + off[l],
                           kConv[l].cout, kConv[l].cin);
    }
    NET_HIP(hipGetLastError());
    return BLZ_OK;
}

int Net::ensure_dims(const NetDims& d) {
    if (d.D % 4 || d.H % 4 || d.W % 4 || d.B < 1 || d.D < 4 || d.H < 4 || d.W < 4) {
        set_error("network grid must be a multiple of 4 (>= 4) on every axis");
        return BLZ_E_INVALID;
    }
    if (d.B == dims.B && d.D == dims.D && d.H == dims.H && d.W == dims.W) return BLZ_OK;
    const long long n1 = d.n(), n2 = d.half().n(), n3 = d.half().half().n();
    // (buffer, elements) in BF16, then FP32 arrays
    struct Item {
        bf16** p;
        long long n;
    };
    const Item items[] = {{&x0, n1 * 16},   {&a1, n1 * 16},     {&cat1, n1 * 48},  {&p1, n2 * 16},    {&a2, n2 * 32},
                          {&cat2, n2 * 96}, {&p2, n3 * 32},     {&a3, n3 * 64},    {&b3, n3 * 64},    {&a4, n2 * 32},
                          {&b4, n2 * 32},   {&a5, n1 * 16},     {&b5, n1 * 16},    {&g_b5, n1 * 16},  {&g_a5, n1 * 16},
                          {&g_cat1, n1 * 48}, {&g_b4, n2 * 32}, {&g_a4, n2 * 32},  {&g_cat2, n2 * 96}, {&g_b3, n3 * 64},
                          {&g_a3, n3 * 64}, {&g_p2, n3 * 32},   {&g_a2, n2 * 32},  {&g_p1, n2 * 16},  {&g_a1, n1 * 16}};
    size_t bytes = 0;
    for (const Item& it : items) bytes += ((size_t)it.n * 2 + 255) / 256 * 256;
    const size_t fbytes = ((size_t)(n1 + 1) * 4 + 255) / 256 * 256; // + gradient clip scale after out
    bytes += 3 * fbytes;
    if (bytes > act_bytes) {
        if (act) NET_HIP(hipFree(act));
        act = nullptr;
        act_bytes = 0;
        if (hipMalloc(&act, bytes) != hipSuccess) {
            set_error("hipMalloc failed for network activations");
            return BLZ_E_NOMEM;
        }
        act_bytes = bytes;
    }
    char* c = (char*)act;
    for (const Item& it : items) {
        *it.p = (bf16*)c;
        c += ((size_t)it.n * 2 + 255) / 256 * 256;
    }
    out = (float*)c;
    c += fbytes;
    target = (float*)c;
    c += fbytes;
    gout = (float*)c;
    // zero everything once: padded input channels stay 0
    hipLaunchKernelGGL(k_fill_bf16, dim3(nblk((long long)(bytes / 2), 256)), dim3(256), 0, stream, (bf16*)act,
                       bytes / 2);
    NET_HIP(hipGetLastError());
    dims = d;
    return BLZ_OK;
}

// The forward pass as a list of steps over a view's buffers (level 0 is the
// input grid, each level half the one before)
struct FwdStep {
    int kind;  // 0 convolution with leaky ReLU, 1 pooling (x fine -> y coarse), 2 head
    int layer; // convolution layer
    int level; // grid level of the input
    int C;     // pooled channels
    TV x, y, up;
};
constexpr int kFwdSteps = 13;

static void forward_steps(const Net& n, FwdStep* st) {
    const TV none{nullptr, 0, 0};
    int k = 0;
    st[k++] = FwdStep{0, L_E1A, 0, 0, TV{n.x0, 16, 0}, TV{n.a1, 16, 0}, none};
    st[k++] = FwdStep{0, L_E1B, 0, 0, TV{n.a1, 16, 0}, TV{n.cat1, 48, 32}, none};
    st[k++] = FwdStep{1, -1, 0, 16, TV{n.cat1, 48, 32}, TV{n.p1, 16, 0}, none};
    st[k++] = FwdStep{0, L_E2A, 1, 0, TV{n.p1, 16, 0}, TV{n.a2, 32, 0}, none};
    st[k++] = FwdStep{0, L_E2B, 1, 0, TV{n.a2, 32, 0}, TV{n.cat2, 96, 64}, none};
    st[k++] = FwdStep{1, -1, 1, 32, TV{n.cat2, 96, 64}, TV{n.p2, 32, 0}, none};
    st[k++] = FwdStep{0, L_E3A, 2, 0, TV{n.p2, 32, 0}, TV{n.a3, 64, 0}, none};
    // b3 and b4 also written upsampled into the skip concatenations
    st[k++] = FwdStep{0, L_E3B, 2, 0, TV{n.a3, 64, 0}, TV{n.b3, 64, 0}, TV{n.cat2, 96, 0}};
    st[k++] = FwdStep{0, L_D2A, 1, 0, TV{n.cat2, 96, 0}, TV{n.a4, 32, 0}, none};
    st[k++] = FwdStep{0, L_D2B, 1, 0, TV{n.a4, 32, 0}, TV{n.b4, 32, 0}, TV{n.cat1, 48, 0}};
    st[k++] = FwdStep{0, L_D1A, 0, 0, TV{n.cat1, 48, 0}, TV{n.a5, 16, 0}, none};
    st[k++] = FwdStep{0, L_D1B, 0, 0, TV{n.a5, 16, 0}, TV{n.b5, 16, 0}, none};
    st[k++] = FwdStep{2, -1, 0, 0, TV{n.b5, NET_HEAD, 0}, none, none};
}

static void level_dims(const NetDims& d, NetDims* lv) {
    lv[0] = d;
    lv[1] = d.half();
    lv[2] = lv[1].half();
    lv[3] = lv[2].half();
}

int Net::forward() {
    FwdStep st[kFwdSteps];
    forward_steps(*this, st);
    NetDims lv[4];
    level_dims(dims, lv);
    const TV none{nullptr, 0, 0};
    for (const FwdStep& f : st) {
        const NetDims& d = lv[f.level];
        if (f.kind == 0) {
            if (int e = conv_fwd(*this, f.layer, f.x, f.y, none, true, d, false, f.up)) return e;
        } else if (f.kind == 1) {
            const long long nc = lv[f.level + 1].n();
            hipLaunchKernelGGL(k_pool_fwd, dim3(nblk(nc * f.C, 256)), dim3(256), 0, stream, f.x.p, f.x.C, f.x.off,
                               f.y.p, f.y.C, f.y.off, f.C, Geo{d.B, d.D, d.H, d.W}, nc);
        } else {
            hipLaunchKernelGGL(k_head_fwd, dim3(nblk(d.n(), 256)), dim3(256), 0, stream, f.x.p, P + off[L_NCONV], out,
                               d.n());
        }
    }
    NET_HIP(hipGetLastError());
    return BLZ_OK;
}

NetMulti::~NetMulti() {
    if (tab) (void)hipFree(tab);
}

int NetMulti::build(const std::vector<Net*>& views, const std::vector<const int*>& stop) {
    const int nv = (int)views.size();
    if (nv < 1 || nv > 65535 || (int)stop.size() != nv) {
        set_error("multi-view forward: 1 to 65535 views, one stop flag each");
        return BLZ_E_INVALID;
    }
    const Net& v0 = *views[0];
    for (const Net* v : views)
        if (v->Wb != v0.Wb || v->P != v0.P || !v->x0) {
            set_error("multi-view forward: views of one network with activations");
            return BLZ_E_INVALID;
        }
    std::vector<FwdStep> steps((size_t)nv * kFwdSteps);
    std::vector<NetDims> lv((size_t)nv * 4);
    for (int q = 0; q < nv; ++q) {
        forward_steps(*views[q], &steps[(size_t)q * kFwdSteps]);
        level_dims(views[q]->dims, &lv[(size_t)q * 4]);
    }
    std::vector<MView> t;
    std::vector<Launch> pl;
    for (int k = 0; k < kFwdSteps; ++k) {
        const FwdStep& f0 = steps[k];
        const int cin = f0.kind == 0 ? kConv[f0.layer].cin : 0, cout = f0.kind == 0 ? kConv[f0.layer].cout : 0;
        // convolutions: one launch per reduction split in use
        for (int ks : {4, 2, 1}) {
            if (f0.kind != 0 && ks != 1) continue;
            // one wave per tile (the splits run in it) where the launch's
            // views give at least 16384 tile waves, 64 per CU
            long long tile_waves = 0;
            if (f0.kind == 0)
                for (int q = 0; q < nv; ++q) {
                    const long long n = lv[(size_t)q * 4 + steps[(size_t)q * kFwdSteps + k].level].n();
                    if (conv_split(n, cin, cout) == ks) tile_waves += (n + 15) / 16 * (cout / 16);
                }
            const bool one = tile_waves >= 16384;
            Launch L{f0.kind, f0.layer, ks, f0.C, one ? 1 : 0, f0.x, f0.y, f0.up,
                     f0.kind == 0 ? v0.Wb + v0.off[f0.layer] : nullptr, (int)t.size(), 0, 0u};
            for (int q = 0; q < nv; ++q) {
                const FwdStep& f = steps[(size_t)q * kFwdSteps + k];
                const NetDims& d = lv[(size_t)q * 4 + f.level];
                MView m;
                std::memset(&m, 0, sizeof m); // padding too: the table is compared bytewise
                m.X = f.x.p;
                m.Y = f.y.p;
                m.U = f.up.p;
                m.out = views[q]->out;
                m.g = Geo{d.B, d.D, d.H, d.W};
                m.n = d.n();
                m.stop = stop[q];
                if (f.kind == 0) {
                    if (conv_split(d.n(), cin, cout) != ks) continue;
                    if (d.n() * std::max(f.x.C, f.y.C) >= (1ll << 31)) {
                        set_error("network activations too large for 32-bit offsets");
                        return BLZ_E_INVALID;
                    }
                    m.nblk = nblk((d.n() + 15) / 16, one ? 4 : 4 / ks);
                } else if (f.kind == 1) {
                    m.n = lv[(size_t)q * 4 + f.level + 1].n();
                    m.nblk = nblk(m.n * f.C, 256);
                } else {
                    m.nblk = nblk(d.n(), 256);
                }
                t.push_back(m);
                L.blocks = std::max(L.blocks, m.nblk);
            }
            L.count = (int)t.size() - L.first;
            if (L.count) pl.push_back(L);
        }
    }
    const size_t bytes = t.size() * sizeof(MView);
    const unsigned char* raw = (const unsigned char*)t.data();
    if (cur.size() != bytes || std::memcmp(cur.data(), raw, bytes) != 0) {
        if (bytes > tab_bytes) {
            if (tab) NET_HIP(hipFree(tab));
            tab = nullptr;
            tab_bytes = 0;
            NET_HIP(hipMalloc(&tab, bytes));
            tab_bytes = bytes;
        }
        NET_HIP(hipMemcpy(tab, raw, bytes, hipMemcpyHostToDevice));
        cur.assign(raw, raw + bytes);
    }
    plan = pl;
    head = v0.P + v0.off[L_NCONV];
    return BLZ_OK;
}

template <int CIN>
static void conv_multi(const NetMulti::Launch& L, const MView* tb, hipStream_t st) {
    const int cout = kConv[L.layer].cout;
    const dim3 gr(L.blocks, (unsigned)(cout / 16), (unsigned)L.count);
#define CONV_MULTI_ONE(KS, UP, ONE)                                                                                \
    hipLaunchKernelGGL((k_conv3_multi<CIN, KS, UP, ONE>), gr, dim3(256), 0, st, tb, L.x.C, L.x.off, L.w, cout,     \
                       L.y.C, L.y.off, L.up.C, L.up.off)
    // (with KS = 1 the two modes are the same code)
#define CONV_MULTI(KS, UP)                                                                                         \
    do {                                                                                                           \
        if (KS > 1 && L.one) CONV_MULTI_ONE(KS, UP, true);                                                         \
        else CONV_MULTI_ONE(KS, UP, false);                                                                        \
    } while (0)
    if (L.up.p) {
        // the decoder's upsampled layers take 32 or 64 input channels
        if constexpr (CIN == 32 || CIN == 64) {
            if (L.ks == 4) CONV_MULTI(4, true);
            else if (L.ks == 2) CONV_MULTI(2, true);
            else CONV_MULTI(1, true);
        }
    } else if (L.ks == 4) {
        CONV_MULTI(4, false);
    } else if (L.ks == 2) {
        CONV_MULTI(2, false);
    } else {
        CONV_MULTI(1, false);
    }
#undef CONV_MULTI
#undef CONV_MULTI_ONE
}

int NetMulti::forward(hipStream_t st) const {
    const MView* t = (const MView*)tab;
    for (const Launch& L : plan) {
        const MView* tb = t + L.first;
        if (L.kind == 0) {
            const int cin = kConv[L.layer].cin;
            if (L.up.p && cin != 32 && cin != 64) {
                set_error("multi-view forward: no upsampled layer of this width");
                return BLZ_E_INVALID;
            }
            switch (cin) {
            case 16: conv_multi<16>(L, tb, st); break;
            case 32: conv_multi<32>(L, tb, st); break;
            case 48: conv_multi<48>(L, tb, st); break;
            case 64: conv_multi<64>(L, tb, st); break;
            case 96: conv_multi<96>(L, tb, st); break;
            default: set_error("unsupported channel count"); return BLZ_E_INVALID;
            }
        } else if (L.kind == 1) {
            hipLaunchKernelGGL(k_pool_multi, dim3(L.blocks, 1, (unsigned)L.count), dim3(256), 0, st, tb, L.x.C,
                               L.x.off, L.y.C, L.y.off, L.C);
        } else {
            hipLaunchKernelGGL(k_head_multi, dim3(L.blocks, 1, (unsigned)L.count), dim3(256), 0, st, tb, head);
        }
    }
    NET_HIP(hipGetLastError());
    return BLZ_OK;
}

int Net::loss_and_grad() {
    const long long n1 = dims.n();
    hipLaunchKernelGGL(k_fill_d, dim3(1), dim3(1), 0, stream, loss);
    hipLaunchKernelGGL(k_mse, dim3((unsigned)std::min<long long>(nblk(n1, 256), 1024)), dim3(256), 0, stream, out,
                       target, gout, loss, n1);
    NET_HIP(hipGetLastError());
    return BLZ_OK;
}

int Net::backward() {
    const NetDims d1 = dims, d2 = d1.half(), d3 = d2.half();
    const Geo g1{d1.B, d1.D, d1.H, d1.W}, g2{d2.B, d2.D, d2.H, d2.W};
    const long long n1 = d1.n(), n2 = d2.n(), n3 = d3.n();
    const TV none{nullptr, 0, 0};
    int e;
    hipLaunchKernelGGL(k_fill_f32, dim3(nblk((long long)nparam, 256)), dim3(256), 0, stream, G, nparam, 0.0f);
    // head (ReLU mask of b5 folded in)
    hipLaunchKernelGGL(k_head_bwd, dim3((unsigned)std::min<long long>(nblk(n1, 256), 512)), dim3(256), 0, stream,
                       gout, b5, P + off[L_NCONV], g_b5, G + off[L_NCONV], n1);
    NET_HIP(hipGetLastError());
    // d1b: a5 -> b5
    if ((e = conv_dw(*this, L_D1B, TV{a5, 16, 0}, TV{g_b5, 16, 0}, d1))) return e;
    if ((e = conv_fwd(*this, L_D1B, TV{g_b5, 16, 0}, TV{g_a5, 16, 0}, TV{a5, 16, 0}, false, d1, true))) return e;
    // d1a: cat1 -> a5; input gradient over all 48 channels (unmasked)
    if ((e = conv_dw(*this, L_D1A, TV{cat1, 48, 0}, TV{g_a5, 16, 0}, d1))) return e;
    if ((e = conv_fwd(*this, L_D1A, TV{g_a5, 16, 0}, TV{g_cat1, 48, 0}, none, false, d1, true))) return e;
    // upsampled b4 (channels 0..31 of cat1) -> b4, ReLU mask of b4
    hipLaunchKernelGGL(k_up_bwd, dim3(nblk(n2 * 32, 256)), dim3(256), 0, stream, g_cat1, 48, 0, g_b4, 32, 0, b4, 32,
                       0, 32, g1, n2);
    // d2b: a4 -> b4
    if ((e = conv_dw(*this, L_D2B, TV{a4, 32, 0}, TV{g_b4, 32, 0}, d2))) return e;
    if ((e = conv_fwd(*this, L_D2B, TV{g_b4, 32, 0}, TV{g_a4, 32, 0}, TV{a4, 32, 0}, false, d2, true))) return e;
    // d2a: cat2 -> a4
    if ((e = conv_dw(*this, L_D2A, TV{cat2, 96, 0}, TV{g_a4, 32, 0}, d2))) return e;
    if ((e = conv_fwd(*this, L_D2A, TV{g_a4, 32, 0}, TV{g_cat2, 96, 0}, none, false, d2, true))) return e;
    // upsampled b3 (channels 0..63 of cat2) -> b3, ReLU mask of b3
    hipLaunchKernelGGL(k_up_bwd, dim3(nblk(n3 * 64, 256)), dim3(256), 0, stream, g_cat2, 96, 0, g_b3, 64, 0, b3, 64,
                       0, 64, g2, n3);
    // e3b: a3 -> b3
    if ((e = conv_dw(*this, L_E3B, TV{a3, 64, 0}, TV{g_b3, 64, 0}, d3))) return e;
    if ((e = conv_fwd(*this, L_E3B, TV{g_b3, 64, 0}, TV{g_a3, 64, 0}, TV{a3, 64, 0}, false, d3, true))) return e;
    // e3a: p2 -> a3
    if ((e = conv_dw(*this, L_E3A, TV{p2, 32, 0}, TV{g_a3, 64, 0}, d3))) return e;
    if ((e = conv_fwd(*this, L_E3A, TV{g_a3, 64, 0}, TV{g_p2, 32, 0}, none, false, d3, true))) return e;
    // pool of s2 (channels 64..95 of cat2): add into s2's gradient, ReLU mask of s2
    hipLaunchKernelGGL(k_pool_bwd, dim3(nblk(n2 * 32, 256)), dim3(256), 0, stream, g_p2, 32, 0, g_cat2, 96, 64, cat2,
                       96, 64, 32, g2, n2);
    // e2b: a2 -> s2
    if ((e = conv_dw(*this, L_E2B, TV{a2, 32, 0}, TV{g_cat2, 96, 64}, d2))) return e;
    if ((e = conv_fwd(*this, L_E2B, TV{g_cat2, 96, 64}, TV{g_a2, 32, 0}, TV{a2, 32, 0}, false, d2, true))) return e;
    // e2a: p1 -> a2
    if ((e = conv_dw(*this, L_E2A, TV{p1, 16, 0}, TV{g_a2, 32, 0}, d2))) return e;
    if ((e = conv_fwd(*this, L_E2A, TV{g_a2, 32, 0}, TV{g_p1, 16, 0}, none, false, d2, true))) return e;
    // pool of s1 (channels 32..47 of cat1)
    hipLaunchKernelGGL(k_pool_bwd, dim3(nblk(n1 * 16, 256)), dim3(256), 0, stream, g_p1, 16, 0, g_cat1, 48, 32, cat1,
                       48, 32, 16, g1, n1);
    // e1b: a1 -> s1
    if ((e = conv_dw(*this, L_E1B, TV{a1, 16, 0}, TV{g_cat1, 48, 32}, d1))) return e;
    if ((e = conv_fwd(*this, L_E1B, TV{g_cat1, 48, 32}, TV{g_a1, 16, 0}, TV{a1, 16, 0}, false, d1, true))) return e;
    // e1a: x0 -> a1 (no input gradient)
    if ((e = conv_dw(*this, L_E1A, TV{x0, 16, 0}, TV{g_a1, 16, 0}, d1))) return e;
    NET_HIP(hipGetLastError());
    return BLZ_OK;
}

int Net::adam(float lr, float b1, float b2, float eps) {
    ++adam_t;
    const float c1 = 1.0f - std::pow(b1, (float)adam_t), c2 = 1.0f - std::pow(b2, (float)adam_t);
    float* gscale = out + dims.n(); // one float after the prediction (see ensure_dims)
    hipLaunchKernelGGL(k_gclip, dim3(1), dim3(1024), 0, stream, G, nparam, 1.0f, gscale);
    hipLaunchKernelGGL(k_adam, dim3(nblk((long long)nparam, 256)), dim3(256), 0, stream, P, G, M, V, nparam, lr, b1,
                       b2, eps, c1, c2, gscale);
    NET_HIP(hipGetLastError());
    return refresh_weights();
}

int net_pack_input(Net& n, const float* res, const float* mask, hipStream_t st) {
    const long long n1 = n.dims.n();
    hipLaunchKernelGGL(k_pack_input, dim3(nblk(n1 * NET_IN_CH, 256)), dim3(256), 0, st, res, mask, n.x0, n1);
    NET_HIP(hipGetLastError());
    return BLZ_OK;
}

} // namespace blz

// ---- C ABI of the network (include/blaze.h) ----
using blz::Net;

extern "C" {

int blz_net_create(unsigned long long seed, blz_net** out) {
    if (!out) {
        blz::set_error("blz_net_create: null argument");
        return BLZ_E_INVALID;
    }
    *out = nullptr;
    blz_net* h = new (std::nothrow) blz_net();
    if (!h) {
        blz::set_error("out of host memory");
        return BLZ_E_NOMEM;
    }
    if (int e = h->n.init(seed)) {
        delete h;
        return e;
    }
    *out = h;
    return BLZ_OK;
}

int blz_net_destroy(blz_net* h) {
    delete h;
    return BLZ_OK;
}

int blz_net_num_params(blz_net* h, long long* n) {
    if (!h || !n) {
        blz::set_error("null argument");
        return BLZ_E_INVALID;
    }
    *n = (long long)h->n.nparam;
    return BLZ_OK;
}

int blz_net_get_params(blz_net* h, float* host) {
    if (!h || !host) {
        blz::set_error("null argument");
        return BLZ_E_INVALID;
    }
    NET_HIP(hipStreamSynchronize(h->n.stream));
    NET_HIP(hipMemcpy(host, h->n.P, h->n.nparam * 4, hipMemcpyDeviceToHost));
    return BLZ_OK;
}

int blz_net_set_params(blz_net* h, const float* host) {
    if (!h || !host) {
        blz::set_error("null argument");
        return BLZ_E_INVALID;
    }
    NET_HIP(hipStreamSynchronize(h->n.stream));
    NET_HIP(hipMemcpy(h->n.P, host, h->n.nparam * 4, hipMemcpyHostToDevice));
    if (int e = h->n.refresh_weights()) return e;
    NET_HIP(hipStreamSynchronize(h->n.stream));
    return BLZ_OK;
}

int blz_net_get_grads(blz_net* h, float* host) {
    if (!h || !host) {
        blz::set_error("null argument");
        return BLZ_E_INVALID;
    }
    NET_HIP(hipStreamSynchronize(h->n.stream));
    NET_HIP(hipMemcpy(host, h->n.G, h->n.nparam * 4, hipMemcpyDeviceToHost));
    return BLZ_OK;
}

static int net_load(Net& n, const float* in2, const float* target, const int dims[4]) {
    if (!in2 || !dims) {
        blz::set_error("null argument");
        return BLZ_E_INVALID;
    }
    const blz::NetDims d{dims[0], dims[1], dims[2], dims[3]};
    if (int e = n.ensure_dims(d)) return e;
    const long long nv = d.n();
    // staging: input channels (2 x nv) into out/gout, then packed into x0
    NET_HIP(hipMemcpyAsync(n.out, in2, nv * 4, hipMemcpyHostToDevice, n.stream));
    NET_HIP(hipMemcpyAsync(n.gout, in2 + nv, nv * 4, hipMemcpyHostToDevice, n.stream));
    if (int e = blz::net_pack_input(n, n.out, n.gout, n.stream)) return e;
    if (target) NET_HIP(hipMemcpyAsync(n.target, target, nv * 4, hipMemcpyHostToDevice, n.stream));
    return BLZ_OK;
}

int blz_net_forward(blz_net* h, const float* in2, const int dims[4], float* out) {
    if (!h || !out) {
        blz::set_error("null argument");
        return BLZ_E_INVALID;
    }
    Net& n = h->n;
    if (int e = net_load(n, in2, nullptr, dims)) return e;
    if (int e = n.forward()) return e;
    NET_HIP(hipMemcpyAsync(out, n.out, n.dims.n() * 4, hipMemcpyDeviceToHost, n.stream));
    NET_HIP(hipStreamSynchronize(n.stream));
    return BLZ_OK;
}

int blz_net_train_step(blz_net* h, const float* in2, const float* target, const int dims[4], float lr, int update,
                       double* loss) {
    if (!h || !target) {
        blz::set_error("null argument");
        return BLZ_E_INVALID;
    }
    Net& n = h->n;
    if (int e = net_load(n, in2, target, dims)) return e;
    if (int e = n.forward()) return e;
    if (int e = n.loss_and_grad()) return e;
    if (int e = n.backward()) return e;
    if (update)
        if (int e = n.adam(lr, 0.9f, 0.999f, 1e-8f)) return e;
    NET_HIP(hipMemcpyAsync(n.loss_host, n.loss, sizeof(double), hipMemcpyDeviceToHost, n.stream));
    NET_HIP(hipStreamSynchronize(n.stream));
    if (loss) *loss = *n.loss_host;
    return BLZ_OK;
}

} // extern "C"
