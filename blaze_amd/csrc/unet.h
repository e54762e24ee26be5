// unet.h — the learned warm start's network: a 3-level 3D U-Net (channels
// 16/32/64, BF16 weights and activations, FP32 accumulation on MFMA) that
// maps a coarse-grid (residual, solid mask) pair to a coarse pressure
// correction. No reference counterpart: the reference solve always starts
// from the previous pressure (gs_solve_kernel, src/FluidSolver.cu:517-599).
//
// Layout: activations are channels-last, element (voxel v, channel c) of a
// view at p[v * C + off + c]; voxels in x-fastest order over (batch, z, y, x).
// Convolution weights are [tap][cout][cin] (tap = 9 dz + 3 dy + dx over the
// 3x3x3 stencil, offsets -1..1). Network (bias-free, so it commutes with a
// positive scaling of its input):
//   x0(16: residual, mask, 14 zero) -e1a-> a1(16) -e1b-> s1(16)
//   s1 -pool-> p1(16) -e2a-> a2(32) -e2b-> s2(32)
//   s2 -pool-> p2(32) -e3a-> a3(64) -e3b-> b3(64)
//   [up(b3), s2](96) -d2a-> a4(32) -d2b-> b4(32)
//   [up(b4), s1](48) -d1a-> a5(16) -d1b-> b5(16) -1x1-> out(1, FP32)
// with a leaky ReLU (slope NET_LEAK below 0) after every 3x3x3
// convolution, 2x2x2 average pooling and nearest-neighbour 2x upsampling.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <vector>

#include "common.h"

namespace blz {

typedef unsigned short bf16;

struct TV {
    bf16* p;
    int C;   // channels per voxel in the buffer
    int off; // first channel of the view
};

struct NetDims {
    int B, D, H, W; // level-1 grid (D, H, W multiples of 4)
    long long n() const { return (long long)B * D * H * W; }
    NetDims half() const { return NetDims{B, D / 2, H / 2, W / 2}; }
};

enum { L_E1A = 0, L_E1B, L_E2A, L_E2B, L_E3A, L_E3B, L_D2A, L_D2B, L_D1A, L_D1B, L_NCONV };

struct ConvSpec {
    int cin, cout;
};
constexpr ConvSpec kConv[L_NCONV] = {{16, 16}, {16, 16}, {16, 32}, {32, 32}, {32, 64},
                                     {64, 64}, {96, 32}, {32, 32}, {48, 16}, {16, 16}};
constexpr int NET_IN_CH = 16;  // x0 channels (2 used)
constexpr int NET_HEAD = 16;   // 1x1 head inputs
constexpr float NET_LEAK = 0.1f; // leaky ReLU slope (keeps f(a x) = a f(x), a > 0)

struct Net {
    hipStream_t stream = nullptr;
    // parameters: the convolutions' [27][cout][cin] blocks, then the head
    size_t off[L_NCONV + 1] = {};
    size_t nparam = 0;
    float* P = nullptr;  // FP32 master copy
    float* G = nullptr;  // gradient
    float* M = nullptr;  // Adam first moment
    float* V = nullptr;  // Adam second moment
    bf16* Wb = nullptr;  // BF16 weights (forward)
    bf16* Wt = nullptr;  // BF16 weights flipped and transposed (input gradients)
    int adam_t = 0;
    // activations and gradients for the current dims
    NetDims dims{0, 0, 0, 0};
    void* act = nullptr;
    size_t act_bytes = 0;
    bf16 *x0 = nullptr, *a1 = nullptr, *cat1 = nullptr, *p1 = nullptr, *a2 = nullptr, *cat2 = nullptr,
         *p2 = nullptr, *a3 = nullptr, *b3 = nullptr, *a4 = nullptr, *b4 = nullptr, *a5 = nullptr, *b5 = nullptr;
    bf16 *g_b5 = nullptr, *g_a5 = nullptr, *g_cat1 = nullptr, *g_b4 = nullptr, *g_a4 = nullptr,
         *g_cat2 = nullptr, *g_b3 = nullptr, *g_a3 = nullptr, *g_p2 = nullptr, *g_a2 = nullptr,
         *g_p1 = nullptr, *g_a1 = nullptr;
    float* out = nullptr;    // FP32 prediction, n voxels
    float* target = nullptr; // FP32 target, n voxels
    float* gout = nullptr;   // d loss / d out
    double* loss = nullptr;  // device accumulator (sum of squared errors)
    double* loss_host = nullptr;

    bool view = false; // parameters borrowed from another Net

    ~Net();
    int init(unsigned long long seed);
    int init_view(const Net& src, hipStream_t st); // share src's weights
    int ensure_dims(const NetDims& d);
    int refresh_weights(); // BF16 copies from the FP32 master
    int forward();         // x0 -> out (stream-ordered)
    int loss_and_grad();   // out, target -> gout, *loss (mean squared error)
    int backward();        // gout -> G (overwritten)
    int adam(float lr, float b1, float b2, float eps);
};

// The forwards of several views of one network (the same weights, each view
// on its own grid with its own activations) as one sequence of launches: a
// layer is one launch per reduction split in use, over a device table of
// the views. Each view computes what its own forward() computes, bit for
// bit (the same per-workgroup code and reduction split). A view whose stop
// flag is set when a launch runs is skipped by it (its output is then
// stale). Rebuilding with other views may move the table: graphs captured
// from forward() hold its address.
struct NetMulti {
    struct Launch {
        int kind;          // 0 convolution, 1 pooling, 2 head
        int layer, ks, C;  // convolution layer and reduction split; pooled channels
        int one;           // each wave runs a whole tile's splits (enough tiles in the launch)
        TV x, y, up;       // channel layouts (the pointers are per view, in the table)
        const bf16* w;     // convolution weights
        int first, count;  // views [first, first + count) of the table
        unsigned blocks;   // grid x extent (the largest view's)
    };
    std::vector<Launch> plan;
    void* tab = nullptr; // device table
    size_t tab_bytes = 0;
    std::vector<unsigned char> cur; // its content
    const float* head = nullptr;

    ~NetMulti();
    int build(const std::vector<Net*>& views, const std::vector<const int*>& stop);
    int forward(hipStream_t st) const;
};

// launchers used by the warm start (stream-ordered)
int net_pack_input(Net& n, const float* res, const float* mask, hipStream_t st);

} // namespace blz

struct blz_net {
    blz::Net n;
};
