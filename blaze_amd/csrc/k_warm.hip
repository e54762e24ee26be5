// k_warm.hip — learned warm start of the pressure solve (no reference
// counterpart: gs_solve_kernel, src/FluidSolver.cu:517-599, always starts
// from the previous pressure).
//
// Before the sweeps of a residual-driven solve:
//  1. restrict the residual the previous pressure leaves,
//       r = scale*rhs + sum_nbr p - n p        (scale = rho dx^2 / dt)
//     to a coarser grid, F = 4 or more (doubled until at most 32 coarse cells
//     per axis, the scale the network was trained at): sum over each F^3
//     block / F, which is the right-hand side of the coarse Poisson problem
//     for a smooth error; with a solid mask (1 where a block touches a domain
//     wall or lies outside);
//  2. normalise by s = RMS of the coarse r and run the U-Net (bias-free with
//     leaky ReLUs, so its output scales with s);
//  3. save p, prolongate the prediction trilinearly to e and add alpha e,
//     alpha = <r, e - mean e> / <e, A e> (A x = n x - sum_nbr x, FP64
//     sums): the scale minimising the error's energy norm, so the network
//     only has to get the shape of the correction right and no step can
//     raise the A-norm of the error;
//  4. the judge: the corrected pressure (home planes) and the uncorrected one
//     (saved to the spare planes) both get the next batch of Gauss-Seidel
//     sweeps, and the one with the lower FP64 residual continues -- on a
//     rejection the solve continues exactly (bitwise) as without the
//     prediction, having spent one batch of sweeps twice. The sweeps then
//     run to the same tolerance as without it.
// All decisions are taken on the device (no host round trip).
#include "solver.h"

namespace blz {

namespace {

constexpr unsigned BS = 256;

__device__ __forceinline__ bf16 f2bf_w(float f) {
    const unsigned u = __float_as_uint(f);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (bf16)((u >> 16) | 0x40u);
    return (bf16)((u + 0x7fffu + ((u >> 16) & 1u)) >> 16);
}

__device__ __forceinline__ float pget(const RB& p, const RBL& L, int i, int j, int k) {
    return p.c[(i + j) & 1][L.at(i, j, k)];
}

// thread per network cell (I, J, K)
__global__ void __launch_bounds__(BS) k_ws_restrict(RB p, RB rhs, RBL L, int W, int H, int D, float scale, int F,
                                                    int Wc, int Hc, int Dc, int Wn, int Hn, int Dn,
                                                    float* __restrict__ res, float* __restrict__ mask,
                                                    const int* __restrict__ stop) {
    const int t = blockIdx.x * BS + threadIdx.x;
    if (t >= Wn * Hn * Dn || *stop) return;
    const int I = t % Wn, J = (t / Wn) % Hn, K = t / (Wn * Hn);
    if (I >= Wc || J >= Hc || K >= Dc) {
        res[t] = 0.0f;
        mask[t] = 1.0f;
        return;
    }
    float sum = 0.0f;
    bool wall = false;
    for (int k = F * K; k < min(F * K + F, D); ++k)
        for (int j = F * J; j < min(F * J + F, H); ++j)
            for (int i = F * I; i < min(F * I + F, W); ++i) {
                const float pc = pget(p, L, i, j, k);
                float r = scale * pget(rhs, L, i, j, k);
                float n = 0.0f;
                if (i > 0) { r += pget(p, L, i - 1, j, k); n += 1.0f; }
                if (i < W - 1) { r += pget(p, L, i + 1, j, k); n += 1.0f; }
                if (j > 0) { r += pget(p, L, i, j - 1, k); n += 1.0f; }
                if (j < H - 1) { r += pget(p, L, i, j + 1, k); n += 1.0f; }
                if (k > 0) { r += pget(p, L, i, j, k - 1); n += 1.0f; }
                if (k < D - 1) { r += pget(p, L, i, j, k + 1); n += 1.0f; }
                sum += r - n * pc;
                wall |= n < 6.0f;
            }
    const float rc = sum / (float)F;
    res[t] = rc;
    mask[t] = wall ? 1.0f : 0.0f;
}

// input scale s = RMS of the coarse residual over the real coarse cells (one
// block, fixed order: deterministic); stored as float bits in *amax
__global__ void __launch_bounds__(1024) k_ws_rms(const float* __restrict__ res, int n, int nreal,
                                                 unsigned* __restrict__ amax, const int* __restrict__ stop) {
    if (*stop) return;
    double a = 0.0;
    for (int t = threadIdx.x; t < n; t += 1024) a += (double)res[t] * (double)res[t];
    for (int o = 32; o >= 1; o >>= 1) a += __shfl_xor(a, o, 64);
    __shared__ double sm[16];
    if ((threadIdx.x & 63) == 0) sm[threadIdx.x >> 6] = a;
    __syncthreads();
    if (threadIdx.x == 0) {
        double tot = 0.0;
        for (int q = 0; q < 16; ++q) tot += sm[q];
        const float r = (float)sqrt(tot / (double)max(nreal, 1));
        *amax = __float_as_uint(r);
    }
}

// network input: (r / s, mask) in BF16, 16 channels (s = input scale)
__global__ void k_ws_pack(const float* __restrict__ res, const float* __restrict__ mask,
                          const unsigned* __restrict__ amax, bf16* __restrict__ x0, int n) {
    const int t = blockIdx.x * BS + threadIdx.x;
    if (t >= n * NET_IN_CH) return;
    const int c = t % NET_IN_CH, v = t / NET_IN_CH;
    const float s = __uint_as_float(*amax);
    float val = 0.0f;
    if (c == 0) val = s > 0.0f ? res[v] / s : 0.0f;
    if (c == 1) val = mask[v];
    x0[t] = f2bf_w(val);
}

// FP32 normalised input for data generation: (r / s, mask)
__global__ void k_ws_export(const float* __restrict__ res, const float* __restrict__ mask,
                            const unsigned* __restrict__ amax, float* __restrict__ out2, int n) {
    const int t = blockIdx.x * BS + threadIdx.x;
    if (t >= n) return;
    const float s = __uint_as_float(*amax);
    out2[t] = s > 0.0f ? res[t] / s : 0.0f;
    out2[n + t] = mask[t];
}

// e = trilinear(prediction) at every fine cell, natural order, into tmp
__global__ void __launch_bounds__(BS) k_ws_prolong(float* __restrict__ e, int W, int H, int D,
                                                   const float* __restrict__ pred, float invF, int Wc, int Hc, int Dc,
                                                   int Wn, int Hn, const SolveCtl* __restrict__ c) {
    const int t = blockIdx.x * BS + threadIdx.x;
    if (t >= W * H * D || c->done) return;
    const int i = t % W, j = (t / W) % H, k = t / (W * H);
    auto axis = [invF](int q, int n, int& a0, int& a1, float& w) {
        float f = ((float)q + 0.5f) * invF - 0.5f;
        f = fminf(fmaxf(f, 0.0f), (float)(n - 1));
        a0 = (int)f;
        a1 = min(a0 + 1, n - 1);
        w = f - (float)a0;
    };
    int x0, x1, y0, y1, z0, z1;
    float wx, wy, wz;
    axis(i, Wc, x0, x1, wx);
    axis(j, Hc, y0, y1, wy);
    axis(k, Dc, z0, z1, wz);
    auto P = [&](int x, int y, int z) { return pred[((size_t)z * Hn + y) * Wn + x]; };
    const float c00 = P(x0, y0, z0) * (1.0f - wx) + P(x1, y0, z0) * wx;
    const float c10 = P(x0, y1, z0) * (1.0f - wx) + P(x1, y1, z0) * wx;
    const float c01 = P(x0, y0, z1) * (1.0f - wx) + P(x1, y0, z1) * wx;
    const float c11 = P(x0, y1, z1) * (1.0f - wx) + P(x1, y1, z1) * wx;
    const float c0 = c00 * (1.0f - wy) + c10 * wy, c1 = c01 * (1.0f - wy) + c11 * wy;
    const float v = c0 * (1.0f - wz) + c1 * wz;
    e[t] = (v == v && fabsf(v) <= 3.4e38f) ? v : 0.0f;
}

// scale of the correction minimising the error's energy norm: with A x =
// n x - sum_nbr x (symmetric, positive semi-definite, constants in its null
// space) and r = scale*rhs - A p, the error p* - p - alpha e has the least
// A-norm at alpha = <r, e'> / <e, A e>, e' = e minus its mean (a constant
// changes no residual). Unlike the residual-norm fit, this step can never
// raise the A-norm of the error, i.e. it cannot add smooth error that the
// sweeps remove slowly. Per-block partial sums (FP64) of <r,e>, <e,Ae>,
// sum r and sum e.
__global__ void __launch_bounds__(BS) k_ws_dots(RB p, RB rhs, RBL L, int W, int H, int D, float scale,
                                                const float* __restrict__ e, double* __restrict__ partial,
                                                const SolveCtl* __restrict__ c) {
    if (c->done) return;
    double v[4] = {0.0, 0.0, 0.0, 0.0};
    const int n = W * H * D;
    for (int t = blockIdx.x * BS + threadIdx.x; t < n; t += gridDim.x * BS) {
        const int i = t % W, j = (t / W) % H, k = t / (W * H);
        const double pc = pget(p, L, i, j, k);
        double r = (double)scale * pget(rhs, L, i, j, k);
        const double ec = e[t];
        double ae = 0.0, cnt = 0.0;
        if (i > 0) { r += pget(p, L, i - 1, j, k); ae -= e[t - 1]; cnt += 1.0; }
        if (i < W - 1) { r += pget(p, L, i + 1, j, k); ae -= e[t + 1]; cnt += 1.0; }
        if (j > 0) { r += pget(p, L, i, j - 1, k); ae -= e[t - W]; cnt += 1.0; }
        if (j < H - 1) { r += pget(p, L, i, j + 1, k); ae -= e[t + W]; cnt += 1.0; }
        if (k > 0) { r += pget(p, L, i, j, k - 1); ae -= e[t - W * H]; cnt += 1.0; }
        if (k < D - 1) { r += pget(p, L, i, j, k + 1); ae -= e[t + W * H]; cnt += 1.0; }
        r -= cnt * pc;
        ae += cnt * ec;
        v[0] += r * ec;
        v[1] += ec * ae;
        v[2] += r;
        v[3] += ec;
    }
#pragma unroll
    for (int q = 0; q < 4; ++q)
        for (int o = 32; o >= 1; o >>= 1) v[q] += __shfl_xor(v[q], o, 64);
    __shared__ double sv[4][BS / 64];
    if ((threadIdx.x & 63) == 0)
        for (int q = 0; q < 4; ++q) sv[q][threadIdx.x >> 6] = v[q];
    __syncthreads();
    if (threadIdx.x < 4) {
        double a = 0.0;
        for (int w = 0; w < (int)(BS / 64); ++w) a += sv[threadIdx.x][w];
        partial[4 * blockIdx.x + threadIdx.x] = a;
    }
}

// p += alpha e (alpha from the partial sums, 0 when undefined)
__global__ void __launch_bounds__(BS) k_ws_apply(RB p, RBL L, int W, int H, int D, const float* __restrict__ e,
                                                 const double* __restrict__ partial, int nb,
                                                 const SolveCtl* __restrict__ c) {
    if (c->done) return;
    __shared__ float alpha;
    if (threadIdx.x == 0) {
        double re = 0.0, eae = 0.0, sr = 0.0, se = 0.0;
        for (int q = 0; q < nb; ++q) {
            re += partial[4 * q];
            eae += partial[4 * q + 1];
            sr += partial[4 * q + 2];
            se += partial[4 * q + 3];
        }
        const double al = eae > 0.0 ? (re - sr * se / ((double)W * H * D)) / eae : 0.0;
        alpha = (al == al && fabs(al) < 1e30) ? (float)al : 0.0f;
    }
    __syncthreads();
    const int n = W * H * D;
    for (int t = blockIdx.x * BS + threadIdx.x; t < n; t += gridDim.x * BS) {
        const int i = t % W, j = (t / W) % H, k = t / (W * H);
        float& dst = p.c[(i + j) & 1][L.at(i, j, k)];
        dst = dst + alpha * e[t];
    }
}

__device__ __forceinline__ double block_max(const double* __restrict__ partial, int n) {
    double mx = 0.0;
    for (int q = threadIdx.x; q < n; q += BS) mx = fmax(mx, partial[q]);
    for (int o = 32; o >= 1; o >>= 1) mx = fmax(mx, __shfl_xor(mx, o, 64));
    __shared__ double sm[BS / 64];
    if ((threadIdx.x & 63) == 0) sm[threadIdx.x >> 6] = mx;
    __syncthreads();
    double r = sm[0];
    for (int q = 1; q < (int)(BS / 64); ++q) r = fmax(r, sm[q]);
    __syncthreads();
    return r;
}

// the judge: residuals of the corrected (A) and uncorrected (B) pressure
// after the same sweeps; the lower one continues (B wins ties)
__global__ void __launch_bounds__(BS) k_ws_judge(const double* __restrict__ pa, int na, const double* __restrict__ pb,
                                                 int nb_, SolveCtl* c, int* host_done, int seq, double tol, int cap,
                                                 int add) {
    if (c->done) return;
    const double ra = block_max(pa, na);
    const double rb = block_max(pb, nb_);
    if (threadIdx.x == 0) {
        c->iters += add;
        c->checks += 2;
        c->res_ws = ra;
        if (ra < rb) {
            c->res = ra;
            c->ws = 1;
            c->ws_kept += 1;
        } else {
            c->res = rb;
            c->ws = 2;
            c->ws_rejected += 1;
        }
        if (!(c->res > tol) || c->iters >= cap) {
            c->done = 1;
            __hip_atomic_store(host_done, seq, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
            __threadfence_system();
        }
    }
}

// the "given prediction" path keeps the immediate comparison (no sweeps)
__global__ void __launch_bounds__(BS) k_res_final_ws(const double* __restrict__ partial, int n, SolveCtl* c,
                                                     int* host_done, int seq, double tol, int cap) {
    if (c->done) return;
    const double r = block_max(partial, n);
    if (threadIdx.x == 0) {
        c->checks += 1;
        c->res_ws = r;
        if (r < c->res) {
            c->res = r;
            c->ws = 1;
            c->ws_kept += 1;
        } else {
            c->ws = 2;
            c->ws_rejected += 1;
        }
    }
}

// continue from the uncorrected pressure when the prediction was rejected
__global__ void __launch_bounds__(BS) k_ws_restore(const SolveCtl* __restrict__ c, const float* __restrict__ backup,
                                                   float* __restrict__ p, size_t n) {
    if (c->ws != 2) return;
    for (size_t t = (size_t)blockIdx.x * BS + threadIdx.x; t < n; t += (size_t)gridDim.x * BS) p[t] = backup[t];
}

__global__ void k_ws_zero(unsigned* a) { *a = 0u; }

// a new application: no decision yet (the restore only follows this one's)
__global__ void k_ws_begin(SolveCtl* c) { c->ws = 0; }

inline unsigned nb(size_t n) { return (unsigned)((n + BS - 1) / BS); }

} // namespace

int warm_start_attach(Solver& s, const Net* net) {
    if (s.ws_net) {
        delete s.ws_net;
        s.ws_net = nullptr;
    }
    s.ws_src = net;
    if (!net) return BLZ_OK;
    // coarsening factor: 4, doubled until the coarse grid has at most 32
    // cells per axis (the scale of the training grids)
    s.ws_f = 4;
    while (std::max(s.W, std::max(s.H, s.D)) > 32 * s.ws_f) s.ws_f *= 2;
    s.Wc = (s.W + s.ws_f - 1) / s.ws_f;
    s.Hc = (s.H + s.ws_f - 1) / s.ws_f;
    s.Dc = (s.D + s.ws_f - 1) / s.ws_f;
    s.Wn = std::max(4, (s.Wc + 3) / 4 * 4);
    s.Hn = std::max(4, (s.Hc + 3) / 4 * 4);
    s.Dn = std::max(4, (s.Dc + 3) / 4 * 4);
    const size_t nn = (size_t)s.Wn * s.Hn * s.Dn;
    if (!s.ws_buf && hipMalloc(&s.ws_buf, (2 * nn + 64) * sizeof(float)) != hipSuccess) {
        set_error("hipMalloc failed for warm-start buffers");
        return BLZ_E_NOMEM;
    }
    if (!fused_available(s)) {
        set_error("grid too large for the warm start's second pressure plane set");
        return BLZ_E_INVALID;
    }
    Net* v = new (std::nothrow) Net();
    if (!v) {
        set_error("out of host memory");
        return BLZ_E_NOMEM;
    }
    v->init_view(*net, s.stream);
    if (int e = v->ensure_dims(NetDims{1, s.Dn, s.Hn, s.Wn})) {
        delete v;
        return e;
    }
    s.ws_net = v;
    return BLZ_OK;
}

static int ws_restrict(Solver& s, float dt, const int* stop) {
    const size_t nn = (size_t)s.Wn * s.Hn * s.Dn;
    float* res = s.ws_buf;
    float* mask = s.ws_buf + nn;
    unsigned* amax = (unsigned*)(s.ws_buf + 2 * nn);
    const float scale = (s.s.density * s.dx * s.dx) / dt;
    hipLaunchKernelGGL(k_ws_zero, dim3(1), dim3(1), 0, s.stream, amax);
    hipLaunchKernelGGL(k_ws_restrict, dim3(nb(nn)), dim3(BS), 0, s.stream, s.p, s.rhs, s.L, s.W, s.H, s.D, scale,
                       s.ws_f, s.Wc, s.Hc, s.Dc, s.Wn, s.Hn, s.Dn, res, mask, stop);
    hipLaunchKernelGGL(k_ws_rms, dim3(1), dim3(1024), 0, s.stream, res, (int)nn, s.Wc * s.Hc * s.Dc, amax, stop);
    BLZ_HIP(hipGetLastError());
    return BLZ_OK;
}

int warm_start_apply(Solver& s, float dt, SolveCtl* c, int* host_done, int seq, double tol, int cap, int judge) {
    if (!s.ws_net) return BLZ_OK;
    const size_t nn = (size_t)s.Wn * s.Hn * s.Dn;
    float* res = s.ws_buf;
    float* mask = s.ws_buf + nn;
    unsigned* amax = (unsigned*)(s.ws_buf + 2 * nn);
    hipLaunchKernelGGL(k_ws_begin, dim3(1), dim3(1), 0, s.stream, c);
    if (int e = ws_restrict(s, dt, &c->done)) return e;
    hipLaunchKernelGGL(k_ws_pack, dim3(nb(nn * NET_IN_CH)), dim3(BS), 0, s.stream, res, mask, amax, s.ws_net->x0,
                       (int)nn);
    BLZ_HIP(hipGetLastError());
    if (int e = s.ws_net->forward()) return e;
    // the uncorrected pressure goes to the spare planes, the corrected one
    // (prediction with its fitted scale) stays in the home planes
    float* spare = s.prb + 4 * s.Np;
    BLZ_HIP(hipMemcpyAsync(spare, s.prb, 2 * s.Np * sizeof(float), hipMemcpyDeviceToDevice, s.stream));
    const size_t N = (size_t)s.W * s.H * s.D;
    float* e = s.tmp;
    hipLaunchKernelGGL(k_ws_prolong, dim3(nb(N)), dim3(BS), 0, s.stream, e, s.W, s.H, s.D, s.ws_net->out,
                       1.0f / (float)s.ws_f, s.Wc, s.Hc, s.Dc, s.Wn, s.Hn, c);
    const float scale = (s.s.density * s.dx * s.dx) / dt;
    const int nbd = (int)std::min<size_t>(nb(N), RESIDUAL_BLOCKS / 4);
    hipLaunchKernelGGL(k_ws_dots, dim3(nbd), dim3(BS), 0, s.stream, s.p, s.rhs, s.L, s.W, s.H, s.D, scale, e,
                       s.red + 8, c);
    hipLaunchKernelGGL(k_ws_apply, dim3((unsigned)std::min<size_t>(nb(N), 4096)), dim3(BS), 0, s.stream, s.p, s.L,
                       s.W, s.H, s.D, e, s.red + 8, nbd, c);
    BLZ_HIP(hipGetLastError());
    // the judge: the same sweeps on both, the lower residual continues
    int launches = 0, kind = 0;
    const int keep = s.rbgs_variant;
    s.rbgs_variant = 8; // in place on the home planes (the spare ones hold the other path)
    int rc = launch_rbgs_sweeps(s, dt, judge, &c->done, &launches, &kind);
    s.rbgs_variant = keep;
    if (rc) return rc;
    if (int e2 = launch_rbgs_spare(s, dt, judge, &c->done)) return e2;
    int na = 0, nb2 = 0;
    if (int e2 = launch_residual_partial(s, dt, false, s.red + 8, &c->done, &na)) return e2;
    if (int e2 = launch_residual_partial(s, dt, true, s.red + 8 + RESIDUAL_BLOCKS, &c->done, &nb2)) return e2;
    hipLaunchKernelGGL(k_ws_judge, dim3(1), dim3(BS), 0, s.stream, s.red + 8, na, s.red + 8 + RESIDUAL_BLOCKS, nb2, c,
                       host_done, seq, tol, cap, judge);
    hipLaunchKernelGGL(k_ws_restore, dim3((unsigned)std::min<size_t>(nb(2 * s.Np), 4096)), dim3(BS), 0, s.stream, c,
                       spare, s.prb, 2 * s.Np);
    BLZ_HIP(hipGetLastError());
    return BLZ_OK;
}

// the correction step with a given coarse prediction (network grid, x
// fastest) instead of the network's: restrict, then steps 3-4 of the
// warm start, synchronously; for measuring what an exact coarse solve gains
int warm_start_apply_given(Solver& s, float dt, const float* host_pred, double* before, double* after, int* kept) {
    if (!s.ws_net) {
        set_error("no warm-start network attached");
        return BLZ_E_INVALID;
    }
    const size_t nn = (size_t)s.Wn * s.Hn * s.Dn;
    SolveCtl* c = s.ctl;
    if (int e = launch_ctl_reset(s, c)) return e;
    if (int e = launch_residual_check(s, dt, c, s.done_host_dev, 0, -1.0, 1 << 30, 0, 1)) return e;
    BLZ_HIP(hipMemcpyAsync(s.ws_net->out, host_pred, nn * sizeof(float), hipMemcpyHostToDevice, s.stream));
    float* backup = s.prb + 4 * s.Np;
    BLZ_HIP(hipMemcpyAsync(backup, s.prb, 2 * s.Np * sizeof(float), hipMemcpyDeviceToDevice, s.stream));
    const size_t N = (size_t)s.W * s.H * s.D;
    float* e = s.tmp;
    hipLaunchKernelGGL(k_ws_prolong, dim3(nb(N)), dim3(BS), 0, s.stream, e, s.W, s.H, s.D, s.ws_net->out,
                       1.0f / (float)s.ws_f, s.Wc, s.Hc, s.Dc, s.Wn, s.Hn, c);
    const float scale = (s.s.density * s.dx * s.dx) / dt;
    const int nbd = (int)std::min<size_t>(nb(N), RESIDUAL_BLOCKS / 4);
    hipLaunchKernelGGL(k_ws_dots, dim3(nbd), dim3(BS), 0, s.stream, s.p, s.rhs, s.L, s.W, s.H, s.D, scale, e,
                       s.red + 8, c);
    hipLaunchKernelGGL(k_ws_apply, dim3((unsigned)std::min<size_t>(nb(N), 4096)), dim3(BS), 0, s.stream, s.p, s.L,
                       s.W, s.H, s.D, e, s.red + 8, nbd, c);
    BLZ_HIP(hipGetLastError());
    if (int e2 = launch_residual_ws(s, dt, c, s.done_host_dev, 0, -1.0, 1 << 30)) return e2;
    hipLaunchKernelGGL(k_ws_restore, dim3((unsigned)std::min<size_t>(nb(2 * s.Np), 4096)), dim3(BS), 0, s.stream, c,
                       backup, s.prb, 2 * s.Np);
    BLZ_HIP(hipMemcpyAsync(s.ctl_host, c, sizeof(SolveCtl), hipMemcpyDeviceToHost, s.stream));
    BLZ_HIP(hipStreamSynchronize(s.stream));
    *before = s.ctl_host[0].res0;
    *after = s.ctl_host[0].res_ws;
    *kept = s.ctl_host[0].ws == 1;
    return BLZ_OK;
}

int warm_start_input(Solver& s, float dt, float* host_in2) {
    const size_t nn = (size_t)s.Wn * s.Hn * s.Dn;
    if (int e = ws_restrict(s, dt, s.zero_flag)) return e;
    float* out2 = (float*)s.tmp;
    hipLaunchKernelGGL(k_ws_export, dim3(nb(nn)), dim3(BS), 0, s.stream, s.ws_buf, s.ws_buf + nn,
                       (const unsigned*)(s.ws_buf + 2 * nn), out2, (int)nn);
    BLZ_HIP(hipGetLastError());
    BLZ_HIP(hipMemcpyAsync(host_in2, out2, 2 * nn * sizeof(float), hipMemcpyDeviceToHost, s.stream));
    BLZ_HIP(hipStreamSynchronize(s.stream));
    return BLZ_OK;
}

// the residual evaluation of the corrected pressure (partials of k_residual)
int launch_residual_ws_final(Solver& s, SolveCtl* c, int* host_done, int seq, double tol, int cap, int nb_partial) {
    hipLaunchKernelGGL(k_res_final_ws, dim3(1), dim3(BS), 0, s.stream, s.red + 8, nb_partial, c, host_done, seq, tol,
                       cap);
    BLZ_HIP(hipGetLastError());
    return BLZ_OK;
}

} // namespace blz
