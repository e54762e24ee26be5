// k_project.hip — pressure projection: divergence, red/black Gauss-Seidel
// pressure solve, FP64 residual, pressure-gradient velocity update.
//
// Reference kernels (asiroisvigneux/Blaze src/FluidSolver.cu):
//   compute_divergence_kernel :475-515, gs_solve_kernel :517-599,
//   pressure_gradient_update_kernel :601-669, project() :823-856.
//
// Pressure and rhs are stored split by colour (red: (i+j) even, black: odd),
// each colour plane (W/2) x H x D with element (i>>1) + j*(W/2) + k*(W/2)*H.
// A sweep is a red pass then a black pass; in a pass every column of that
// colour walks k = 0..D-1 updating in place, reading the other colour (fixed
// during the pass) and its own k-1 value just computed. This equals the
// reference's per-slab red/black order inside a thread block, without its
// inter-block races, and every access of a pass is unit-stride in the split
// layout.
#include "solver.h"

namespace blz {

namespace {

constexpr unsigned BS = 256;

__device__ __forceinline__ int rb_index(int i, int j, int k, int Wh, int H) { return (i >> 1) + j * Wh + k * Wh * H; }

__device__ __forceinline__ float p_at(const RB& p, int i, int j, int k, int Wh, int H) {
    return p.c[(i + j) & 1][rb_index(i, j, k, Wh, H)];
}

// compute_divergence_kernel :475-515, written into the split layout
__global__ void __launch_bounds__(BS) k_divergence(const float* __restrict__ U, const float* __restrict__ V,
                                                   const float* __restrict__ Wf, RB rhs, int W, int H, int D,
                                                   float sc, int4 cbA, int2 cbB) {
    const int t = blockIdx.x * BS + threadIdx.x;
    if (t >= W * H * D) return;
    const int i = t % W, j = (t / W) % H, k = t / (W * H);
    const int u0 = i + j * (W + 1) + k * (W + 1) * H;
    const int v0 = i + j * W + k * W * (H + 1);
    const int w0 = t;
    const float u_lo = U[u0], u_hi = U[u0 + 1];
    const float v_lo = V[v0], v_hi = V[v0 + W];
    const float w_lo = Wf[w0], w_hi = Wf[w0 + W * H];
    float r = -sc * (u_hi - u_lo + v_hi - v_lo + w_hi - w_lo);
    // closedBounds: cbA = (xMin, xMax, yMin, yMax), cbB = (zMin, zMax)
    if (i == 0) {
        if (cbA.x) r -= sc * u_lo;
    } else if (i == W - 1) {
        if (cbA.y) r += sc * u_hi;
    }
    if (j == 0) {
        if (cbA.w) r -= sc * v_lo;
    } else if (j == H - 1) {
        if (cbA.z) r += sc * v_hi;
    }
    if (k == 0) {
        if (cbB.x) r -= sc * w_lo;
    } else if (k == D - 1) {
        if (cbB.y) r += sc * w_hi;
    }
    rhs.c[(i + j) & 1][rb_index(i, j, k, W >> 1, H)] = r;
}

// One colour pass. Thread per column (m, j) of the colour; software
// pipelined PF slabs ahead (all loads of a pass are independent of the
// values the pass writes, except the in-register k-1 value).
template <int PF>
__global__ void __launch_bounds__(BS) k_rbgs_pass(float* __restrict__ A, const float* __restrict__ O,
                                                  const float* __restrict__ Rc, int W, int H, int D, int colour,
                                                  float scale) {
    const int Wh = W >> 1;
    const int t = blockIdx.x * BS + threadIdx.x;
    if (t >= Wh * H) return;
    const int m = t % Wh, j = t / Wh;
    const int q = (j + colour) & 1;
    const int i = 2 * m + q;
    const bool hL = i > 0, hR = i < W - 1, hN = j > 0, hS = j < H - 1;
    const int plane = Wh * H;
    const int oL = t + q - 1, oR = t + q, oN = t - Wh, oS = t + Wh;
    const float cnt_xy = (float)((int)hL + (int)hR + (int)hN + (int)hS);

    float r_rhs[PF], r_l[PF], r_r[PF], r_n[PF], r_s[PF], r_up[PF];
#pragma unroll
    for (int u = 0; u < PF; ++u) {
        const int k = u;
        if (k < D) {
            const int b = k * plane;
            r_rhs[u] = Rc[t + b];
            r_l[u] = hL ? O[oL + b] : 0.0f;
            r_r[u] = hR ? O[oR + b] : 0.0f;
            r_n[u] = hN ? O[oN + b] : 0.0f;
            r_s[u] = hS ? O[oS + b] : 0.0f;
            r_up[u] = (k < D - 1) ? A[t + b + plane] : 0.0f;
        }
    }
    float prev = 0.0f;
    for (int kb = 0; kb < D; kb += PF) {
#pragma unroll
        for (int u = 0; u < PF; ++u) {
            const int k = kb + u;
            if (k < D) {
                float num = scale * r_rhs[u];
                if (hL) num += r_l[u];
                if (hR) num += r_r[u];
                if (hN) num += r_n[u];
                if (hS) num += r_s[u];
                float den = cnt_xy;
                if (k > 0) { num += prev; den += 1.0f; }
                if (k < D - 1) { num += r_up[u]; den += 1.0f; }
                const float pn = num / den;
                A[t + k * plane] = pn;
                prev = pn;
                const int kn = k + PF;
                if (kn < D) {
                    const int b = kn * plane;
                    r_rhs[u] = Rc[t + b];
                    r_l[u] = hL ? O[oL + b] : 0.0f;
                    r_r[u] = hR ? O[oR + b] : 0.0f;
                    r_n[u] = hN ? O[oN + b] : 0.0f;
                    r_s[u] = hS ? O[oS + b] : 0.0f;
                    r_up[u] = (kn < D - 1) ? A[t + b + plane] : 0.0f;
                }
            }
        }
    }
}

__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
    return v;
}

// block max of non-negative doubles -> atomic max on the bit pattern
__device__ __forceinline__ void block_max_to(double v, double* out) {
    __shared__ double sm[BS / 64];
    v = wave_max(v);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) sm[wid] = v;
    __syncthreads();
    if (wid == 0) {
        double w = lane < (int)(BS / 64) ? sm[lane] : 0.0;
        w = wave_max(w);
        if (lane == 0) atomicMax((unsigned long long*)out, (unsigned long long)__double_as_longlong(w));
    }
}

// ||div u||_inf proxy: max over cells not adjacent to an open face of
// |rhs + c (sum_nbr p - n p)|, c = dt/(rho dx^2), all in FP64; NaN -> +inf
__global__ void __launch_bounds__(BS) k_residual(RB p, RB rhs, int W, int H, int D, double c, int4 cbA, int2 cbB,
                                                 double* out) {
    const int t = blockIdx.x * BS + threadIdx.x;
    double r = 0.0;
    if (t < W * H * D) {
        const int i = t % W, j = (t / W) % H, k = t / (W * H);
        const bool skip = (i == 0 && !cbA.x) || (i == W - 1 && !cbA.y) || (j == 0 && !cbA.w) ||
                          (j == H - 1 && !cbA.z) || (k == 0 && !cbB.x) || (k == D - 1 && !cbB.y);
        if (!skip) {
            const int Wh = W >> 1;
            double sum = 0.0, n = 0.0;
            if (i > 0) { sum += (double)p_at(p, i - 1, j, k, Wh, H); n += 1.0; }
            if (i < W - 1) { sum += (double)p_at(p, i + 1, j, k, Wh, H); n += 1.0; }
            if (j > 0) { sum += (double)p_at(p, i, j - 1, k, Wh, H); n += 1.0; }
            if (j < H - 1) { sum += (double)p_at(p, i, j + 1, k, Wh, H); n += 1.0; }
            if (k > 0) { sum += (double)p_at(p, i, j, k - 1, Wh, H); n += 1.0; }
            if (k < D - 1) { sum += (double)p_at(p, i, j, k + 1, Wh, H); n += 1.0; }
            const double pc = (double)p_at(p, i, j, k, Wh, H);
            const double rc = (double)rhs.c[(i + j) & 1][rb_index(i, j, k, Wh, H)];
            r = fabs(rc + c * (sum - n * pc));
            if (r != r) r = __longlong_as_double(0x7ff0000000000000LL);
        }
    }
    block_max_to(r, out);
}

// pressure_gradient_update_kernel :601-669 for the faces along AXIS,
// out-of-place: interior faces subtract s*(p_hi - p_lo); boundary faces are
// 0 when closed, else the updated value of the adjacent interior face.
template <int AXIS>
__global__ void __launch_bounds__(BS) k_pupdate(const float* __restrict__ in, float* __restrict__ out, RB p, int W,
                                                int H, int D, float sc, int closedLo, int closedHi) {
    const int fx = W + (AXIS == 0), fy = H + (AXIS == 1), fz = D + (AXIS == 2);
    const int t = blockIdx.x * BS + threadIdx.x;
    if (t >= fx * fy * fz) return;
    const int i = t % fx, j = (t / fx) % fy, k = t / (fx * fy);
    const int f = AXIS == 0 ? i : (AXIS == 1 ? j : k);
    const int nA = AXIS == 0 ? W : (AXIS == 1 ? H : D);
    const int sF = AXIS == 0 ? 1 : (AXIS == 1 ? fx : fx * fy); // face stride along the axis
    const int Wh = W >> 1;
    int g, fi = i, fj = j, fk = k; // face whose updated value is taken
    if (f == 0) {
        if (closedLo) { out[t] = 0.0f; return; }
        g = t + sF;
        if (AXIS == 0) fi = 1; else if (AXIS == 1) fj = 1; else fk = 1;
    } else if (f == nA) {
        if (closedHi) { out[t] = 0.0f; return; }
        g = t - sF;
        if (AXIS == 0) fi = nA - 1; else if (AXIS == 1) fj = nA - 1; else fk = nA - 1;
    } else {
        g = t;
    }
    // interior face (fi,fj,fk) lies between cells (f-1) and f along AXIS
    const float p_hi = p_at(p, fi, fj, fk, Wh, H);
    const float p_lo = AXIS == 0 ? p_at(p, fi - 1, fj, fk, Wh, H)
                                 : (AXIS == 1 ? p_at(p, fi, fj - 1, fk, Wh, H) : p_at(p, fi, fj, fk - 1, Wh, H));
    out[t] = in[g] - sc * (p_hi - p_lo);
}

__global__ void __launch_bounds__(BS) k_rb_to_natural(RB p, float* __restrict__ out, int W, int H, int D) {
    const int t = blockIdx.x * BS + threadIdx.x;
    if (t >= W * H * D) return;
    const int i = t % W, j = (t / W) % H, k = t / (W * H);
    out[t] = p_at(p, i, j, k, W >> 1, H);
}

__global__ void __launch_bounds__(BS) k_natural_to_rb(RB p, const float* __restrict__ in, int W, int H, int D) {
    const int t = blockIdx.x * BS + threadIdx.x;
    if (t >= W * H * D) return;
    const int i = t % W, j = (t / W) % H, k = t / (W * H);
    p.c[(i + j) & 1][rb_index(i, j, k, W >> 1, H)] = in[t];
}

// max |a| over an array (NaN counts as +inf), into a shared double slot
__global__ void __launch_bounds__(BS) k_maxabs(const float* __restrict__ a, size_t n, double* out) {
    double m = 0.0;
    for (size_t t = (size_t)blockIdx.x * BS + threadIdx.x; t < n; t += (size_t)gridDim.x * BS) {
        double v = (double)fabsf(a[t]);
        if (v != v) v = __longlong_as_double(0x7ff0000000000000LL);
        m = fmax(m, v);
    }
    block_max_to(m, out);
}

__global__ void k_fill_doubles(double* d, int n, double v) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) d[t] = v;
}

} // namespace

#define LAUNCH(kern, n, ...)                                                                  \
    do {                                                                                      \
        if ((n) > 0) {                                                                        \
            hipLaunchKernelGGL(kern, dim3(blocks_for((n), BS)), dim3(BS), 0, s.stream, __VA_ARGS__); \
            BLZ_HIP(hipGetLastError());                                                       \
        }                                                                                     \
    } while (0)

static inline int4 cb_a(const Solver& s) {
    const int* c = s.s.closed_bounds;
    return make_int4(c[0], c[1], c[2], c[3]);
}
static inline int2 cb_b(const Solver& s) {
    const int* c = s.s.closed_bounds;
    return make_int2(c[4], c[5]);
}

int launch_divergence(Solver& s) {
    const float sc = 1.0f / s.dx;
    LAUNCH(k_divergence, (size_t)s.W * s.H * s.D, s.f[F_U].front(), s.f[F_V].front(), s.f[F_W].front(), s.rhs, s.W,
           s.H, s.D, sc, cb_a(s), cb_b(s));
    return BLZ_OK;
}

int launch_rbgs_sweeps(Solver& s, float dt, int sweeps) {
    const float scale = (s.s.density * s.dx * s.dx) / dt; // Bridson p.75, :530
    const size_t cols = (size_t)(s.W / 2) * s.H;
    for (int l = 0; l < sweeps; ++l) {
        for (int c = 0; c < 2; ++c) {
            LAUNCH((k_rbgs_pass<8>), cols, s.p.c[c], s.p.c[c ^ 1], s.rhs.c[c], s.W, s.H, s.D, c, scale);
        }
    }
    return BLZ_OK;
}

int launch_clear_doubles(Solver& s, double* dev, int n) {
    hipLaunchKernelGGL(k_fill_doubles, dim3(1), dim3(64), 0, s.stream, dev, n, 0.0);
    BLZ_HIP(hipGetLastError());
    return BLZ_OK;
}

int launch_residual(Solver& s, float dt, double* dev_out) {
    const double c = (double)dt / ((double)s.s.density * (double)s.dx * (double)s.dx);
    if (int e = launch_clear_doubles(s, dev_out, 1)) return e;
    LAUNCH(k_residual, (size_t)s.W * s.H * s.D, s.p, s.rhs, s.W, s.H, s.D, c, cb_a(s), cb_b(s), dev_out);
    return BLZ_OK;
}

int launch_max_abs_velocity(Solver& s, double* dev_out) {
    if (int e = launch_clear_doubles(s, dev_out, 1)) return e;
    for (int g = F_U; g <= F_W; ++g) {
        const size_t n = s.f[g].count();
        const unsigned nb = std::min<size_t>(blocks_for(n, BS), 2048);
        hipLaunchKernelGGL(k_maxabs, dim3(nb), dim3(BS), 0, s.stream, s.f[g].front(), n, dev_out);
        BLZ_HIP(hipGetLastError());
    }
    return BLZ_OK;
}

int launch_pressure_update(Solver& s, float dt) {
    const float sc = dt / (s.s.density * s.dx);
    const int* cb = s.s.closed_bounds;
    Field& U = s.f[F_U];
    Field& V = s.f[F_V];
    Field& W = s.f[F_W];
    LAUNCH(k_pupdate<0>, U.count(), U.front(), U.back(), s.p, s.W, s.H, s.D, sc, cb[0], cb[1]);
    LAUNCH(k_pupdate<1>, V.count(), V.front(), V.back(), s.p, s.W, s.H, s.D, sc, cb[3], cb[2]);
    LAUNCH(k_pupdate<2>, W.count(), W.front(), W.back(), s.p, s.W, s.H, s.D, sc, cb[4], cb[5]);
    U.swap();
    V.swap();
    W.swap();
    return BLZ_OK;
}

int launch_p_to_natural(Solver& s, float* out, const RB& src) {
    LAUNCH(k_rb_to_natural, (size_t)s.W * s.H * s.D, src, out, s.W, s.H, s.D);
    return BLZ_OK;
}

int launch_natural_to_rb(Solver& s, RB& dst, const float* in) {
    LAUNCH(k_natural_to_rb, (size_t)s.W * s.H * s.D, dst, in, s.W, s.H, s.D);
    return BLZ_OK;
}

} // namespace blz
