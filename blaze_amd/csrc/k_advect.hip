// k_advect.hip — semi-Lagrangian advection: Ralston RK3 back-trace with
// trilinear velocity, tricubic sample of the advected field.
//
// Reference: advect_RK3_kernel (asiroisvigneux/Blaze src/FluidSolver.cu:431-456),
// launched for T (clamped), U, V, W (:942-945), then swap_grids_kernel
// (:458-465). Every grid reads the same front velocities and writes its back
// buffer; the host swaps front/back pointers afterwards (no device kernel).
// Positions are scaled to grid units by p / dx, or by p * 2^k when dx =
// 2^-k (the same bits, without 30 IEEE divisions per cell).
#include "solver.h"

namespace blz {

namespace {

constexpr unsigned BS = 256;

// Ralston RK3 back-trace of cell (i, j, k) of grid G through the front
// velocities (trilinear), in object space
template <class S>
__device__ __forceinline__ f3 backtrace(const Gv& U, const Gv& V, const Gv& W, const Gv& G, const S& sc, float dx,
                                        int i, int j, int k, float a, float b, float c1, float c2, float c3,
                                        float dt) {
    f3 pos = g2o(G, dx, i, j, k);
    const f3 k1 = mk3(trilinear_s(U, sc, pos), trilinear_s(V, sc, pos), trilinear_s(W, sc, pos));
    const f3 p1 = mk3(pos.x - a * k1.x, pos.y - a * k1.y, pos.z - a * k1.z);
    const f3 k2 = mk3(trilinear_s(U, sc, p1), trilinear_s(V, sc, p1), trilinear_s(W, sc, p1));
    const f3 p2 = mk3(pos.x - b * k2.x, pos.y - b * k2.y, pos.z - b * k2.z);
    const f3 k3 = mk3(trilinear_s(U, sc, p2), trilinear_s(V, sc, p2), trilinear_s(W, sc, p2));
    pos.x -= dt * (c1 * k1.x + c2 * k2.x + c3 * k3.x);
    pos.y -= dt * (c1 * k1.y + c2 * k2.y + c3 * k3.y);
    pos.z -= dt * (c1 * k1.z + c2 * k2.z + c3 * k3.z);
    return pos;
}

// Ralston's weights (src/FluidSolver.cu:437-452)
constexpr float RK_C1 = 2.0f / 9.0f, RK_C2 = 3.0f / 9.0f, RK_C3 = 4.0f / 9.0f;

template <bool CLAMP, class S>
__device__ __forceinline__ void advect_cell(const Gv& U, const Gv& V, const Gv& W, const Gv& G,
                                            float* __restrict__ out, const S& sc, float dx, float a, float b,
                                            float c1, float c2, float c3, float dt, int t) {
    const int n = G.nx * G.ny * G.nz;
    if (t >= n) return;
    const int i = t % G.nx, j = (t / G.nx) % G.ny, k = t / (G.nx * G.ny);
    const f3 pos = backtrace(U, V, W, G, sc, dx, i, j, k, a, b, c1, c2, c3, dt);
    as_global(out)[t] = tricubic_s<CLAMP>(G, sc, pos);
}

template <bool CLAMP, class S>
__global__ void __launch_bounds__(BS) k_advect(Gv U, Gv V, Gv W, Gv G, float* __restrict__ out, S sc, float dx,
                                               float a, float b, float c1, float c2, float c3, float dt) {
    advect_cell<CLAMP>(U, V, W, G, out, sc, dx, a, b, c1, c2, c3, dt, (int)(blockIdx.x * BS + threadIdx.x));
}

// temperature and smoke density share the cell-centred grid, hence the
// back-trace: one trace, two clamped tricubic samples (the same values as
// two separate launches)
template <class S>
__device__ __forceinline__ void advect_pair_cell(const Gv& U, const Gv& V, const Gv& W, const Gv& G1, const Gv& G2,
                                                 float* __restrict__ out1, float* __restrict__ out2, const S& sc,
                                                 float dx, float a, float b, float c1, float c2, float c3, float dt,
                                                 int t) {
    const int n = G1.nx * G1.ny * G1.nz;
    if (t >= n) return;
    const int i = t % G1.nx, j = (t / G1.nx) % G1.ny, k = t / (G1.nx * G1.ny);
    const f3 pos = backtrace(U, V, W, G1, sc, dx, i, j, k, a, b, c1, c2, c3, dt);
    as_global(out1)[t] = tricubic_s<true>(G1, sc, pos);
    as_global(out2)[t] = tricubic_s<true>(G2, sc, pos);
}

template <class S>
__global__ void __launch_bounds__(BS) k_advect_pair(Gv U, Gv V, Gv W, Gv G1, Gv G2, float* __restrict__ out1,
                                                    float* __restrict__ out2, S sc, float dx, float a, float b,
                                                    float c1, float c2, float c3, float dt) {
    advect_pair_cell(U, V, W, G1, G2, out1, out2, sc, dx, a, b, c1, c2, c3, dt, (int)(blockIdx.x * BS + threadIdx.x));
}

// the job of block b in a table of first blocks (ascending)
__device__ __forceinline__ const AdvJob& adv_job_of(const AdvJob* __restrict__ tab, int n, int b) {
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (tab[mid].first_block <= b) lo = mid;
        else hi = mid - 1;
    }
    return tab[lo];
}

// the advection of one field (or a pair) of every grid in a table
template <int KIND, class S>
__global__ void __launch_bounds__(BS) k_advect_multi(const AdvJob* __restrict__ tab, int n) {
    const AdvJob& e = adv_job_of(tab, n, (int)blockIdx.x);
    const int t = (int)((blockIdx.x - e.first_block) * BS + threadIdx.x);
    const S sc{e.s};
    if constexpr (KIND == ADV_PAIR)
        advect_pair_cell(e.U, e.V, e.W, e.G1, e.G2, e.out1, e.out2, sc, e.dx, e.a, e.b, RK_C1, RK_C2, RK_C3, e.dt, t);
    else
        advect_cell<KIND == ADV_CLAMPED>(e.U, e.V, e.W, e.G1, e.out1, sc, e.dx, e.a, e.b, RK_C1, RK_C2, RK_C3, e.dt,
                                         t);
}

template <class S>
int advect_all(Solver& s, const S& sc, float dt) {
    const Gv U = s.f[F_U].view(), V = s.f[F_V].view(), W = s.f[F_W].view();
    const float a = 0.5f * dt, b = 0.75f * dt;
    const float c1 = RK_C1, c2 = RK_C2, c3 = RK_C3;
    for (int g = F_T; g < F_N; ++g) {
        Field& G = s.f[g];
        const size_t n = G.count();
        if (g == F_S && !s.smoke) continue;
        const Gv tv = s.f[F_T].view(), dv = s.f[F_S].view();
        const bool pair = s.smoke && tv.nx == dv.nx && tv.ny == dv.ny && tv.nz == dv.nz && tv.ox == dv.ox &&
                          tv.oy == dv.oy && tv.oz == dv.oz;
        if (g == F_S && pair) continue; // with T
        if (g == F_T && pair) {
            Field& D = s.f[F_S];
            hipLaunchKernelGGL(k_advect_pair<S>, dim3(blocks_for(n, BS)), dim3(BS), 0, s.stream, U, V, W, G.view(),
                               D.view(), G.back(), D.back(), sc, s.dx, a, b, c1, c2, c3, dt);
        } else if (g == F_T || g == F_S) {
            hipLaunchKernelGGL((k_advect<true, S>), dim3(blocks_for(n, BS)), dim3(BS), 0, s.stream, U, V, W, G.view(),
                               G.back(), sc, s.dx, a, b, c1, c2, c3, dt);
        } else {
            hipLaunchKernelGGL((k_advect<false, S>), dim3(blocks_for(n, BS)), dim3(BS), 0, s.stream, U, V, W,
                               G.view(), G.back(), sc, s.dx, a, b, c1, c2, c3, dt);
        }
        BLZ_HIP(hipGetLastError());
    }
    return BLZ_OK;
}

} // namespace

// the jobs of launch_advect (appended to jobs[2 * kind + (scale multiplies)],
// first blocks left to the caller), with the same field swaps
void advect_jobs(Solver& s, float dt, std::vector<AdvJob>* jobs) {
    float inv = 0.0f;
    const bool mul = dx_pow2(s.dx, &inv);
    const Gv U = s.f[F_U].view(), V = s.f[F_V].view(), W = s.f[F_W].view();
    const Gv tv = s.f[F_T].view(), dv = s.f[F_S].view();
    const bool pair = s.smoke && tv.nx == dv.nx && tv.ny == dv.ny && tv.nz == dv.nz && tv.ox == dv.ox &&
                      tv.oy == dv.oy && tv.oz == dv.oz;
    for (int g = F_T; g < F_N; ++g) {
        if (g == F_S && (!s.smoke || pair)) continue;
        Field& G = s.f[g];
        AdvJob j;
        j.U = U;
        j.V = V;
        j.W = W;
        j.G1 = G.view();
        j.G2 = (g == F_T && pair) ? dv : G.view();
        j.out1 = G.back();
        j.out2 = (g == F_T && pair) ? s.f[F_S].back() : nullptr;
        j.dx = s.dx;
        j.s = mul ? inv : s.dx;
        j.a = 0.5f * dt;
        j.b = 0.75f * dt;
        j.dt = dt;
        j.first_block = 0;
        j.n = (int)G.count();
        const int kind = (g == F_T && pair) ? ADV_PAIR : ((g == F_T || g == F_S) ? ADV_CLAMPED : ADV_FREE);
        jobs[2 * kind + (mul ? 1 : 0)].push_back(j);
    }
    for (int g = F_T; g < F_N; ++g)
        if (g != F_S || s.smoke) s.f[g].swap();
}

int launch_advect_multi(hipStream_t st, int kind, bool mul, const AdvJob* tab, int n, int blocks) {
    if (n < 1 || blocks < 1) return BLZ_OK;
    const dim3 gr((unsigned)blocks), bl(BS);
#define ADV_MULTI(K)                                                                        \
    do {                                                                                    \
        if (mul) hipLaunchKernelGGL((k_advect_multi<K, MulDx>), gr, bl, 0, st, tab, n);    \
        else hipLaunchKernelGGL((k_advect_multi<K, DivDx>), gr, bl, 0, st, tab, n);        \
    } while (0)
    if (kind == ADV_PAIR) ADV_MULTI(ADV_PAIR);
    else if (kind == ADV_CLAMPED) ADV_MULTI(ADV_CLAMPED);
    else ADV_MULTI(ADV_FREE);
#undef ADV_MULTI
    BLZ_HIP(hipGetLastError());
    return BLZ_OK;
}

int launch_advect(Solver& s, float dt) {
    float inv = 0.0f;
    const int rc = dx_pow2(s.dx, &inv) ? advect_all(s, MulDx{inv}, dt) : advect_all(s, DivDx{s.dx}, dt);
    if (rc) return rc;
    for (int g = F_T; g < F_N; ++g)
        if (g != F_S || s.smoke) s.f[g].swap();
    return BLZ_OK;
}

} // namespace blz
