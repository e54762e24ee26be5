// k_advect.hip — semi-Lagrangian advection: Ralston RK3 back-trace with
// trilinear velocity, tricubic sample of the advected field.
//
// Reference: advect_RK3_kernel (asiroisvigneux/Blaze src/FluidSolver.cu:431-456),
// launched for T (clamped), U, V, W (:942-945), then swap_grids_kernel
// (:458-465). Every grid reads the same front velocities and writes its back
// buffer; the host swaps front/back pointers afterwards (no device kernel).
#include "solver.h"

namespace blz {

namespace {

constexpr unsigned BS = 256;

template <bool CLAMP>
__global__ void __launch_bounds__(BS) k_advect(Gv U, Gv V, Gv W, Gv G, float* __restrict__ out, float dx, float a,
                                               float b, float c1, float c2, float c3, float dt) {
    const int n = G.nx * G.ny * G.nz;
    const int t = blockIdx.x * BS + threadIdx.x;
    if (t >= n) return;
    const int i = t % G.nx, j = (t / G.nx) % G.ny, k = t / (G.nx * G.ny);
    f3 pos = g2o(G, dx, i, j, k);
    const f3 k1 = mk3(trilinear(U, dx, pos), trilinear(V, dx, pos), trilinear(W, dx, pos));
    const f3 p1 = mk3(pos.x - a * k1.x, pos.y - a * k1.y, pos.z - a * k1.z);
    const f3 k2 = mk3(trilinear(U, dx, p1), trilinear(V, dx, p1), trilinear(W, dx, p1));
    const f3 p2 = mk3(pos.x - b * k2.x, pos.y - b * k2.y, pos.z - b * k2.z);
    const f3 k3 = mk3(trilinear(U, dx, p2), trilinear(V, dx, p2), trilinear(W, dx, p2));
    pos.x -= dt * (c1 * k1.x + c2 * k2.x + c3 * k3.x);
    pos.y -= dt * (c1 * k1.y + c2 * k2.y + c3 * k3.y);
    pos.z -= dt * (c1 * k1.z + c2 * k2.z + c3 * k3.z);
    out[t] = tricubic<CLAMP>(G, dx, pos);
}

} // namespace

int launch_advect(Solver& s, float dt) {
    const Gv U = s.f[F_U].view(), V = s.f[F_V].view(), W = s.f[F_W].view();
    const float a = 0.5f * dt, b = 0.75f * dt;
    const float c1 = 2.0f / 9.0f, c2 = 3.0f / 9.0f, c3 = 4.0f / 9.0f;
    for (int g = F_T; g < F_N; ++g) {
        if (g == F_S && !s.smoke) continue;
        Field& G = s.f[g];
        const size_t n = G.count();
        const bool clamp = (g == F_T || g == F_S);
        if (clamp)
            hipLaunchKernelGGL(k_advect<true>, dim3(blocks_for(n, BS)), dim3(BS), 0, s.stream, U, V, W, G.view(),
                               G.back(), s.dx, a, b, c1, c2, c3, dt);
        else
            hipLaunchKernelGGL(k_advect<false>, dim3(blocks_for(n, BS)), dim3(BS), 0, s.stream, U, V, W, G.view(),
                               G.back(), s.dx, a, b, c1, c2, c3, dt);
        BLZ_HIP(hipGetLastError());
    }
    for (int g = F_T; g < F_N; ++g)
        if (g != F_S || s.smoke) s.f[g].swap();
    return BLZ_OK;
}

} // namespace blz
