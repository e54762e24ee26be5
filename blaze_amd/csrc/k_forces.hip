// k_forces.hip — body forces of the solver step: cooling, drag, buoyancy,
// vorticity confinement, wind and curl-noise turbulence.
//
// Reference kernels (asiroisvigneux/Blaze src/FluidSolver.cu):
//   temperature_cooldown_kernel :118-124, apply_drag_kernel :126-140,
//   add_buoyancy_kernel :142-156, computeVorticity :174-201,
//   vorticity_confinement_kernel :203-248, add_curl_noise_kernel :271-293,
//   compute_turbulence_kernel :295-335, add_wind_kernel :337-390.
//
// The reference scatters vorticity and curl-noise face increments with float
// atomics. Here each cell writes its increments once and a face-centred
// gather adds them in a fixed order, (face + lower cell) + upper cell, so
// results are deterministic and match the serial CPU restatement bit for bit.
#include "solver.h"

namespace blz {

namespace {

constexpr unsigned BS = 256;

__global__ void __launch_bounds__(BS) k_scale(float* __restrict__ a, size_t n, float f) {
    size_t t = (size_t)blockIdx.x * BS + threadIdx.x;
    if (t < n) a[t] *= f;
}

// V faces: V += dt * ((beta * T~) * g), T~ trilinear at the face position
template <class S>
__device__ __forceinline__ void buoyancy_cell(const Gv& T, float* __restrict__ V, int vnx, int vny, int vnz,
                                              float vox, float voy, float voz, const S& sc, float dx, float beta,
                                              float g, float dt, int t) {
    const int n = vnx * vny * vnz;
    if (t >= n) return;
    const int i = t % vnx, j = (t / vnx) % vny, k = t / (vnx * vny);
    const f3 pos = mk3(((float)i + vox) * dx, ((float)j + voy) * dx, ((float)k + voz) * dx);
    const float temp = trilinear_s(T, sc, pos);
    const float force = beta * temp * g;
    as_global(V)[t] += dt * force;
}

template <class S>
__global__ void __launch_bounds__(BS) k_buoyancy(Gv T, float* __restrict__ V, int vnx, int vny, int vnz, float vox,
                                                 float voy, float voz, S sc, float dx, float beta, float g,
                                                 float dt) {
    buoyancy_cell(T, V, vnx, vny, vnz, vox, voy, voz, sc, dx, beta, g, dt, (int)(blockIdx.x * BS + threadIdx.x));
}

// the buoyancy of every grid in a table (blocks mapped through first blocks)
template <class S>
__global__ void __launch_bounds__(BS) k_buoyancy_multi(const BuoyJob* __restrict__ tab, int n) {
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (tab[mid].first_block <= (int)blockIdx.x) lo = mid;
        else hi = mid - 1;
    }
    const BuoyJob& e = tab[lo];
    const S sc{e.s};
    buoyancy_cell(e.T, e.V, e.vnx, e.vny, e.vnz, e.vox, e.voy, e.voz, sc, e.dx, e.beta, e.g, e.dt,
                  (int)((blockIdx.x - e.first_block) * BS + threadIdx.x));
}

// computeVorticity :174-201 at cell (i,j,k) of the T grid samples U, V and W
// (trilinear) at the centres of the six neighbouring cells, positions
// ((float)(i-+1) + ox) * dx etc. Each such position is, in the same float
// operations, the centre g2o(T, dx, ...) of that neighbour, so every sample
// is a pure function of one cell: k_vort_centres samples the three
// components once per cell (tmp planes 4..6) and k_vort_omega differences
// them, with the reference's expression order, the same bits as twelve
// samples per cell at a quarter of the gathers.
template <class S>
__global__ void __launch_bounds__(BS) k_vort_centres(Gv T, Gv U, Gv V, Gv W, S sc, float dx, float* __restrict__ tmp,
                                                     int nW, int nH, int nD) {
    const int t = blockIdx.x * BS + threadIdx.x;
    if (t >= nW * nH * nD) return;
    const int i = t % nW, j = (t / nW) % nH, k = t / (nW * nH);
    const size_t N = (size_t)nW * nH * nD;
    const f3 pos = g2o(T, dx, i, j, k);
    tmp[4 * N + t] = trilinear_s(U, sc, pos);
    tmp[5 * N + t] = trilinear_s(V, sc, pos);
    tmp[6 * N + t] = trilinear_s(W, sc, pos);
}

// omega and |omega| for cells 1..N-2 on each axis (tmp planes 0..3). The
// reference evaluates omega 7 times per cell; it is a pure function of the
// cell, so evaluating it once and reading it back gives the same bits.
__global__ void __launch_bounds__(BS) k_vort_omega(float dx, float* __restrict__ tmp, int nW, int nH, int nD) {
    const int iw = nW - 2, ih = nH - 2, id = nD - 2;
    const int t = blockIdx.x * BS + threadIdx.x;
    if (t >= iw * ih * id) return;
    const int i = 1 + t % iw, j = 1 + (t / iw) % ih, k = 1 + t / (iw * ih);
    const size_t N = (size_t)nW * nH * nD;
    const size_t sy = nW, sz = (size_t)nW * nH;
    const size_t c = (size_t)i + (size_t)j * sy + (size_t)k * sz;
    const float* Uc = tmp + 4 * N;
    const float* Vc = tmp + 5 * N;
    const float* Wc = tmp + 6 * N;
    const float w121 = Wc[c + sy], w101 = Wc[c - sy];
    const float v112 = Vc[c + sz], v110 = Vc[c - sz];
    const float u112 = Uc[c + sz], u110 = Uc[c - sz];
    const float w211 = Wc[c + 1], w011 = Wc[c - 1];
    const float v211 = Vc[c + 1], v011 = Vc[c - 1];
    const float u121 = Uc[c + sy], u101 = Uc[c - sy];
    const float t2 = 2.0f * dx;
    const f3 om = mk3((w121 - w101 - (v112 - v110)) / t2, (u112 - u110 - (w211 - w011)) / t2,
                      (v211 - v011 - (u121 - u101)) / t2);
    tmp[c] = om.x;
    tmp[N + c] = om.y;
    tmp[2 * N + c] = om.z;
    tmp[3 * N + c] = len3(om);
}

// confinement increments 0.5*dt*f for active cells 2..N-3 (tmp planes 4..6)
__global__ void __launch_bounds__(BS) k_vort_force(float dx, float* __restrict__ tmp, int nW, int nH, int nD, float conf,
                                                   float dt, float eps) {
    const int iw = nW - 4, ih = nH - 4, id = nD - 4;
    const int t = blockIdx.x * BS + threadIdx.x;
    if (t >= iw * ih * id) return;
    const int i = 2 + t % iw, j = 2 + (t / iw) % ih, k = 2 + t / (iw * ih);
    const size_t N = (size_t)nW * nH * nD;
    const size_t sy = nW, sz = (size_t)nW * nH;
    const size_t c = (size_t)i + j * sy + k * sz;
    const float* nrm = tmp + 3 * N;
    f3 om = mk3(tmp[c], tmp[N + c], tmp[2 * N + c]);
    const float t2 = 2.0f * dx;
    const f3 gn = mk3((nrm[c + 1] - nrm[c - 1]) / t2, (nrm[c + sy] - nrm[c - sy]) / t2,
                      (nrm[c + sz] - nrm[c - sz]) / t2);
    // the velocity at the cell centre: this cell's k_vort_centres samples
    // (read before this thread overwrites them with the increments)
    const f3 vel = mk3(tmp[4 * N + c], tmp[5 * N + c], tmp[6 * N + c]);
    const float speed = len3(vel);
    const float onorm = nrm[c] + eps;
    if (onorm > speed) om = mk3(speed * om.x / onorm, speed * om.y / onorm, speed * om.z / onorm);
    const float cd = conf * dx;
    const f3 cr = cross3(gn, om);
    const float h = 0.5f * dt;
    tmp[4 * N + c] = h * (cd * cr.x);
    tmp[5 * N + c] = h * (cd * cr.y);
    tmp[6 * N + c] = h * (cd * cr.z);
}

// Face gather along AXIS: face f (index along the axis) receives the
// increment of cell f-1 (its upper face) then of cell f (its lower face).
// RANGE2: only cells 2..N-3 on every axis carry increments (vorticity);
// otherwise every cell does (curl noise).
template <int AXIS, bool RANGE2>
__global__ void __launch_bounds__(BS) k_face_gather(float* __restrict__ face, const float* __restrict__ inc,
                                                    int nW, int nH, int nD) {
    const int fx = nW + (AXIS == 0), fy = nH + (AXIS == 1), fz = nD + (AXIS == 2);
    const int t = blockIdx.x * BS + threadIdx.x;
    if (t >= fx * fy * fz) return;
    const int i = t % fx, j = (t / fx) % fy, k = t / (fx * fy);
    const int f = AXIS == 0 ? i : (AXIS == 1 ? j : k);
    const int nA = AXIS == 0 ? nW : (AXIS == 1 ? nH : nD);
    bool lo = f >= 1, hi = f < nA;
    if (RANGE2) {
        const bool oi = AXIS == 0 || (i >= 2 && i <= nW - 3);
        const bool oj = AXIS == 1 || (j >= 2 && j <= nH - 3);
        const bool ok = AXIS == 2 || (k >= 2 && k <= nD - 3);
        const bool other = oi && oj && ok;
        lo = lo && other && (f - 1 >= 2) && (f - 1 <= nA - 3);
        hi = hi && other && (f >= 2) && (f <= nA - 3);
    }
    const size_t sA = AXIS == 0 ? 1 : (AXIS == 1 ? (size_t)nW : (size_t)nW * nH);
    // cell (i,j,k) of the T grid; for f == nA it is out of range and unused
    const size_t c = (size_t)i + (size_t)j * nW + (size_t)k * nW * nH;
    float v = face[t];
    if (lo) v = v + inc[c - sA];
    if (hi) v = v + inc[c];
    face[t] = v;
}

// add_wind_kernel :337-390, restricted to the border slabs it touches
__global__ void __launch_bounds__(BS) k_wind(float* __restrict__ U, float* __restrict__ V, float* __restrict__ W,
                                             int nW, int nH, int nD, float dx, int xwind, int4 slabs, int nslab,
                                             float offset, float freq, float tamp, float amp, float dirv, float dt,
                                             const float* __restrict__ offp) {
    // offp: the offset from device memory (captured steps), else `offset`
    if (offp) offset = *offp;
    // slab axis is x (xwind) or z; the other two axes span [0, n+1)
    const int a = xwind ? nH + 1 : nW + 1; // fastest of the remaining axes
    const int b = xwind ? nD + 1 : nH + 1;
    const int t = blockIdx.x * BS + threadIdx.x;
    if (t >= a * b * nslab) return;
    const int q0 = t % a, q1 = (t / a) % b, sl = t / (a * b);
    const int sv = sl == 0 ? slabs.x : (sl == 1 ? slabs.y : (sl == 2 ? slabs.z : slabs.w));
    int i, j, k;
    if (xwind) { i = sv; j = q0; k = q1; } else { i = q0; j = q1; k = sv; }
    // U origin (0,.5,.5), V origin (.5,0,.5), W origin (.5,.5,0)
    const f3 pos = mk3(((float)i + 0.0f) * dx, ((float)j + 0.0f) * dx, ((float)k + 0.0f) * dx + offset);
    const float nU = tamp * perlin(pos, freq, 123);
    const float nV = tamp * perlin(pos, freq, 456);
    const float nW_ = tamp * perlin(pos, freq, 789);
    if (j < nH && k < nD) {
        float& u = U[i + j * (nW + 1) + k * (nW + 1) * nH];
        u = u * (1.0f - dt) + (amp * (xwind ? (dirv + nU) : nU)) * dt;
    }
    if (i < nW && k < nD) {
        float& v = V[i + j * nW + k * nW * (nH + 1)];
        v = v * (1.0f - dt) + (amp * nV) * dt;
    }
    if (i < nW && j < nH) {
        float& w = W[i + j * nW + k * nW * nH];
        w = w * (1.0f - dt) + (amp * (xwind ? nW_ : (dirv + nW_))) * dt;
    }
}

// compute_turbulence_kernel :295-335 for the faces of one velocity grid G
template <int AXIS>
__global__ void __launch_bounds__(BS) k_turb_noise(Gv T, Gv U, Gv V, Gv W, Gv G, float* __restrict__ out, float dx,
                                                   float offset, float freq, int seed, float amp, float tr0, float tr1,
                                                   float vr0, float vr1, const float* __restrict__ offp) {
    if (offp) offset = *offp;
    const int n = G.nx * G.ny * G.nz;
    const int t = blockIdx.x * BS + threadIdx.x;
    if (t >= n) return;
    const int i = t % G.nx, j = (t / G.nx) % G.ny, k = t / (G.nx * G.ny);
    const f3 pos = g2o(G, dx, i, j, k);
    const float tm = fitf(trilinear(T, dx, pos), tr0, tr1, 1.0f, 0.0f);
    const f3 vel = mk3(trilinear(U, dx, pos), trilinear(V, dx, pos), trilinear(W, dx, pos));
    const float vm = fitf(len3(vel), vr0, vr1, 0.0f, 1.0f);
    const float mask = tm * vm;
    f3 np = mk3((float)i, (float)j, (float)k);
    if (AXIS == 0) np.x = (float)i + offset;
    if (AXIS == 1) np.y = (float)j + offset;
    if (AXIS == 2) np.z = (float)k + offset;
    out[t] = mask * amp * perlin(np, freq, seed);
}

// add_curl_noise_kernel :271-293 (per-cell increments into tmp planes 4..6)
__global__ void __launch_bounds__(BS) k_curl(const float* __restrict__ Ub, const float* __restrict__ Vb,
                                             const float* __restrict__ Wb, float* __restrict__ tmp, int nW, int nH,
                                             int nD, float dx, float dt, float eps) {
    const int t = blockIdx.x * BS + threadIdx.x;
    const int n = nW * nH * nD;
    if (t >= n) return;
    const int i = t % nW, j = (t / nW) % nH, k = t / (nW * nH);
    const int u0 = i + j * (nW + 1) + k * (nW + 1) * nH;
    const int v0 = i + j * nW + k * nW * (nH + 1);
    const int w0 = t;
    const float vb = Vb[v0];
    const float dudx = (Ub[u0 + 1] - vb) / dx;
    const float dvdy = (Vb[v0 + nW] - vb) / dx;
    const float dwdz = (Wb[w0 + nW * nH] - vb) / dx;
    f3 c = mk3(dvdy - dwdz, dwdz - dudx, dudx - dvdy);
    const f3 a = mk3((Ub[u0] + Ub[u0 + 1]) / 2.0f, (vb + Vb[v0 + nW]) / 2.0f, (Wb[w0] + Wb[w0 + nW * nH]) / 2.0f);
    const float avgSpeed = len3(a);
    const float den = len3(c) + eps;
    c = mk3(avgSpeed * c.x / den, avgSpeed * c.y / den, avgSpeed * c.z / den);
    const float h = 0.5f * dt;
    const size_t N = (size_t)n;
    tmp[4 * N + t] = h * c.x;
    tmp[5 * N + t] = h * c.y;
    tmp[6 * N + t] = h * c.z;
}

} // namespace

#define LAUNCH(kern, n, ...)                                                                  \
    do {                                                                                      \
        if ((n) > 0) {                                                                        \
            hipLaunchKernelGGL(kern, dim3(blocks_for((n), BS)), dim3(BS), 0, s.stream, __VA_ARGS__); \
            BLZ_HIP(hipGetLastError());                                                       \
        }                                                                                     \
    } while (0)

int launch_cooldown(Solver& s, float dt) {
    const float f = 1.0f - dt * s.s.cooling_rate;
    LAUNCH(k_scale, s.f[F_T].count(), s.f[F_T].front(), s.f[F_T].count(), f);
    return BLZ_OK;
}

int launch_drag(Solver& s, float dt) {
    const float f = 1.0f - dt * s.s.drag;
    for (int g = F_U; g <= F_W; ++g) LAUNCH(k_scale, s.f[g].count(), s.f[g].front(), s.f[g].count(), f);
    return BLZ_OK;
}

template <class S>
static int buoyancy(Solver& s, const S& sc, float dt) {
    const Field& V = s.f[F_V];
    LAUNCH(k_buoyancy<S>, V.count(), s.f[F_T].view(), V.front(), V.nx, V.ny, V.nz, V.ox, V.oy, V.oz, sc, s.dx,
           s.s.buoyancy, s.s.gravity, dt);
    return BLZ_OK;
}

int launch_buoyancy(Solver& s, float dt) {
    float inv = 0.0f;
    return dx_pow2(s.dx, &inv) ? buoyancy(s, MulDx{inv}, dt) : buoyancy(s, DivDx{s.dx}, dt);
}

void buoyancy_jobs(Solver& s, float dt, std::vector<BuoyJob>* jobs) {
    float inv = 0.0f;
    const bool mul = dx_pow2(s.dx, &inv);
    const Field& V = s.f[F_V];
    BuoyJob j;
    j.T = s.f[F_T].view();
    j.V = V.front();
    j.vnx = V.nx;
    j.vny = V.ny;
    j.vnz = V.nz;
    j.vox = V.ox;
    j.voy = V.oy;
    j.voz = V.oz;
    j.s = mul ? inv : s.dx;
    j.dx = s.dx;
    j.beta = s.s.buoyancy;
    j.g = s.s.gravity;
    j.dt = dt;
    j.first_block = 0;
    j.n = (int)V.count();
    jobs[mul ? 1 : 0].push_back(j);
}

int launch_buoyancy_multi(hipStream_t st, bool mul, const BuoyJob* tab, int n, int blocks) {
    if (n < 1 || blocks < 1) return BLZ_OK;
    if (mul) hipLaunchKernelGGL(k_buoyancy_multi<MulDx>, dim3((unsigned)blocks), dim3(BS), 0, st, tab, n);
    else hipLaunchKernelGGL(k_buoyancy_multi<DivDx>, dim3((unsigned)blocks), dim3(BS), 0, st, tab, n);
    BLZ_HIP(hipGetLastError());
    return BLZ_OK;
}

static int gather3(Solver& s, bool range2) {
    const size_t N = (size_t)s.W * s.H * s.D;
    const float* inc = s.tmp + 4 * N;
    if (range2) {
        LAUNCH((k_face_gather<0, true>), s.f[F_U].count(), s.f[F_U].front(), inc, s.W, s.H, s.D);
        LAUNCH((k_face_gather<1, true>), s.f[F_V].count(), s.f[F_V].front(), inc + N, s.W, s.H, s.D);
        LAUNCH((k_face_gather<2, true>), s.f[F_W].count(), s.f[F_W].front(), inc + 2 * N, s.W, s.H, s.D);
    } else {
        LAUNCH((k_face_gather<0, false>), s.f[F_U].count(), s.f[F_U].front(), inc, s.W, s.H, s.D);
        LAUNCH((k_face_gather<1, false>), s.f[F_V].count(), s.f[F_V].front(), inc + N, s.W, s.H, s.D);
        LAUNCH((k_face_gather<2, false>), s.f[F_W].count(), s.f[F_W].front(), inc + 2 * N, s.W, s.H, s.D);
    }
    return BLZ_OK;
}

int launch_vorticity(Solver& s, float dt) {
    if (s.W < 5 || s.H < 5 || s.D < 5) return BLZ_OK; // no active cell
    const float eps = 1e-20f * (1.0f / (s.dx * dt));
    const Gv T = s.f[F_T].view(), U = s.f[F_U].view(), V = s.f[F_V].view(), W = s.f[F_W].view();
    float inv = 0.0f;
    const size_t n = (size_t)s.W * s.H * s.D;
    if (dx_pow2(s.dx, &inv)) {
        const MulDx sc{inv};
        LAUNCH(k_vort_centres<MulDx>, n, T, U, V, W, sc, s.dx, s.tmp, s.W, s.H, s.D);
    } else {
        const DivDx sc{s.dx};
        LAUNCH(k_vort_centres<DivDx>, n, T, U, V, W, sc, s.dx, s.tmp, s.W, s.H, s.D);
    }
    LAUNCH(k_vort_omega, (size_t)(s.W - 2) * (s.H - 2) * (s.D - 2), s.dx, s.tmp, s.W, s.H, s.D);
    LAUNCH(k_vort_force, (size_t)(s.W - 4) * (s.H - 4) * (s.D - 4), s.dx, s.tmp, s.W, s.H, s.D,
           s.s.vorticity_conf, dt, eps);
    return gather3(s, true);
}

// host offsets of the wind and turbulence noise at `time` (the values a
// captured step reads from device memory)
void noise_offsets(const Solver& s, float time, float out[2]) {
    const float sx = s.s.domain_max[0] - s.s.domain_min[0];
    const float sy = s.s.domain_max[1] - s.s.domain_min[1];
    const float sz = s.s.domain_max[2] - s.s.domain_min[2];
    const float avgDom = (sx + sy + sz) / 3.0f;
    out[0] = fmodf(time * avgDom * s.s.wind_speed, 25000.0f);
    const float avgGrid = (float)(s.W + s.H + s.D) / 3.0f;
    out[1] = fmodf(time * avgGrid, 25000.0f);
}

int launch_wind(Solver& s, float time, float dt) {
    const bool xwind = s.s.wind_dir <= 1;
    const float freq = 1.0f / s.s.wind_turb_scale;
    const float sx = s.s.domain_max[0] - s.s.domain_min[0];
    const float sy = s.s.domain_max[1] - s.s.domain_min[1];
    const float sz = s.s.domain_max[2] - s.s.domain_min[2];
    const float avgDom = (sx + sy + sz) / 3.0f;
    const float offset = fmodf(time * avgDom * s.s.wind_speed, 25000.0f);
    const float dirv = xwind ? (s.s.wind_dir == 0 ? 1.0f : -1.0f) : (s.s.wind_dir == 2 ? 1.0f : -1.0f);
    const int N = xwind ? s.W : s.D; // slab axis has N+1 face planes
    int sl[4] = {0, 0, 0, 0}, ns = 0;
    for (int v = 0; v <= N && ns < 4; ++v)
        if (v < 2 || v >= N - 1) sl[ns++] = v;
    const size_t a = xwind ? (size_t)s.H + 1 : (size_t)s.W + 1;
    const size_t b = xwind ? (size_t)s.D + 1 : (size_t)s.H + 1;
    LAUNCH(k_wind, a * b * ns, s.f[F_U].front(), s.f[F_V].front(), s.f[F_W].front(), s.W, s.H, s.D, s.dx,
           xwind ? 1 : 0, make_int4(sl[0], sl[1], sl[2], sl[3]), ns, offset, freq, s.s.wind_turb_amp, s.s.wind_amp,
           dirv, dt, s.dev_off ? s.dev_off : nullptr);
    return BLZ_OK;
}

int launch_turbulence(Solver& s, float time, float dt) {
    const float freq = 1.0f / s.s.turbulence_scale;
    const float avgGrid = (float)(s.W + s.H + s.D) / 3.0f;
    const float offset = fmodf(time * avgGrid, 25000.0f);
    const float amp = s.s.turbulence_amp;
    const float* tr = s.s.turb_mask_temp_ramp;
    const float* vr = s.s.turb_mask_vel_ramp;
    const Gv T = s.f[F_T].view(), U = s.f[F_U].view(), V = s.f[F_V].view(), W = s.f[F_W].view();
    LAUNCH(k_turb_noise<0>, s.f[F_U].count(), T, U, V, W, U, s.f[F_U].back(), s.dx, offset, freq, s.s.noise_seed, amp,
           tr[0], tr[1], vr[0], vr[1], s.dev_off ? s.dev_off + 1 : nullptr);
    LAUNCH(k_turb_noise<1>, s.f[F_V].count(), T, U, V, W, V, s.f[F_V].back(), s.dx, offset, freq, s.s.noise_seed, amp,
           tr[0], tr[1], vr[0], vr[1], s.dev_off ? s.dev_off + 1 : nullptr);
    LAUNCH(k_turb_noise<2>, s.f[F_W].count(), T, U, V, W, W, s.f[F_W].back(), s.dx, offset, freq, s.s.noise_seed, amp,
           tr[0], tr[1], vr[0], vr[1], s.dev_off ? s.dev_off + 1 : nullptr);
    const float eps = 1e-20f * (1.0f / (s.dx * dt));
    LAUNCH(k_curl, (size_t)s.W * s.H * s.D, s.f[F_U].back(), s.f[F_V].back(), s.f[F_W].back(), s.tmp, s.W, s.H, s.D,
           s.dx, dt, eps);
    return gather3(s, false);
}

} // namespace blz
