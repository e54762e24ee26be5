// capi.hip — solver handle lifecycle, step orchestration and the C ABI.
//
// Reference: FluidSolver (asiroisvigneux/Blaze src/FluidSolver.cu:671-955).
// The reference keeps device-heap Grid objects and synchronises after every
// launch; here the host owns front/back pointers (a swap is a pointer swap),
// all work of a step is enqueued on the handle's stream, and the host only
// waits where it needs a value (CFL substep count) or to stay a few sweep
// batches ahead of a residual-driven solve, whose convergence test runs on
// the device.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <new>

#include "solver.h"

namespace blz {

static thread_local std::string g_err;
void set_error(const std::string& msg) { g_err = msg; }

void free_p2g_scratch(Solver& s);

Solver::~Solver() {
    if (stream) (void)hipStreamSynchronize(stream);
    for (auto& F : f)
        for (auto& b : F.buf)
            if (b) (void)hipFree(b);
    if (prb) (void)hipFree(prb);
    if (tmp) (void)hipFree(tmp);
    if (red) (void)hipFree(red);
    if (red_host) (void)hipHostFree(red_host);
    if (ctl) (void)hipFree(ctl);
    if (ctl_host) (void)hipHostFree(ctl_host);
    if (trace) (void)hipFree(trace);
    if (ws_net) delete ws_net;
    if (ws_buf) (void)hipFree(ws_buf);
    if (rnd_buf) (void)hipFree(rnd_buf);
    for (int q = 0; q < 2; ++q) {
        if (stage_ev[q]) (void)hipEventSynchronize(stage_ev[q]);
        if (stage_host[q]) (void)hipHostFree(stage_host[q]);
        if (stage_dev[q]) (void)hipFree(stage_dev[q]);
        if (stage_ev[q]) (void)hipEventDestroy(stage_ev[q]);
    }
    if (copy_stream) (void)hipStreamDestroy(copy_stream);
    if (zero_flag) (void)hipFree(zero_flag);
    if (done_host) (void)hipHostFree((void*)done_host);
    for (hipEvent_t e : batch_ev)
        if (e) (void)hipEventDestroy(e);
    if (part) (void)hipFree(part);
    free_p2g_scratch(*this);
    for (auto& e : evpool)
        if (e) (void)hipEventDestroy(e);
    if (stream) (void)hipStreamDestroy(stream);
}

static int validate(const blz_settings& s) {
    for (int a = 0; a < 3; ++a)
        if (s.grid_res[a] < 4 || s.grid_res[a] > 4096) {
            set_error("grid_res must be in [4, 4096] on every axis");
            return BLZ_E_INVALID;
        }
    if (s.grid_res[0] % 2) {
        set_error("grid_res[0] (W) must be even (red/black split layout)");
        return BLZ_E_INVALID;
    }
    const double cells = (double)(s.grid_res[0] + 1) * (s.grid_res[1] + 1) * (s.grid_res[2] + 1);
    if (cells * 5.0 >= 4294967296.0 || cells >= 2147483648.0) {
        set_error("grid too large for 32-bit cell indices");
        return BLZ_E_INVALID;
    }
    const double padded = (double)(s.grid_res[0] / 2 + 2) * (s.grid_res[1] + 2) * ((s.grid_res[2] + 3) / 4 + 2) * 4;
    if (16.0 * padded >= 4294967296.0) {
        set_error("grid too large for 32-bit pressure plane offsets");
        return BLZ_E_INVALID;
    }
    if (!(s.dx > 0.0f) || !(s.dt > 0.0f) || !(s.density > 0.0f)) {
        set_error("dx, dt and density must be positive");
        return BLZ_E_INVALID;
    }
    if (s.max_iter_solve < 0) {
        set_error("max_iter_solve must be >= 0");
        return BLZ_E_INVALID;
    }
    if (s.max_substeps > MAX_SOLVES) {
        set_error("max_substeps must be <= 64");
        return BLZ_E_INVALID;
    }
    return BLZ_OK;
}

int Solver::init(const blz_settings& st_) {
    if (int e = validate(st_)) return e;
    s = st_;
    if (s.residual_check_every <= 0) s.residual_check_every = 10;
    if (s.max_substeps <= 0) s.max_substeps = 8;
    W = s.grid_res[0];
    H = s.grid_res[1];
    D = s.grid_res[2];
    dx = s.dx;
    smoke = s.smoke_density > 0.0f;
    BLZ_HIP(hipGetDevice(&device));
    BLZ_HIP(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
    // create_grids_kernel origins (src/FluidSolver.cu:36-43)
    const int dims[F_N][3] = {{W, H, D}, {W + 1, H, D}, {W, H + 1, D}, {W, H, D + 1}, {W, H, D}};
    const float org[F_N][3] = {{.5f, .5f, .5f}, {0.f, .5f, .5f}, {.5f, 0.f, .5f}, {.5f, .5f, 0.f}, {.5f, .5f, .5f}};
    for (int g = 0; g < F_N; ++g) {
        if (g == F_S && !smoke) continue;
        Field& F = f[g];
        F.nx = dims[g][0];
        F.ny = dims[g][1];
        F.nz = dims[g][2];
        F.ox = org[g][0];
        F.oy = org[g][1];
        F.oz = org[g][2];
        F.flip = (g == F_V);
        for (int b = 0; b < 2; ++b) {
            if (hipMalloc(&F.buf[b], F.count() * sizeof(float)) != hipSuccess) {
                set_error("hipMalloc failed for a grid buffer");
                return BLZ_E_NOMEM;
            }
            BLZ_HIP(hipMemset(F.buf[b], 0, F.count() * sizeof(float)));
        }
    }
    const size_t N = (size_t)W * H * D;
    L.Wh = W / 2;
    L.H = H;
    L.D = D;
    L.Wp = W / 2 + 2;
    L.Hp = H + 2;
    L.Dg = (D + 3) / 4;
    Np = (size_t)L.Wp * L.Hp * (L.Dg + 2) * 4;
    // eight planes when their byte offsets (plus a tail that absorbs the
    // fused sweep's reads two rows past the last group) fit 32 bits, else
    // six (the warm start's judge then sweeps in place), else four (fused
    // sweeps unavailable)
    const size_t tail = (size_t)2 * L.Wp * L.Hp * 4;
    const bool fit8 = (8 * Np + tail) * 4 < ((size_t)1 << 32) && W <= 4096;
    const bool fit6 = (6 * Np + tail) * 4 < ((size_t)1 << 32) && W <= 4096;
    prb_floats = fit8 ? 8 * Np + tail : (fit6 ? 6 * Np + tail : 4 * Np);
    judge_set = fit8 ? 6 : 4;
    // tmp also holds the two home pressure planes while an autotuner times
    // its candidates; on small grids the padded planes outgrow 7 * N
    tmp_floats = std::max(7 * N, 2 * Np);
    if (hipMalloc(&prb, prb_floats * sizeof(float)) != hipSuccess || hipMalloc(&tmp, tmp_floats * sizeof(float)) != hipSuccess ||
        hipMalloc(&red, (8 + 2 * RESIDUAL_BLOCKS) * sizeof(double)) != hipSuccess) {
        set_error("hipMalloc failed for solver scratch");
        return BLZ_E_NOMEM;
    }
    p.c[0] = prb;
    p.c[1] = prb + Np;
    rhs.c[0] = prb + 2 * Np;
    rhs.c[1] = prb + 3 * Np;
    // pressure zeroed once (:715-716): +0 inside, -0 in the ghost shell
    if (int e = launch_init_pressure(*this)) return e;
    BLZ_HIP(hipMemset(red, 0, (8 + 2 * RESIDUAL_BLOCKS) * sizeof(double)));
    BLZ_HIP(hipStreamSynchronize(stream));
    BLZ_HIP(hipHostMalloc((void**)&red_host, 8 * sizeof(double), hipHostMallocDefault));
    if (hipMalloc(&ctl, MAX_SOLVES * sizeof(SolveCtl)) != hipSuccess ||
        hipMalloc(&zero_flag, sizeof(int)) != hipSuccess) {
        set_error("hipMalloc failed for solver control blocks");
        return BLZ_E_NOMEM;
    }
    BLZ_HIP(hipMemset(ctl, 0, MAX_SOLVES * sizeof(SolveCtl)));
    BLZ_HIP(hipMemset(zero_flag, 0, sizeof(int)));
    BLZ_HIP(hipHostMalloc((void**)&ctl_host, MAX_SOLVES * sizeof(SolveCtl), hipHostMallocDefault));
    void* dh = nullptr;
    BLZ_HIP(hipHostMalloc(&dh, sizeof(int), hipHostMallocMapped | hipHostMallocCoherent));
    done_host = (volatile int*)dh;
    *done_host = 0;
    BLZ_HIP(hipHostGetDevicePointer((void**)&done_host_dev, dh, 0));
    for (hipEvent_t& e : batch_ev) BLZ_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    BLZ_HIP(hipDeviceSynchronize());
    std::memset(&st, 0, sizeof(st));
    st.residual = -1.0;
    st.residual_initial = -1.0;
    return BLZ_OK;
}

int Solver::max_abs_velocity(float* out) {
    if (int e = launch_max_abs_velocity(*this, red + 1)) return e;
    BLZ_HIP(hipMemcpyAsync(red_host + 1, red + 1, sizeof(double), hipMemcpyDeviceToHost, stream));
    BLZ_HIP(hipStreamSynchronize(stream));
    *out = (float)red_host[1];
    return BLZ_OK;
}

int Solver::span_begin(int* e0) {
    *e0 = -1;
    if (!timing) return BLZ_OK;
    if (ev_used == evpool.size()) {
        hipEvent_t e;
        BLZ_HIP(hipEventCreate(&e));
        evpool.push_back(e);
    }
    *e0 = (int)ev_used++;
    BLZ_HIP(hipEventRecord(evpool[*e0], stream));
    return BLZ_OK;
}

int Solver::span_end(int kind, int e0, int launches, int solve, int batch, int sweeps) {
    if (!timing || e0 < 0) return BLZ_OK;
    int e1 = -1;
    if (int e = span_begin(&e1)) return e;
    spans.push_back(Span{kind, e0, e1, launches, solve, batch, sweeps});
    return BLZ_OK;
}

// after a stream synchronisation: fold the recorded spans into the stats
int Solver::resolve_spans() {
    for (const Span& sp : spans) {
        float ms = 0.0f;
        BLZ_HIP(hipEventElapsedTime(&ms, evpool[sp.e0], evpool[sp.e1]));
        if (sp.kind < BLZ_NUM_STAGES) st.stage_ms[sp.kind] += ms;
        else if (sp.kind == SPAN_SOURCE) st.source_ms += ms;
        else if (sp.kind == SPAN_STEP) st.step_ms += ms;
        else if (sp.kind == SPAN_RBGS) {
            // a batch enqueued after its solve had converged ran as no-ops
            if (sp.solve >= 0 && sp.solve < (int)solves.size() && solves[sp.solve] >= 0 &&
                sp.batch >= ctl_host[sp.solve].checks - 1)
                continue;
            st.rbgs_kernel_ms += ms;
            st.rbgs_launches += sp.launches;
            st.rbgs_sweeps_timed += sp.sweeps;
        }
    }
    spans.clear();
    ev_used = 0;
    return BLZ_OK;
}

// copy the control blocks of this step's residual-driven solves to the host
int Solver::enqueue_solve_stats() {
    if (!solves.empty())
        BLZ_HIP(hipMemcpyAsync(ctl_host, ctl, solves.size() * sizeof(SolveCtl), hipMemcpyDeviceToHost, stream));
    return BLZ_OK;
}

int Solver::finalize() {
    if (!solves.empty()) st.warm_start_used = 0;
    for (size_t q = 0; q < solves.size(); ++q) {
        int it = solves[q];
        double r = -1.0, r0 = -1.0;
        if (it < 0) {
            it = -it - 1;
            st.rbgs_sweeps_executed += it;
        } else {
            st.rbgs_sweeps_executed += ctl_host[q].swept;
            it = ctl_host[q].iters;
            r = ctl_host[q].res;
            r0 = ctl_host[q].res0;
            st.residual_checks += ctl_host[q].checks;
        }
        st.rbgs_iters += it;
        st.rbgs_iters_last = it;
        st.residual = r;
        st.residual_initial = r0;
        if (solves[q] >= 0) {
            st.warm_start_used += ctl_host[q].ws_kept;
            st.warm_start_fallbacks += ctl_host[q].ws_rejected;
        }
    }
    // non-finite state (the residual kernels map NaN to +inf): the
    // reference prints and exits on device errors (src/utils/helper_cuda.h:
    // 576-591); here the step reports BLZ_E_NUMERIC and the caller decides
    int bad = -1;
    for (size_t q = 0; q < solves.size() && bad < 0; ++q)
        if (solves[q] >= 0 && !(std::isfinite(ctl_host[q].res0) && std::isfinite(ctl_host[q].res))) bad = (int)q;
    const int e = resolve_spans();
    if (!solves.empty()) last_nsolves = (int)solves.size();
    solves.clear();
    if (e) return e;
    if (bad >= 0) {
        set_error("non-finite state: the pressure residual of solve " + std::to_string(bad) +
                  " is not finite (NaN or infinity in the velocity, pressure or right-hand side)");
        return BLZ_E_NUMERIC;
    }
    return BLZ_OK;
}

void clear_step_stats(blz_stats& st) {
    for (double& v : st.stage_ms) v = 0.0;
    st.step_ms = 0.0;
    st.rbgs_kernel_ms = 0.0;
    st.rbgs_launches = 0;
    st.rbgs_sweeps_timed = 0;
    st.rbgs_iters = 0;
    st.rbgs_sweeps_executed = 0;
    st.residual_checks = 0;
}

// Time `sweeps` sweeps of every RBGS variant on the solver's rhs with the
// pressure set to 1.0 (the pressure is restored afterwards; all variants
// give the same bits); *best = the fastest. Synchronises the stream.
int Solver::autotune(int sweeps, int* best, float* ms_per_sweep) {
    static const int cand[] = {2, 4, 8, 102, 103, 202, 203, 402, 403, 1103, 1104, 1105, 1106, 1203, 1204, 1403, 1404};
    sweeps = std::max(2, sweeps & ~1);
    const float dt = s.dt;
    if (tmp_floats < 2 * Np) {
        set_error("autotune: scratch smaller than the pressure planes it saves");
        return BLZ_E_INVALID;
    }
    BLZ_HIP(hipStreamSynchronize(stream));
    BLZ_HIP(hipMemcpyAsync(tmp, prb, 2 * Np * sizeof(float), hipMemcpyDeviceToDevice, stream));
    // timing state: pressure 1.0 everywhere (the solver's rhs), so no
    // numerator takes the exact division: at a first solve the pressure
    // front decays through tiny values that send waves to the exact run for
    // a few sweeps, which would time that transient instead of the kernels
    if (int e = launch_fill_pressure(*this, 1.0f)) return e;
    hipEvent_t e0, e1;
    BLZ_HIP(hipEventCreate(&e0));
    BLZ_HIP(hipEventCreate(&e1));
    const int keep = rbgs_variant;
    int bestv = 8, rc = BLZ_OK, k = 0;
    float bestms = 1e30f;
    for (int v : cand) {
        if (v >= 100 && !fused_available(*this)) continue;
        rbgs_variant = v;
        int launches = 0, kind = 0;
        rc = launch_rbgs_sweeps(*this, dt, 2, zero_flag, &launches, &kind);
        if (!rc) rc = hipEventRecord(e0, stream) == hipSuccess ? BLZ_OK : BLZ_E_HIP;
        if (!rc) rc = launch_rbgs_sweeps(*this, dt, sweeps, zero_flag, &launches, &kind);
        if (!rc) rc = hipEventRecord(e1, stream) == hipSuccess ? BLZ_OK : BLZ_E_HIP;
        if (!rc) rc = hipEventSynchronize(e1) == hipSuccess ? BLZ_OK : BLZ_E_HIP;
        if (rc) break;
        float ms = 0.0f;
        (void)hipEventElapsedTime(&ms, e0, e1);
        ms /= (float)sweeps;
        if (ms_per_sweep) ms_per_sweep[k] = ms;
        ++k;
        if (ms < bestms) {
            bestms = ms;
            bestv = v;
        }
    }
    rbgs_variant = keep;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipMemcpyAsync(prb, tmp, 2 * Np * sizeof(float), hipMemcpyDeviceToDevice, stream);
    BLZ_HIP(hipStreamSynchronize(stream));
    if (rc) {
        set_error("RBGS autotune failed");
        return rc;
    }
    *best = bestv;
    return BLZ_OK;
}

// gs_solve_kernel (src/FluidSolver.cu:517-599): a fixed sweep count, or
// with residual_tol > 0 the loop
//     r = residual; it = 0
//     while (r > tol && it < cap) { n = min(every, cap - it); n sweeps; it += n; r = residual }
// run on the device: every batch of sweeps is followed by a residual check
// that updates the solve's control block, and once it is done the remaining
// kernels return at entry. The host enqueues batches while the solve is not
// known to have converged, at most SOLVE_LOOKAHEAD past the last batch it
// has seen finish. Statistics: solves[q] records the solve, ctl_host[q]
// its result after the step.
int Solver::project_solve(float dt_) {
    const int cap = s.max_iter_solve;
    // variant 0: the autotuner's choice for this grid, made at the first solve
    if (rbgs_variant == 0 && rbgs_auto == 0 && cap > 0)
        if (int e = autotune(10, &rbgs_auto, nullptr)) return e;
    const int q = (int)solves.size();
    if (q >= MAX_SOLVES) {
        set_error("too many pressure solves enqueued without synchronisation");
        return BLZ_E_INVALID;
    }
    auto sweeps = [&](int n, const int* stop, int batch) -> int {
        int e0 = -1, launches = 0, kind = 0;
        if (int e = span_begin(&e0)) return e;
        if (int e = launch_rbgs_sweeps(*this, dt_, n, stop, &launches, &kind)) return e;
        st.rbgs_kernel = kind;
        return span_end(SPAN_RBGS, e0, launches, q, batch, n);
    };
    if (!(s.residual_tol > 0.0f)) {
        solves.push_back(-cap - 1);
        return sweeps(cap, zero_flag, 0);
    }
    solves.push_back(cap);
    const double tol = (double)s.residual_tol;
    const int every = s.residual_check_every;
    const int seq = ++solve_seq;
    SolveCtl* c = ctl + q;
    // adaptive check spacing for plain solves (a warm start's judge reads
    // every batch's check); the fixed-point stop for all
    const bool warm = ws_net && s.warm_start && cap > 0;
    if (int e = launch_ctl_reset(*this, c, s.residual_check_adaptive && !warm, s.residual_check_adaptive != 0))
        return e;
    if (int e = launch_residual_check(*this, dt_, c, done_host_dev, seq, tol, cap, 0, 1)) return e;
    int it0 = 0;
    if (ws_net && s.warm_start && cap > 0) {
        it0 = std::min(every, cap);
        if (int e = warm_start_apply(*this, dt_, c, done_host_dev, seq, tol, cap, it0)) return e;
    }
    int b = 0;
    for (int it = it0; it < cap; ++b) {
        if (*done_host == seq) break;
        const int n = std::min(every, cap - it);
        if (int e = sweeps(n, &c->done, b)) return e;
        it += n;
        if (int e = launch_residual_check(*this, dt_, c, done_host_dev, seq, tol, cap, n, 0)) return e;
        // warm_start = N >= 2: the learned correction again every N sweeps
        if (ws_net && s.warm_start >= 2 && it % s.warm_start == 0 && it < cap) {
            const int k = std::min(every, cap - it);
            if (int e = warm_start_apply(*this, dt_, c, done_host_dev, seq, tol, cap, k, true)) return e;
            it += k;
        }
        BLZ_HIP(hipEventRecord(batch_ev[b % (SOLVE_LOOKAHEAD + 1)], stream));
        if (b >= SOLVE_LOOKAHEAD) BLZ_HIP(hipEventSynchronize(batch_ev[(b - SOLVE_LOOKAHEAD) % (SOLVE_LOOKAHEAD + 1)]));
    }
    return BLZ_OK;
}

int Solver::run_stage(int stage, float time, float dt_) {
    int e0 = -1;
    if (int e = span_begin(&e0)) return e;
    int rc = BLZ_OK;
    switch (stage) {
    case BLZ_STAGE_COOLDOWN: rc = launch_cooldown(*this, dt_); break;
    case BLZ_STAGE_DRAG: rc = launch_drag(*this, dt_); break;
    case BLZ_STAGE_VORTICITY: rc = launch_vorticity(*this, dt_); break;
    case BLZ_STAGE_BUOYANCY: rc = launch_buoyancy(*this, dt_); break;
    case BLZ_STAGE_WIND: rc = launch_wind(*this, time, dt_); break;
    case BLZ_STAGE_TURBULENCE: rc = launch_turbulence(*this, time, dt_); break;
    case BLZ_STAGE_DIVERGENCE: rc = launch_divergence(*this); break;
    case BLZ_STAGE_PRESSURE: rc = project_solve(dt_); break;
    case BLZ_STAGE_PROJECT_UPDATE: rc = launch_pressure_update(*this, dt_); break;
    case BLZ_STAGE_ADVECT: rc = launch_advect(*this, dt_); break;
    default: set_error("unknown stage"); return BLZ_E_INVALID;
    }
    if (rc) return rc;
    return span_end(stage, e0, 0);
}

// FluidSolver::step, src/FluidSolver.cu:858-955 (stage gates :868-925)
int Solver::substep(float time, float dt_) {
    if (std::fabs(s.cooling_rate) > 0.0f)
        if (int e = run_stage(BLZ_STAGE_COOLDOWN, time, dt_)) return e;
    if (s.drag > 0.0f)
        if (int e = run_stage(BLZ_STAGE_DRAG, time, dt_)) return e;
    if (std::fabs(s.vorticity_conf) > 0.0f)
        if (int e = run_stage(BLZ_STAGE_VORTICITY, time, dt_)) return e;
    if (std::fabs(s.buoyancy) > 0.0f)
        if (int e = run_stage(BLZ_STAGE_BUOYANCY, time, dt_)) return e;
    if (s.wind_amp > 0.0f)
        if (int e = run_stage(BLZ_STAGE_WIND, time, dt_)) return e;
    if (s.turbulence_amp > 0.0f)
        if (int e = run_stage(BLZ_STAGE_TURBULENCE, time, dt_)) return e;
    if (int e = run_stage(BLZ_STAGE_DIVERGENCE, time, dt_)) return e;
    if (int e = run_stage(BLZ_STAGE_PRESSURE, time, dt_)) return e;
    if (int e = run_stage(BLZ_STAGE_PROJECT_UPDATE, time, dt_)) return e;
    if (int e = run_stage(BLZ_STAGE_ADVECT, time, dt_)) return e;
    return BLZ_OK;
}

// One frame. With cfl_number > 0: n = ceil(max|u| dt / (cfl dx)) substeps
// (clamped to [1, max_substeps]) of dt/n at times time + q*(dt/n).
int Solver::step_enqueue(float time, bool sync) {
    clear_step_stats(st);
    spans.clear();
    solves.clear();
    ev_used = 0;
    const float dtf = s.dt;
    int n = 1;
    if (s.cfl_number > 0.0f) {
        float umax = 0.0f;
        if (int e = max_abs_velocity(&umax)) return e;
        if (!std::isfinite(umax)) {
            set_error("non-finite state: max |velocity| is not finite (NaN or infinity in U, V or W)");
            return BLZ_E_NUMERIC;
        }
        const float q = std::ceil(umax * dtf / (s.cfl_number * dx));
        n = !(q >= 1.0f) ? 1 : (q > (float)s.max_substeps ? s.max_substeps : (int)q);
    }
    const float dts = dtf / (float)n;
    int e0 = -1;
    if (int e = span_begin(&e0)) return e;
    for (int q = 0; q < n; ++q)
        if (int e = substep(time + (float)q * dts, dts)) return e;
    st.substeps = n;
    if (int e = span_end(SPAN_STEP, e0, 0)) return e;
    if (int e = enqueue_solve_stats()) return e;
    if (sync) {
        BLZ_HIP(hipStreamSynchronize(stream));
        if (int e = finalize()) return e;
    }
    return BLZ_OK;
}

int Solver::add_source(const float* part8, int n) {
    if (part8) {
        if (n < 0) {
            set_error("particle count must be >= 0");
            return BLZ_E_INVALID;
        }
        if (n > part_cap) {
            if (part) BLZ_HIP(hipFree(part));
            part = nullptr;
            part_cap = 0;
            if (hipMalloc(&part, (size_t)n * 8 * sizeof(float)) != hipSuccess) {
                set_error("hipMalloc failed for particles");
                return BLZ_E_NOMEM;
            }
            part_cap = n;
        }
        if (n > 0)
            BLZ_HIP(hipMemcpyAsync(part, part8, (size_t)n * 8 * sizeof(float), hipMemcpyHostToDevice, stream));
        npart = n;
        p2g_bound = p2g_entry_bound(part8, n, dx, smoke ? 5 : 4);
    }
    // the call returns with the splat done (its span resolved on its own)
    std::vector<Span> keep;
    keep.swap(spans);
    const size_t used = ev_used;
    int e0 = -1;
    if (int e = span_begin(&e0)) return e;
    p2g_overflow = nullptr;
    if (int e = launch_p2g(*this)) return e;
    if (int e = span_end(SPAN_SOURCE, e0, 0)) return e;
    BLZ_HIP(hipStreamSynchronize(stream));
    if (p2g_overflow) {
        unsigned ov = 0;
        BLZ_HIP(hipMemcpy(&ov, p2g_overflow, sizeof(unsigned), hipMemcpyDeviceToHost));
        if (ov) {
            set_error("particle-to-grid: footprint entries exceeded their bound");
            return BLZ_E_INVALID;
        }
    }
    st.source_ms = 0.0;
    if (int e = resolve_spans()) return e;
    spans.swap(keep);
    ev_used = used;
    return BLZ_OK;
}

// Stage the next source frame: copy into the pinned slot and start its
// upload on the copy stream (returns before the upload ends). The slot's
// previous upload is waited for first; the solver stream is not touched,
// so this may run on another host thread while a step is in flight.
int Solver::stage_source(const float* part8, int n) {
    if (n < 0 || (n > 0 && !part8)) {
        set_error("stage_source: bad particle array");
        return BLZ_E_INVALID;
    }
    if (!copy_stream) BLZ_HIP(hipStreamCreate(&copy_stream));
    const int q = stage_next;
    if (!stage_ev[q]) BLZ_HIP(hipEventCreateWithFlags(&stage_ev[q], hipEventDisableTiming));
    BLZ_HIP(hipEventSynchronize(stage_ev[q]));
    if (n > stage_cap[q]) {
        if (stage_host[q]) BLZ_HIP(hipHostFree(stage_host[q]));
        if (stage_dev[q]) BLZ_HIP(hipFree(stage_dev[q]));
        stage_host[q] = nullptr;
        stage_dev[q] = nullptr;
        stage_cap[q] = 0;
        if (hipHostMalloc(&stage_host[q], (size_t)n * 8 * sizeof(float), hipHostMallocDefault) != hipSuccess ||
            hipMalloc(&stage_dev[q], (size_t)n * 8 * sizeof(float)) != hipSuccess) {
            set_error("allocation failed for staged particles");
            return BLZ_E_NOMEM;
        }
        stage_cap[q] = n;
    }
    if (n > 0) {
        std::memcpy(stage_host[q], part8, (size_t)n * 8 * sizeof(float));
        stage_bound[q] = p2g_entry_bound(part8, n, dx, smoke ? 5 : 4);
        BLZ_HIP(hipMemcpyAsync(stage_dev[q], stage_host[q], (size_t)n * 8 * sizeof(float), hipMemcpyHostToDevice,
                               copy_stream));
    }
    BLZ_HIP(hipEventRecord(stage_ev[q], copy_stream));
    stage_count[q] = n;
    stage_ready = q;
    stage_next = q ^ 1;
    return BLZ_OK;
}

// addSource with the frame staged last: the solver stream waits for its
// upload (device-side), takes it over, and splats as add_source does
int Solver::add_source_staged() {
    if (stage_ready < 0) {
        set_error("add_source_staged: nothing staged");
        return BLZ_E_INVALID;
    }
    const int q = stage_ready, n = stage_count[q];
    if (n > part_cap) {
        if (part) BLZ_HIP(hipFree(part));
        part = nullptr;
        part_cap = 0;
        if (hipMalloc(&part, (size_t)n * 8 * sizeof(float)) != hipSuccess) {
            set_error("hipMalloc failed for particles");
            return BLZ_E_NOMEM;
        }
        part_cap = n;
    }
    BLZ_HIP(hipStreamWaitEvent(stream, stage_ev[q], 0));
    if (n > 0)
        BLZ_HIP(hipMemcpyAsync(part, stage_dev[q], (size_t)n * 8 * sizeof(float), hipMemcpyDeviceToDevice, stream));
    npart = n;
    p2g_bound = n > 0 ? stage_bound[q] : 0;
    stage_ready = -1;
    return add_source(nullptr, 0);
}

} // namespace blz

using blz::Solver;


#define CHECK_HANDLE(h)                              \
    do {                                             \
        if (!(h)) {                                  \
            blz::set_error("null solver handle");    \
            return BLZ_E_INVALID;                    \
        }                                            \
    } while (0)

extern "C" {

const char* blz_last_error(void) { return blz::g_err.c_str(); }

int blz_create(const blz_settings* s, blz_solver** out) {
    if (!s || !out) {
        blz::set_error("blz_create: null argument");
        return BLZ_E_INVALID;
    }
    *out = nullptr;
    blz_solver* h = new (std::nothrow) blz_solver();
    if (!h) {
        blz::set_error("out of host memory");
        return BLZ_E_NOMEM;
    }
    if (int e = h->s.init(*s)) {
        delete h;
        return e;
    }
    *out = h;
    return BLZ_OK;
}

int blz_destroy(blz_solver* h) {
    delete h;
    return BLZ_OK;
}

int blz_add_source(blz_solver* h, const float* part8, int count) {
    CHECK_HANDLE(h);
    return h->s.add_source(part8, count);
}

int blz_step(blz_solver* h, float time) {
    CHECK_HANDLE(h);
    return h->s.step_enqueue(time, true);
}

int blz_step_async(blz_solver* h, float time) {
    CHECK_HANDLE(h);
    return h->s.step_enqueue(time, false);
}

int blz_sync(blz_solver* h) {
    CHECK_HANDLE(h);
    BLZ_HIP(hipStreamSynchronize(h->s.stream));
    return h->s.finalize();
}

int blz_run_stage(blz_solver* h, int stage, float time, float dt) {
    CHECK_HANDLE(h);
    if (stage < 0 || stage >= BLZ_NUM_STAGES) {
        blz::set_error("stage out of range");
        return BLZ_E_INVALID;
    }
    blz::clear_step_stats(h->s.st);
    h->s.spans.clear();
    h->s.solves.clear();
    h->s.ev_used = 0;
    if (int e = h->s.run_stage(stage, time, dt)) return e;
    if (int e = h->s.enqueue_solve_stats()) return e;
    BLZ_HIP(hipStreamSynchronize(h->s.stream));
    return h->s.finalize();
}

static int field_geom(Solver& s, int which, float** dev, size_t* count, int* dims) {
    if (which == BLZ_FIELD_P || which == BLZ_FIELD_RHS) {
        // split padded layout (two planes of Np floats each)
        *dev = which == BLZ_FIELD_P ? s.p.c[0] : s.rhs.c[0];
        *count = (size_t)s.W * s.H * s.D;
        if (dims) { dims[0] = s.W; dims[1] = s.H; dims[2] = s.D; }
        return BLZ_OK;
    }
    if (which < 0 || which > BLZ_FIELD_DENSITY || (which == BLZ_FIELD_DENSITY && !s.smoke)) {
        blz::set_error("field not available");
        return BLZ_E_INVALID;
    }
    const blz::Field& F = s.f[which];
    *dev = F.front();
    *count = F.count();
    if (dims) { dims[0] = F.nx; dims[1] = F.ny; dims[2] = F.nz; }
    return BLZ_OK;
}

int blz_get_field(blz_solver* h, int which, void** dev_ptr, int dims[3]) {
    CHECK_HANDLE(h);
    float* d = nullptr;
    size_t n = 0;
    if (int e = field_geom(h->s, which, &d, &n, dims)) return e;
    if (dev_ptr) *dev_ptr = d;
    return BLZ_OK;
}

int blz_copy_field(blz_solver* h, int which, float* host) {
    CHECK_HANDLE(h);
    if (!host) {
        blz::set_error("null host buffer");
        return BLZ_E_INVALID;
    }
    Solver& s = h->s;
    float* d = nullptr;
    size_t n = 0;
    if (int e = field_geom(s, which, &d, &n, nullptr)) return e;
    if (which == BLZ_FIELD_P || which == BLZ_FIELD_RHS) {
        if (int e = blz::launch_p_to_natural(s, s.tmp, which == BLZ_FIELD_P ? s.p : s.rhs)) return e;
        d = s.tmp;
    }
    BLZ_HIP(hipMemcpyAsync(host, d, n * sizeof(float), hipMemcpyDeviceToHost, s.stream));
    BLZ_HIP(hipStreamSynchronize(s.stream));
    return BLZ_OK;
}

int blz_set_field(blz_solver* h, int which, const float* host) {
    CHECK_HANDLE(h);
    if (!host) {
        blz::set_error("null host buffer");
        return BLZ_E_INVALID;
    }
    Solver& s = h->s;
    float* d = nullptr;
    size_t n = 0;
    if (int e = field_geom(s, which, &d, &n, nullptr)) return e;
    if (which == BLZ_FIELD_P || which == BLZ_FIELD_RHS) {
        BLZ_HIP(hipMemcpyAsync(s.tmp, host, n * sizeof(float), hipMemcpyHostToDevice, s.stream));
        if (int e = blz::launch_natural_to_rb(s, which == BLZ_FIELD_P ? s.p : s.rhs, s.tmp)) return e;
    } else {
        BLZ_HIP(hipMemcpyAsync(d, host, n * sizeof(float), hipMemcpyHostToDevice, s.stream));
    }
    BLZ_HIP(hipStreamSynchronize(s.stream));
    return BLZ_OK;
}

int blz_get_stats(blz_solver* h, blz_stats* out) {
    CHECK_HANDLE(h);
    if (!out) {
        blz::set_error("null stats");
        return BLZ_E_INVALID;
    }
    *out = h->s.st;
    return BLZ_OK;
}

int blz_set_timing(blz_solver* h, int on) {
    CHECK_HANDLE(h);
    h->s.timing = on != 0;
    return BLZ_OK;
}

int blz_set_solve_limits(blz_solver* h, float residual_tol, int max_iter_solve, int residual_check_every) {
    CHECK_HANDLE(h);
    if (!(residual_tol >= 0.0f) || max_iter_solve < 0 || residual_check_every <= 0) {
        blz::set_error("blz_set_solve_limits: need residual_tol >= 0, max_iter_solve >= 0, residual_check_every > 0");
        return BLZ_E_INVALID;
    }
    BLZ_HIP(hipStreamSynchronize(h->s.stream));
    h->s.s.residual_tol = residual_tol;
    h->s.s.max_iter_solve = max_iter_solve;
    h->s.s.residual_check_every = residual_check_every;
    return BLZ_OK;
}

int blz_set_residual_trace(blz_solver* h, int max_entries) {
    CHECK_HANDLE(h);
    if (max_entries < 0 || max_entries > (1 << 20)) {
        blz::set_error("blz_set_residual_trace: max_entries must be in [0, 2^20]");
        return BLZ_E_INVALID;
    }
    Solver& s = h->s;
    BLZ_HIP(hipStreamSynchronize(s.stream));
    if (s.trace) BLZ_HIP(hipFree(s.trace));
    s.trace = nullptr;
    s.trace_cap = 0;
    if (max_entries > 0) {
        if (hipMalloc(&s.trace, (size_t)blz::MAX_SOLVES * max_entries * sizeof(blz_trace_entry)) != hipSuccess) {
            s.trace = nullptr;
            blz::set_error("hipMalloc failed for the residual history");
            return BLZ_E_NOMEM;
        }
        s.trace_cap = max_entries;
    }
    return BLZ_OK;
}

int blz_get_residual_trace(blz_solver* h, int solve, blz_trace_entry* out, int max, int* n) {
    CHECK_HANDLE(h);
    Solver& s = h->s;
    if (!n || max < 0 || (max > 0 && !out) || solve < 0 || solve >= s.last_nsolves) {
        blz::set_error("blz_get_residual_trace: bad arguments");
        return BLZ_E_INVALID;
    }
    if (!s.trace) {
        blz::set_error("blz_get_residual_trace: recording is off (blz_set_residual_trace)");
        return BLZ_E_INVALID;
    }
    // the control blocks of the last step were copied to the host with it
    const int have = std::min(s.ctl_host[solve].ntrace, s.trace_cap);
    *n = have;
    const int m = std::min(have, max);
    if (m > 0)
        BLZ_HIP(hipMemcpy(out, s.trace + (size_t)solve * s.trace_cap, (size_t)m * sizeof(blz_trace_entry),
                          hipMemcpyDeviceToHost));
    return BLZ_OK;
}

int blz_selftest_division(int divisor, unsigned long long* mismatches) {
    if (divisor < 3 || divisor > 6 || !mismatches) {
        blz::set_error("blz_selftest_division: divisor must be 3..6");
        return BLZ_E_INVALID;
    }
    hipStream_t st = nullptr;
    unsigned long long* d = nullptr;
    BLZ_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    int rc = BLZ_OK;
    if (hipMalloc(&d, sizeof(*d)) != hipSuccess) rc = BLZ_E_NOMEM;
    if (!rc && hipMemsetAsync(d, 0, sizeof(*d), st) != hipSuccess) rc = BLZ_E_HIP;
    if (!rc) rc = blz::selftest_division(st, divisor, d);
    if (!rc && hipMemcpyAsync(mismatches, d, sizeof(*d), hipMemcpyDeviceToHost, st) != hipSuccess) rc = BLZ_E_HIP;
    if (!rc && hipStreamSynchronize(st) != hipSuccess) rc = BLZ_E_HIP;
    if (d) (void)hipFree(d);
    (void)hipStreamDestroy(st);
    if (rc) blz::set_error("blz_selftest_division failed");
    return rc;
}

int blz_set_warm_start(blz_solver* h, blz_net* net) {
    CHECK_HANDLE(h);
    BLZ_HIP(hipStreamSynchronize(h->s.stream));
    return blz::warm_start_attach(h->s, net ? &net->n : nullptr);
}

int blz_stage_source(blz_solver* h, const float* part8, int count) {
    CHECK_HANDLE(h);
    return h->s.stage_source(part8, count);
}

int blz_add_source_staged(blz_solver* h) {
    CHECK_HANDLE(h);
    return h->s.add_source_staged();
}

int blz_render(blz_solver* h, const blz_render_settings* r, float* rgb) {
    CHECK_HANDLE(h);
    if (!r || !rgb) {
        blz::set_error("blz_render: null argument");
        return BLZ_E_INVALID;
    }
    return blz::render_frame(h->s, *r, rgb);
}

int blz_render_noalloc(blz_solver* h, const blz_render_settings* r, float* rgb) {
    CHECK_HANDLE(h);
    if (!r || !rgb) {
        blz::set_error("blz_render_noalloc: null argument");
        return BLZ_E_INVALID;
    }
    return blz::render_frame_noalloc(h->s, *r, rgb);
}

int blz_set_warm_start_parts(blz_solver* h, int parts) {
    CHECK_HANDLE(h);
    return blz::warm_start_set_parts(h->s, parts);
}

int blz_warm_start_input(blz_solver* h, float dt, float* in2, int net_dims[3], int coarse_dims[3]) {
    CHECK_HANDLE(h);
    Solver& s = h->s;
    if (!s.ws_net) {
        blz::set_error("no warm-start network attached (blz_set_warm_start)");
        return BLZ_E_INVALID;
    }
    if (net_dims) {
        net_dims[0] = s.Wn;
        net_dims[1] = s.Hn;
        net_dims[2] = s.Dn;
    }
    if (coarse_dims) {
        coarse_dims[0] = s.Wc;
        coarse_dims[1] = s.Hc;
        coarse_dims[2] = s.Dc;
    }
    if (!in2) return BLZ_OK;
    if (!(dt > 0.0f)) {
        blz::set_error("dt must be positive");
        return BLZ_E_INVALID;
    }
    return blz::warm_start_input(s, dt, in2);
}

int blz_warm_start_input_global(blz_solver* h, float dt, float* in2) {
    CHECK_HANDLE(h);
    Solver& s = h->s;
    if (!s.ws_net) {
        blz::set_error("no warm-start network attached (blz_set_warm_start)");
        return BLZ_E_INVALID;
    }
    if (!in2 || !(dt > 0.0f)) {
        blz::set_error("blz_warm_start_input_global: need an output array and dt > 0");
        return BLZ_E_INVALID;
    }
    return blz::warm_start_input_global(s, dt, in2);
}

int blz_autotune_rbgs(blz_solver* h, int sweeps, int* variant, float* ms_per_sweep) {
    CHECK_HANDLE(h);
    int best = 0;
    if (int e = h->s.autotune(sweeps, &best, ms_per_sweep)) return e;
    h->s.rbgs_variant = best;
    if (variant) *variant = best;
    return BLZ_OK;
}

int blz_warm_start_global(blz_solver* h, float dt) {
    CHECK_HANDLE(h);
    if (!(dt > 0.0f)) {
        blz::set_error("dt must be positive");
        return BLZ_E_INVALID;
    }
    if (!h->s.ws_net) {
        blz::set_error("no warm-start network attached (blz_set_warm_start)");
        return BLZ_E_INVALID;
    }
    return blz::warm_start_global(h->s, dt);
}

int blz_warm_start_apply(blz_solver* h, float dt, const float* pred, double* before, double* after, int* kept) {
    CHECK_HANDLE(h);
    if (!pred || !before || !after || !kept || !(dt > 0.0f)) {
        blz::set_error("blz_warm_start_apply: bad argument");
        return BLZ_E_INVALID;
    }
    return blz::warm_start_apply_given(h->s, dt, pred, before, after, kept);
}

int blz_set_rbgs_variant(blz_solver* h, int variant) {
    CHECK_HANDLE(h);
    if (!blz::rbgs_variant_valid(variant)) {
        blz::set_error("unknown RBGS variant");
        return BLZ_E_INVALID;
    }
    h->s.rbgs_auto = 0;
    h->s.rbgs_variant = variant;
    return BLZ_OK;
}

} // extern "C"
