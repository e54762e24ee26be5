// batch.hip — many independent scenes of uneven grids stepped together: one
// step of every scene is one replay of a captured HIP graph, and the
// pressure sweeps and residual checks, the pressure update, the advection
// and the warm start's network of all scenes run as single launches over
// tables of grids (k_rbgs_fused_multi, k_residual_multi, k_pupdate_multi,
// k_advect_multi, NetMulti), so small grids share the GPU instead of being
// launch-bound one by one. The sweep kernel's tile parameters (slab groups
// in flight, waves per block) come from an autotuner that times each
// candidate on the batch itself.
//
// No reference counterpart (the reference runs one scene per process,
// src/main.cu:72-117). Each scene computes exactly what blz_step computes
// for it alone with the same settings (the same per-cell code in the same
// stage order).
#include <algorithm>
#include <chrono>
#include <cstring>
#include <new>
#include <vector>

#include "solver.h"

namespace blz {

int launch_fused_multi(hipStream_t st, int variant, const FusedScene* tab, int n, int blocks);
int launch_residual_multi(hipStream_t st, const ResScene* tab, int n, int blocks, int add, int first);
int fused_waves(const Solver& s);
int pair_waves(const Solver& s);
constexpr int NTUNE = BLZ_BATCH_TUNE;
int residual_blocks(const Solver& s);

struct Batch {
    std::vector<blz_solver*> sc;
    hipStream_t root = nullptr;
    // the graphs of one step, per parity of steps, launched in order. When
    // the solve loop runs more than CHUNK_SWEEPS sweeps, its later batches
    // are cut into graphs of CHUNK_SWEEPS sweeps each and the pressure
    // update and advection go to a last graph of their own: once every
    // scene's solve has ended, the step skips the remaining sweep graphs, as
    // a single solve stops enqueuing sweeps (their launches would all return
    // at entry)
    std::vector<hipGraphExec_t> exec[2];
    std::vector<hipGraph_t> cap_graphs; // graphs of the capture in progress
    hipEvent_t chunk_ev[2] = {nullptr, nullptr};
    int graphs_run = 0; // graphs launched by the last step
    int next = 0; // parity of the next replay
    int field_cur[2][1024][F_N] = {};
    hipEvent_t fork = nullptr;
    std::vector<hipEvent_t> join;
    float* off_host = nullptr; // 2 per scene (pinned)
    float* off_dev = nullptr;
    FusedScene* fs = nullptr;  // 2n entries: home -> spare, then spare -> home
    ResScene* rs = nullptr;    // n entries
    // the warm start's first judge of every warm-started scene in multi-grid
    // launches: sweeps of the corrected (home) and uncorrected (judge
    // planes) paths, their residual partials, the decisions and restores
    FusedScene* fsw = nullptr; // 2n: home path pairs of the warm scenes
    FusedScene* fsj = nullptr; // 2n: judge-plane path pairs (planes 6, 7 <-> spare 4, 5)
    ResScene* rsw = nullptr;   // n: home planes, partials at red + 8
    ResScene* rsj = nullptr;   // n: judge planes, partials at red + 8 + RESIDUAL_BLOCKS
    JudgeScene* js = nullptr;  // n
    FusedScene* fsj2 = nullptr; // 2n: fsj gated on the predicted judge's rejection (skip flag)
    ResScene* rsj2 = nullptr;   // n: rsj gated the same way
    std::vector<int> warm_q;   // scenes of these tables (empty: per-scene judges)
    bool multi_reapply = false; // re-applications also as multi-grid launches (one common interval and cap)
    // the network forwards of the warm-started scenes as one launch sequence
    // on the root stream, between the two halves of each scene's preparation
    // (when every scene of warm_q runs the default order with one network)
    NetMulti netm;
    bool joint_fwd = false;
    // with joint_fwd, the corrections around the forward as multi-grid
    // launches too, for the scenes warm_start_multi_ok admits (ws_multi[q])
    WsMulti wsm;
    std::vector<char> ws_multi;
    // job tables of the stages run as multi-grid launches, per captured
    // graph (the fields' buffers alternate): filled on the host while
    // capturing, uploaded after
    char* jobs_dev[2] = {nullptr, nullptr};
    std::vector<char> jobs_host[2];
    size_t jobs_bytes = 0; // per graph
    size_t jobs_off = 0;   // fill level of the table set being captured
    int cap_which = 0;     // graph being captured
    // a table's place in the set of graph cap_which (nullptr when full)
    template <class J>
    const J* put_jobs(const std::vector<J>& v) {
        const size_t b = v.size() * sizeof(J);
        if (jobs_off + b > jobs_bytes) {
            set_error("batch job tables overflow");
            return nullptr;
        }
        if (b) std::memcpy(jobs_host[cap_which].data() + jobs_off, v.data(), b);
        const J* d = (const J*)(jobs_dev[cap_which] + jobs_off);
        jobs_off += (b + 255) / 256 * 256;
        return d;
    }
    // first blocks of a table of 256-thread jobs over n elements; the total
    template <class J>
    static int first_blocks(std::vector<J>& v) {
        int fb = 0;
        for (J& j : v) {
            j.first_block = fb;
            fb += (int)blocks_for((size_t)j.n, 256);
        }
        return fb;
    }
    int w_fs_blocks = 0, w_rs_blocks = 0, w_restore_blocks = 0;
    std::vector<FusedScene> fs_h;
    std::vector<ResScene> rs_h;
    int fs_blocks = 0, rs_blocks = 0;
    int every = 10;
    int variant = 103;
    int nodes = 0;
    double capture_ms = 0.0, step_ms = 0.0;
    double tune_ms[NTUNE] = {};
    const Net* net = nullptr;

    ~Batch();
    int init(const blz_settings* st, int n);
    int build_tables(int v);
    int build_judge_tables();
    int capture(int which);
    int cut(hipStream_t st);
    void drop_graphs();
    int enqueue_step(hipStream_t st);
    int joint_forward_and_finish(hipStream_t st);
    int enqueue_update_and_advect(hipStream_t st);
    int step(float time);
    int autotune(int sweeps, int* best);
};

// fused (one row per wave) and row-pair candidates; waves per block = the
// hundreds digit
static const int kVariants[NTUNE] = {102, 103, 202, 203, 402, 403, 1103, 1104, 1203, 1204};
static int wpb_of(int v) { return (v / 100) % 10; }
// sweeps per graph of a step's later sweeps (about 30 ms of sweeps of the
// 64-scene benchmark batch)
constexpr int CHUNK_SWEEPS = 160;

Batch::~Batch() {
    if (root) (void)hipStreamSynchronize(root);
    drop_graphs();
    if (fork) (void)hipEventDestroy(fork);
    for (hipEvent_t e : chunk_ev)
        if (e) (void)hipEventDestroy(e);
    for (hipEvent_t e : join)
        if (e) (void)hipEventDestroy(e);
    if (off_host) (void)hipHostFree(off_host);
    if (off_dev) (void)hipFree(off_dev);
    if (fs) (void)hipFree(fs);
    if (rs) (void)hipFree(rs);
    for (void* q : {(void*)fsw, (void*)fsj, (void*)rsw, (void*)rsj, (void*)js, (void*)fsj2, (void*)rsj2,
                    (void*)jobs_dev[0], (void*)jobs_dev[1]})
        if (q) (void)hipFree(q);
    for (blz_solver* h : sc) delete h;
    if (root) (void)hipStreamDestroy(root);
}

int Batch::init(const blz_settings* st, int n) {
    if (n < 1 || n > 1024) {
        set_error("batch size must be in [1, 1024]");
        return BLZ_E_INVALID;
    }
    every = st[0].residual_check_every > 0 ? st[0].residual_check_every : 10;
    for (int q = 0; q < n; ++q) {
        const blz_settings& s = st[q];
        const int ev = s.residual_check_every > 0 ? s.residual_check_every : 10;
        if (!(s.residual_tol > 0.0f) || s.cfl_number > 0.0f || ev != every || (every & 1) ||
            s.max_iter_solve % every != 0) {
            set_error("batched scenes need residual_tol > 0, cfl_number 0, one even residual_check_every "
                      "dividing max_iter_solve");
            return BLZ_E_INVALID;
        }
    }
    BLZ_HIP(hipStreamCreateWithFlags(&root, hipStreamNonBlocking));
    BLZ_HIP(hipEventCreateWithFlags(&fork, hipEventDisableTiming));
    for (hipEvent_t& e : chunk_ev) BLZ_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    BLZ_HIP(hipHostMalloc((void**)&off_host, 2 * n * sizeof(float), hipHostMallocDefault));
    if (hipMalloc(&off_dev, 2 * n * sizeof(float)) != hipSuccess || hipMalloc(&fs, 2 * n * sizeof(FusedScene)) != hipSuccess ||
        hipMalloc(&rs, n * sizeof(ResScene)) != hipSuccess || hipMalloc(&fsw, 2 * n * sizeof(FusedScene)) != hipSuccess ||
        hipMalloc(&fsj, 2 * n * sizeof(FusedScene)) != hipSuccess || hipMalloc(&rsw, n * sizeof(ResScene)) != hipSuccess ||
        hipMalloc(&rsj, n * sizeof(ResScene)) != hipSuccess || hipMalloc(&js, n * sizeof(JudgeScene)) != hipSuccess ||
        hipMalloc(&fsj2, 2 * n * sizeof(FusedScene)) != hipSuccess || hipMalloc(&rsj2, n * sizeof(ResScene)) != hipSuccess) {
        set_error("hipMalloc failed for batch tables");
        return BLZ_E_NOMEM;
    }
    // per scene at most 2 stage jobs before the solve, 3 pressure-update and
    // 5 advection jobs, and a control block reset, in up to 13 tables of
    // 256-byte aligned starts
    jobs_bytes = n * (sizeof(BuoyJob) + sizeof(DivJob) + sizeof(CtlJob) + 3 * sizeof(PupJob) + 5 * sizeof(AdvJob)) +
                 13 * 256;
    for (auto& d : jobs_dev)
        if (hipMalloc(&d, jobs_bytes) != hipSuccess) {
            set_error("hipMalloc failed for batch tables");
            return BLZ_E_NOMEM;
        }
    for (int q = 0; q < n; ++q) {
        blz_solver* h = new (std::nothrow) blz_solver();
        if (!h) {
            set_error("out of host memory");
            return BLZ_E_NOMEM;
        }
        sc.push_back(h);
        if (int e = h->s.init(st[q])) return e;
        if (!fused_available(h->s)) {
            set_error("batched scene too large for the fused sweep planes");
            return BLZ_E_INVALID;
        }
        hipEvent_t e;
        BLZ_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        join.push_back(e);
    }
    return build_tables(variant);
}

int Batch::build_tables(int v) {
    const int n = (int)sc.size();
    const int wpb = wpb_of(v);
    fs_h.assign(2 * n, FusedScene{});
    rs_h.assign(n, ResScene{});
    int fb = 0, rb = 0;
    for (int q = 0; q < n; ++q) {
        Solver& s = sc[q]->s;
        const unsigned pb = (unsigned)(s.Np * 4);
        const float dt = s.s.dt;
        const int nw = v >= 1000 ? pair_waves(s) : fused_waves(s);
        for (int h = 0; h < 2; ++h) {
            FusedScene& e = fs_h[h * n + q];
            const unsigned src = h == 0 ? 0u : 4u * pb, dst = h == 0 ? 4u * pb : 0u;
            e.base = s.prb;
            e.nbytes = (unsigned)(s.prb_floats * 4);
            e.srcR = src;
            e.srcB = src + pb;
            e.rhsR = 2u * pb;
            e.rhsB = 3u * pb;
            e.dstR = dst;
            e.dstB = dst + pb;
            e.W = s.W;
            e.H = s.H;
            e.D = s.D;
            e.Wp = s.L.Wp;
            e.Hp = s.L.Hp;
            e.scale = (s.s.density * s.dx * s.dx) / dt; // gs_solve_kernel :530
            e.stop = &s.ctl[0].done;
            e.first_block = fb;
            e.nwaves = nw;
        }
        fb += (nw + wpb - 1) / wpb;
        ResScene& r = rs_h[q];
        r.p = s.p;
        r.rhs = s.rhs;
        r.L = s.L;
        r.W = s.W;
        r.H = s.H;
        r.D = s.D;
        r.c = (double)dt / ((double)s.s.density * (double)s.dx * (double)s.dx);
        for (int a = 0; a < 6; ++a) r.cb[a] = s.s.closed_bounds[a];
        r.partial = s.red + 8;
        r.ctl = s.ctl;
        r.first_block = rb;
        r.nblocks = residual_blocks(s);
        r.tol = (double)s.s.residual_tol;
        r.cap = s.s.max_iter_solve;
        r.seq = 0;
        r.host_done = s.done_host_dev;
        r.gate = &s.ctl[0].done;
        rb += r.nblocks;
    }
    fs_blocks = fb;
    rs_blocks = rb;
    // the device table takes the second sweep of a pair (spare -> home) in
    // reverse scene order: the blocks of a launch start roughly in index
    // order, so each sweep begins with the scenes the previous one ended
    // with, whose planes are the likeliest to be still in the Infinity Cache
    // when the batch's planes exceed it (the scenes are independent, so
    // the order changes no value)
    std::vector<FusedScene> dev(fs_h);
    for (int r = 0, f1 = 0; r < n; ++r) {
        FusedScene e = fs_h[n + (n - 1 - r)];
        e.first_block = f1;
        f1 += (e.nwaves + wpb - 1) / wpb;
        dev[n + r] = e;
    }
    BLZ_HIP(hipMemcpy(fs, dev.data(), 2 * n * sizeof(FusedScene), hipMemcpyHostToDevice));
    BLZ_HIP(hipMemcpy(rs, rs_h.data(), n * sizeof(ResScene), hipMemcpyHostToDevice));
    // the scenes' own sweeps (the warm start's judge, on each scene's stream)
    // use the same variant as a single-grid launch: one launch per sweep
    for (int q = 0; q < n; ++q) sc[q]->s.rbgs_auto = v;
    variant = v;
    return BLZ_OK;
}

// Tables of the first warm-start judge of each step for the scenes that
// apply one (network attached, warm_start != 0, a sweep cap), in multi-grid
// launches: the same sweeps, residuals and decisions as warm_start_judge
// runs per scene. Needs judge planes of their own and a fused variant
// (else warm_q stays empty and each scene judges on its own stream).
int Batch::build_judge_tables() {
    const int n = (int)sc.size();
    warm_q.clear();
    bool ok = variant >= 100;
    for (int q = 0; q < n; ++q) {
        const Solver& s = sc[q]->s;
        if (s.ws_net && s.s.warm_start && s.s.max_iter_solve > 0) {
            warm_q.push_back(q);
            ok = ok && s.judge_set == 6;
        }
    }
    if (!ok) warm_q.clear();
    const int nw = (int)warm_q.size();
    multi_reapply = false;
    joint_fwd = false;
    if (!nw) return BLZ_OK;
    bool joint = true;
    std::vector<Net*> views;
    std::vector<const int*> stops;
    for (int q : warm_q) {
        const Solver& s = sc[q]->s;
        joint = joint && warm_start_splits(s) && s.ws_src == sc[warm_q[0]]->s.ws_src;
        views.push_back(s.ws_net);
        stops.push_back(&s.ctl[0].done);
    }
    ws_multi.assign(n, 0);
    if (joint) {
        if (int e = netm.build(views, stops)) return e;
        joint_fwd = true;
        std::vector<Solver*> grids;
        std::vector<float> dts;
        for (int q : warm_q) {
            Solver& s = sc[q]->s;
            if (!warm_start_multi_ok(s)) continue;
            ws_multi[q] = 1;
            grids.push_back(&s);
            dts.push_back(s.s.dt);
        }
        if (int e = wsm.build(grids, dts)) return e;
    }
    multi_reapply = true;
    for (int q : warm_q) {
        const Solver& s = sc[q]->s;
        const Solver& s0 = sc[warm_q[0]]->s;
        multi_reapply = multi_reapply && s.s.warm_start >= 2 && s.s.warm_start == s0.s.warm_start &&
                        s.s.max_iter_solve == s0.s.max_iter_solve;
    }
    const int wpb = wpb_of(variant);
    std::vector<FusedScene> hw(2 * nw), hj(2 * nw), hj2(2 * nw);
    std::vector<ResScene> rw(nw), rj(nw), rj2(nw);
    std::vector<JudgeScene> jh(nw);
    int fb = 0, rb = 0, cb = 0;
    for (int w = 0; w < nw; ++w) {
        const int q = warm_q[w];
        Solver& s = sc[q]->s;
        const unsigned pb = (unsigned)(s.Np * 4);
        for (int h = 0; h < 2; ++h) {
            FusedScene e = fs_h[h * n + q];
            e.first_block = fb;
            hw[h * nw + w] = e;
            // judge planes (6, 7) -> spare (4, 5), then back
            FusedScene j = e;
            const unsigned src = h == 0 ? 6u * pb : 4u * pb, dst = h == 0 ? 4u * pb : 6u * pb;
            j.srcR = src;
            j.srcB = src + pb;
            j.dstR = dst;
            j.dstB = dst + pb;
            hj[h * nw + w] = j;
            j.stop = &s.ctl[0].skip_spare;
            hj2[h * nw + w] = j;
        }
        fb += (fs_h[q].nwaves + wpb - 1) / wpb;
        ResScene a = rs_h[q];
        a.first_block = rb;
        rw[w] = a;
        ResScene b = a;
        b.p.c[0] = s.prb + 6 * s.Np;
        b.p.c[1] = s.prb + 7 * s.Np;
        b.partial = s.red + 8 + RESIDUAL_BLOCKS;
        rj[w] = b;
        b.gate = &s.ctl[0].skip_spare;
        rj2[w] = b;
        rb += a.nblocks;
        JudgeScene& J = jh[w];
        J.pa = s.red + 8;
        J.pb = s.red + 8 + RESIDUAL_BLOCKS;
        J.na = J.nb = a.nblocks;
        J.ctl = s.ctl;
        J.host_done = s.done_host_dev;
        J.tol = (double)s.s.residual_tol;
        J.cap = s.s.max_iter_solve;
        J.backup = s.prb + 6 * s.Np;
        J.p = s.prb;
        J.n = 2 * s.Np;
        J.first_block = cb;
        J.nblocks = (int)std::min<size_t>((J.n + 255) / 256, 1024);
        cb += J.nblocks;
    }
    w_fs_blocks = fb;
    w_rs_blocks = rb;
    w_restore_blocks = cb;
    BLZ_HIP(hipMemcpy(fsw, hw.data(), 2 * nw * sizeof(FusedScene), hipMemcpyHostToDevice));
    BLZ_HIP(hipMemcpy(fsj, hj.data(), 2 * nw * sizeof(FusedScene), hipMemcpyHostToDevice));
    BLZ_HIP(hipMemcpy(rsw, rw.data(), nw * sizeof(ResScene), hipMemcpyHostToDevice));
    BLZ_HIP(hipMemcpy(rsj, rj.data(), nw * sizeof(ResScene), hipMemcpyHostToDevice));
    BLZ_HIP(hipMemcpy(js, jh.data(), nw * sizeof(JudgeScene), hipMemcpyHostToDevice));
    BLZ_HIP(hipMemcpy(fsj2, hj2.data(), 2 * nw * sizeof(FusedScene), hipMemcpyHostToDevice));
    BLZ_HIP(hipMemcpy(rsj2, rj2.data(), nw * sizeof(ResScene), hipMemcpyHostToDevice));
    return BLZ_OK;
}

// the corrections of all warm-started scenes from the point where the
// scenes' streams joined into `st`: the first halves of the multi-grid
// scenes (the others packed their inputs on their own streams), the network
// forwards of all, the second halves of the multi-grid scenes, then those
// of the others on their streams, joined back
int Batch::joint_forward_and_finish(hipStream_t st) {
    if (int e = wsm.input(st)) return e;
    if (int e = netm.forward(st)) return e;
    if (int e = wsm.finish(st)) return e;
    BLZ_HIP(hipEventRecord(fork, st));
    for (int q : warm_q) {
        if (ws_multi[q]) continue;
        Solver& s = sc[q]->s;
        BLZ_HIP(hipStreamWaitEvent(s.stream, fork, 0));
        if (int e = warm_start_prepare_finish(s, s.ctl)) return e;
        BLZ_HIP(hipEventRecord(join[q], s.stream));
        BLZ_HIP(hipStreamWaitEvent(st, join[q], 0));
    }
    return BLZ_OK;
}

// one step of every scene: per-scene stages on the scenes' streams, the
// solves of all scenes on `st`
int Batch::enqueue_step(hipStream_t st) {
    const int n = (int)sc.size();
    jobs_host[cap_which].assign(jobs_bytes, 0);
    jobs_off = 0;
    BLZ_HIP(hipMemcpyAsync(off_dev, off_host, 2 * n * sizeof(float), hipMemcpyHostToDevice, st));
    // FluidSolver::step's stages before the solve (src/FluidSolver.cu:868-925),
    // each over the scenes that run it: buoyancy and divergence as
    // multi-grid launches, the others on the scenes' streams
    auto per_scene = [&](int stage, bool (*runs)(const blz_settings&)) -> int {
        bool any = false;
        for (int q = 0; q < n && !any; ++q) any = runs(sc[q]->s.s);
        if (!any) return BLZ_OK;
        BLZ_HIP(hipEventRecord(fork, st));
        for (int q = 0; q < n; ++q) {
            Solver& s = sc[q]->s;
            if (!runs(s.s)) continue;
            BLZ_HIP(hipStreamWaitEvent(s.stream, fork, 0));
            if (int e = s.run_stage(stage, 0.0f, s.s.dt)) return e;
            BLZ_HIP(hipEventRecord(join[q], s.stream));
            BLZ_HIP(hipStreamWaitEvent(st, join[q], 0));
        }
        return BLZ_OK;
    };
    if (int e = per_scene(BLZ_STAGE_COOLDOWN, [](const blz_settings& b) { return std::fabs(b.cooling_rate) > 0.0f; }))
        return e;
    if (int e = per_scene(BLZ_STAGE_DRAG, [](const blz_settings& b) { return b.drag > 0.0f; })) return e;
    if (int e = per_scene(BLZ_STAGE_VORTICITY, [](const blz_settings& b) { return std::fabs(b.vorticity_conf) > 0.0f; }))
        return e;
    {
        std::vector<BuoyJob> bj[2];
        for (int q = 0; q < n; ++q) {
            Solver& s = sc[q]->s;
            if (std::fabs(s.s.buoyancy) > 0.0f) buoyancy_jobs(s, s.s.dt, bj);
        }
        for (int m = 0; m < 2; ++m) {
            const int blocks = first_blocks(bj[m]);
            const BuoyJob* t = put_jobs(bj[m]);
            if (!t) return BLZ_E_INVALID;
            if (int e = launch_buoyancy_multi(st, m == 1, t, (int)bj[m].size(), blocks)) return e;
        }
    }
    if (int e = per_scene(BLZ_STAGE_WIND, [](const blz_settings& b) { return b.wind_amp > 0.0f; })) return e;
    if (int e = per_scene(BLZ_STAGE_TURBULENCE, [](const blz_settings& b) { return b.turbulence_amp > 0.0f; }))
        return e;
    {
        std::vector<DivJob> dj;
        std::vector<CtlJob> cj;
        for (int q = 0; q < n; ++q) {
            Solver& s = sc[q]->s;
            divergence_jobs(s, dj);
            // adaptive check spacing and the fixed-point stop as
            // Solver::project_solve sets them
            const bool warm = s.ws_net && s.s.warm_start && s.s.max_iter_solve > 0;
            ctl_reset_jobs(s, s.ctl, cj, s.s.residual_check_adaptive && !warm, s.s.residual_check_adaptive != 0);
        }
        const int blocks = first_blocks(dj);
        const DivJob* dt = put_jobs(dj);
        const CtlJob* ct = put_jobs(cj);
        if (!dt || !ct) return BLZ_E_INVALID;
        if (int e = launch_divergence_multi(st, dt, n, blocks)) return e;
        // the solve's control blocks, then its first residual check (as
        // Solver::project_solve's launch_ctl_reset and launch_residual_check)
        if (int e = launch_ctl_reset_multi(st, ct, n)) return e;
        if (int e = launch_residual_multi(st, rs, n, rs_blocks, 0, 1)) return e;
    }
    // the warm start's corrections: on the scenes' streams where they do not
    // run in multi-grid launches
    bool any_ws = false;
    for (int q = 0; q < n && !any_ws; ++q) {
        const Solver& s = sc[q]->s;
        any_ws = s.ws_net && s.s.warm_start && s.s.max_iter_solve > 0 && !(joint_fwd && ws_multi[q]);
    }
    if (any_ws) {
        BLZ_HIP(hipEventRecord(fork, st));
        for (int q = 0; q < n; ++q) {
            Solver& s = sc[q]->s;
            SolveCtl* c = s.ctl;
            const double tol = (double)s.s.residual_tol;
            const int cap = s.s.max_iter_solve;
            if (!(s.ws_net && s.s.warm_start && cap > 0) || (joint_fwd && ws_multi[q])) continue;
            BLZ_HIP(hipStreamWaitEvent(s.stream, fork, 0));
            if (joint_fwd) {
                // network forward and the rest below
                if (int e = warm_start_prepare_input(s, s.s.dt, c)) return e;
            } else if (!warm_q.empty()) {
                if (int e = warm_start_prepare(s, s.s.dt, c)) return e; // judged below, with the other scenes
            } else if (int e = warm_start_apply(s, s.s.dt, c, s.done_host_dev, 0, tol, cap, every)) {
                return e;
            }
            BLZ_HIP(hipEventRecord(join[q], s.stream));
            BLZ_HIP(hipStreamWaitEvent(st, join[q], 0));
        }
    }
    if (joint_fwd)
        if (int e = joint_forward_and_finish(st)) return e;
    if (!warm_q.empty()) {
        // the first judge of every warm-started scene: `every` sweeps on the
        // corrected pressure (home planes), the same on the uncorrected one
        // (judge planes), both residuals, the decisions and restores
        const int nw = (int)warm_q.size();
        for (int l = 0; l < every / 2; ++l) {
            if (int e = launch_fused_multi(st, variant, fsw, nw, w_fs_blocks)) return e;
            if (int e = launch_fused_multi(st, variant, fsw + nw, nw, w_fs_blocks)) return e;
        }
        for (int l = 0; l < every / 2; ++l) {
            if (int e = launch_fused_multi(st, variant, fsj, nw, w_fs_blocks)) return e;
            if (int e = launch_fused_multi(st, variant, fsj + nw, nw, w_fs_blocks)) return e;
        }
        if (int e = launch_residual_partials_multi(st, rsw, nw, w_rs_blocks)) return e;
        if (int e = launch_residual_partials_multi(st, rsj, nw, w_rs_blocks)) return e;
        if (int e = launch_ws_judge_multi(st, js, nw, w_restore_blocks, every)) return e;
    }
    // every scene's sweeps and checks; a converged scene's blocks return at
    // entry. it[q] follows scene q's sweep count as Solver::project_solve
    // counts it (the warm start's judge batch included), so re-applications
    // of the learned correction happen after the same sweeps as alone.
    std::vector<int> it(n);
    for (int q = 0; q < n; ++q) {
        const Solver& s = sc[q]->s;
        it[q] = (s.ws_net && s.s.warm_start && s.s.max_iter_solve > 0) ? every : 0;
    }
    auto pending = [&]() {
        for (int q = 0; q < n; ++q)
            if (it[q] < sc[q]->s.s.max_iter_solve) return true;
        return false;
    };
    // a new graph every chunk_iters batches (each whole: its forked streams
    // have joined the root stream again)
    const int chunk_iters = std::max(1, CHUNK_SWEEPS / every);
    int batches = 0, cuts = 0;
    while (pending()) {
        if (batches > 0 && batches % chunk_iters == 0) {
            if (int e = cut(st)) return e;
            ++cuts;
        }
        ++batches;
        for (int l = 0; l < every / 2; ++l) {
            if (int e = launch_fused_multi(st, variant, fs, n, fs_blocks)) return e;
            if (int e = launch_fused_multi(st, variant, fs + n, n, fs_blocks)) return e;
        }
        if (int e = launch_residual_multi(st, rs, n, rs_blocks, every, 0)) return e;
        // warm_start = N >= 2: each scene's learned correction every N sweeps
        std::vector<char> again(n, 0);
        bool any = false;
        for (int q = 0; q < n; ++q) {
            const Solver& s = sc[q]->s;
            if (it[q] >= s.s.max_iter_solve) continue;
            it[q] += every;
            again[q] = s.ws_net && s.s.warm_start >= 2 && it[q] % s.s.warm_start == 0 && it[q] < s.s.max_iter_solve;
            any |= again[q] != 0;
        }
        if (!any) continue;
        bool together = multi_reapply;
        for (int q = 0; q < n && together; ++q)
            together = (again[q] != 0) == (std::find(warm_q.begin(), warm_q.end(), q) != warm_q.end());
        if (together) {
            // every warm-started scene re-applies here (one common interval
            // and cap): the corrections per scene, then the predicted judges
            // of all of them as multi-grid launches, as warm_start_judge does
            // per scene
            const int nw = (int)warm_q.size();
            BLZ_HIP(hipEventRecord(fork, st));
            for (int q : warm_q) {
                Solver& s = sc[q]->s;
                BLZ_HIP(hipStreamWaitEvent(s.stream, fork, 0));
                if (joint_fwd) {
                    if (!ws_multi[q])
                        if (int e = warm_start_prepare_input(s, s.s.dt, s.ctl)) return e;
                } else if (int e = warm_start_prepare(s, s.s.dt, s.ctl)) {
                    return e;
                }
                BLZ_HIP(hipEventRecord(join[q], s.stream));
                BLZ_HIP(hipStreamWaitEvent(st, join[q], 0));
            }
            if (joint_fwd)
                if (int e = joint_forward_and_finish(st)) return e;
            for (int l = 0; l < every / 2; ++l) {
                if (int e = launch_fused_multi(st, variant, fsw, nw, w_fs_blocks)) return e;
                if (int e = launch_fused_multi(st, variant, fsw + nw, nw, w_fs_blocks)) return e;
            }
            if (int e = launch_residual_partials_multi(st, rsw, nw, w_rs_blocks)) return e;
            if (int e = launch_ws_judge_pred_multi(st, js, nw, every)) return e;
            for (int l = 0; l < every / 2; ++l) {
                if (int e = launch_fused_multi(st, variant, fsj2, nw, w_fs_blocks)) return e;
                if (int e = launch_fused_multi(st, variant, fsj2 + nw, nw, w_fs_blocks)) return e;
            }
            if (int e = launch_residual_partials_multi(st, rsj2, nw, w_rs_blocks)) return e;
            if (int e = launch_ws_judge_fin_multi(st, js, nw, w_restore_blocks, every)) return e;
            for (int q : warm_q) it[q] += every;
            continue;
        }
        BLZ_HIP(hipEventRecord(fork, st));
        for (int q = 0; q < n; ++q) {
            Solver& s = sc[q]->s;
            BLZ_HIP(hipStreamWaitEvent(s.stream, fork, 0));
            if (again[q]) {
                if (int e = warm_start_apply(s, s.s.dt, s.ctl, s.done_host_dev, 0, (double)s.s.residual_tol,
                                             s.s.max_iter_solve, every, true))
                    return e;
                it[q] += every;
            }
            BLZ_HIP(hipEventRecord(join[q], s.stream));
            BLZ_HIP(hipStreamWaitEvent(st, join[q], 0));
        }
    }
    if (cuts)
        if (int e = cut(st)) return e;
    if (int e = enqueue_update_and_advect(st)) return e;
    BLZ_HIP(hipEventRecord(fork, st));
    for (int q = 0; q < n; ++q) {
        Solver& s = sc[q]->s;
        BLZ_HIP(hipStreamWaitEvent(s.stream, fork, 0));
        BLZ_HIP(hipMemcpyAsync(s.ctl_host, s.ctl, sizeof(SolveCtl), hipMemcpyDeviceToHost, s.stream));
        BLZ_HIP(hipEventRecord(join[q], s.stream));
        BLZ_HIP(hipStreamWaitEvent(st, join[q], 0));
    }
    return BLZ_OK;
}

// The pressure update that ends project() (src/FluidSolver.cu:626-665) and
// the advection (:940-950) of every scene: the update's three face axes,
// then the advected fields,
// each as one launch over a job table of all scenes (the tables of graph
// cap_which, uploaded once the capture ends); the host's field swaps follow
// each scene's own launchers
int Batch::enqueue_update_and_advect(hipStream_t st) {
    const int n = (int)sc.size();
    std::vector<PupJob> pj[3];
    std::vector<AdvJob> aj[6];
    for (int q = 0; q < n; ++q) {
        Solver& s = sc[q]->s;
        pupdate_jobs(s, s.s.dt, pj);
        advect_jobs(s, s.s.dt, aj);
    }
    for (int a = 0; a < 3; ++a) {
        const int blocks = first_blocks(pj[a]);
        const PupJob* t = put_jobs(pj[a]);
        if (!t) return BLZ_E_INVALID;
        if (int e = launch_pupdate_multi(st, a, t, (int)pj[a].size(), blocks)) return e;
    }
    for (int k = 0; k < 6; ++k) {
        const int blocks = first_blocks(aj[k]);
        const AdvJob* t = put_jobs(aj[k]);
        if (!t) return BLZ_E_INVALID;
        if (int e = launch_advect_multi(st, k / 2, k & 1, t, (int)aj[k].size(), blocks)) return e;
    }
    return BLZ_OK;
}

// Capture graph `which`. The fields' front/back buffers alternate within a
// step, so the pointers baked into a capture hold for one parity of steps:
// graph 0 starts from the state at capture time, graph 1 from the state
// graph 0 leaves; replays alternate.
int Batch::capture(int which) {
    const int n = (int)sc.size();
    for (int q = 0; q < n; ++q) {
        BLZ_HIP(hipStreamSynchronize(sc[q]->s.stream));
        sc[q]->s.timing = false;
        for (int g = 0; g < F_N; ++g) field_cur[which][q][g] = sc[q]->s.f[g].cur;
    }
    auto t0 = std::chrono::steady_clock::now();
    if (int e = build_judge_tables()) return e;
    // captured noise kernels read their offsets from device memory; outside
    // the graph (blz_step on a scene handle) they take the time argument
    for (int q = 0; q < n; ++q) sc[q]->s.dev_off = off_dev + 2 * q;
    cap_which = which;
    cap_graphs.clear();
    auto drop_captured = [&]() {
        for (hipGraph_t g : cap_graphs) (void)hipGraphDestroy(g);
        cap_graphs.clear();
    };
    BLZ_HIP(hipStreamBeginCapture(root, hipStreamCaptureModeThreadLocal));
    int e = enqueue_step(root);
    hipGraph_t graph = nullptr;
    hipError_t ce = hipStreamEndCapture(root, &graph);
    if (graph) cap_graphs.push_back(graph);
    for (int q = 0; q < n; ++q) sc[q]->s.dev_off = nullptr;
    if (!e && ce != hipSuccess) {
        set_error(std::string("batch capture: ") + hipGetErrorString(ce));
        e = BLZ_E_HIP;
    }
    // the job tables the captured launches read
    if (!e && hipMemcpy(jobs_dev[which], jobs_host[which].data(), jobs_bytes, hipMemcpyHostToDevice) != hipSuccess) {
        set_error("batch job table upload failed");
        e = BLZ_E_HIP;
    }
    for (hipGraphExec_t x : exec[which]) (void)hipGraphExecDestroy(x);
    exec[which].clear();
    size_t total = 0;
    for (size_t k = 0; k < cap_graphs.size() && !e; ++k) {
        size_t nn = 0;
        hipGraphExec_t x = nullptr;
        hipError_t ge = hipGraphGetNodes(cap_graphs[k], nullptr, &nn);
        if (ge == hipSuccess) ge = hipGraphInstantiate(&x, cap_graphs[k], nullptr, nullptr, 0);
        if (ge != hipSuccess) {
            set_error(std::string("batch graph instantiation: ") + hipGetErrorString(ge));
            e = BLZ_E_HIP;
            break;
        }
        exec[which].push_back(x);
        total += nn;
    }
    drop_captured();
    if (e) {
        for (hipGraphExec_t x : exec[which]) (void)hipGraphExecDestroy(x);
        exec[which].clear();
        return e;
    }
    nodes = (int)total;
    capture_ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return BLZ_OK;
}

// End the capture in progress on `st` as one graph of the step and start
// capturing the next
int Batch::cut(hipStream_t st) {
    hipGraph_t g = nullptr;
    BLZ_HIP(hipStreamEndCapture(st, &g));
    cap_graphs.push_back(g);
    BLZ_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    return BLZ_OK;
}

void Batch::drop_graphs() {
    for (auto& v : exec) {
        for (hipGraphExec_t x : v) (void)hipGraphExecDestroy(x);
        v.clear();
    }
}

int Batch::step(float time) {
    const int n = (int)sc.size();
    if (exec[0].empty() || exec[1].empty()) {
        // capture both parities from the current state; the enqueue swapped
        // the host pointers once per captured step
        if (int e = capture(0)) return e;
        if (int e = capture(1)) return e;
        for (int q = 0; q < n; ++q)
            for (int g = 0; g < F_N; ++g) sc[q]->s.f[g].cur = field_cur[0][q][g];
        next = 0;
    }
    for (int q = 0; q < n; ++q) {
        Solver& s = sc[q]->s;
        noise_offsets(s, time, off_host + 2 * q);
        BLZ_HIP(hipStreamSynchronize(s.stream)); // sourcing or field access before the step
        clear_step_stats(s.st);
    }
    // a batched solve reports its end through its mapped flag with sequence
    // number 0
    for (int q = 0; q < n; ++q) *sc[q]->s.done_host = -1;
    auto all_done = [&]() {
        for (int q = 0; q < n; ++q)
            if (*sc[q]->s.done_host != 0) return false;
        return true;
    };
    auto t0 = std::chrono::steady_clock::now();
    const std::vector<hipGraphExec_t>& g = exec[next];
    const int m = (int)g.size();
    BLZ_HIP(hipGraphLaunch(g[0], root));
    int run = 1;
    if (m > 1) {
        // graphs 1 .. m-2 hold the later sweeps, graph m-1 the update and
        // advection; one sweep graph stays queued ahead of the host's check
        BLZ_HIP(hipEventRecord(chunk_ev[0], root));
        for (int i = 1; i < m - 1; ++i) {
            if (i >= 2) {
                BLZ_HIP(hipEventSynchronize(chunk_ev[i & 1])); // the end of graph i - 2
                if (all_done()) break;
            }
            BLZ_HIP(hipGraphLaunch(g[i], root));
            BLZ_HIP(hipEventRecord(chunk_ev[i & 1], root));
            ++run;
        }
        BLZ_HIP(hipGraphLaunch(g[m - 1], root));
        ++run;
    }
    BLZ_HIP(hipStreamSynchronize(root));
    graphs_run = run;
    step_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    next ^= 1;
    for (int q = 0; q < n; ++q) {
        Solver& s = sc[q]->s;
        for (int g = 0; g < F_N; ++g) s.f[g].cur = field_cur[next][q][g];
        s.solves.assign(1, s.s.max_iter_solve);
        s.spans.clear();
        s.ev_used = 0;
        s.st.substeps = 1;
        if (int e = s.finalize()) return e;
    }
    return BLZ_OK;
}

// time `sweeps` sweeps of every candidate on the batch (pressure set to 1.0,
// restored afterwards), keep the fastest
int Batch::autotune(int sweeps, int* best) {
    const int n = (int)sc.size();
    sweeps = std::max(2, sweeps & ~1);
    for (int q = 0; q < n; ++q) BLZ_HIP(hipStreamSynchronize(sc[q]->s.stream));
    // timing state as for a single grid (Solver::autotune): pressure 1.0,
    // so no numerator takes the exact division
    for (int q = 0; q < n; ++q)
        if (sc[q]->s.tmp_floats < 2 * sc[q]->s.Np) {
            set_error("batch autotune: scratch smaller than the pressure planes it saves");
            return BLZ_E_INVALID;
        }
    for (int q = 0; q < n; ++q) {
        Solver& s = sc[q]->s;
        BLZ_HIP(hipMemcpyAsync(s.tmp, s.prb, 2 * s.Np * sizeof(float), hipMemcpyDeviceToDevice, root));
        BLZ_HIP(hipMemsetAsync(s.ctl, 0, sizeof(SolveCtl), root));
        if (int e = launch_fill_pressure(s, 1.0f, root)) return e;
    }
    hipEvent_t e0, e1;
    BLZ_HIP(hipEventCreate(&e0));
    BLZ_HIP(hipEventCreate(&e1));
    int bestv = variant;
    double bestms = 1e30;
    int rc = BLZ_OK;
    for (int c = 0; c < NTUNE && !rc; ++c) {
        const int v = kVariants[c];
        if ((rc = build_tables(v))) break;
        // one untimed pair, then the timed sweeps
        if ((rc = launch_fused_multi(root, v, fs, n, fs_blocks)) || (rc = launch_fused_multi(root, v, fs + n, n, fs_blocks)))
            break;
        (void)hipEventRecord(e0, root);
        for (int l = 0; l < sweeps / 2 && !rc; ++l) {
            rc = launch_fused_multi(root, v, fs, n, fs_blocks);
            if (!rc) rc = launch_fused_multi(root, v, fs + n, n, fs_blocks);
        }
        (void)hipEventRecord(e1, root);
        if (hipEventSynchronize(e1) != hipSuccess) rc = BLZ_E_HIP;
        float ms = 0.0f;
        (void)hipEventElapsedTime(&ms, e0, e1);
        tune_ms[c] = ms / (float)sweeps;
        if (ms < bestms) {
            bestms = ms;
            bestv = v;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    for (int q = 0; q < n; ++q) {
        Solver& s = sc[q]->s;
        (void)hipMemcpyAsync(s.prb, s.tmp, 2 * s.Np * sizeof(float), hipMemcpyDeviceToDevice, root);
    }
    BLZ_HIP(hipStreamSynchronize(root));
    if (rc) return rc;
    if (int e = build_tables(bestv)) return e;
    drop_graphs();
    if (best) *best = bestv;
    return BLZ_OK;
}

} // namespace blz

using blz::Batch;

struct blz_batch {
    Batch b;
};

extern "C" {

int blz_batch_create(const blz_settings* settings, int n, blz_batch** out) {
    if (!settings || !out) {
        blz::set_error("blz_batch_create: null argument");
        return BLZ_E_INVALID;
    }
    *out = nullptr;
    blz_batch* h = new (std::nothrow) blz_batch();
    if (!h) {
        blz::set_error("out of host memory");
        return BLZ_E_NOMEM;
    }
    if (int e = h->b.init(settings, n)) {
        delete h;
        return e;
    }
    *out = h;
    return BLZ_OK;
}

int blz_batch_destroy(blz_batch* b) {
    delete b;
    return BLZ_OK;
}

int blz_batch_solver(blz_batch* b, int i, blz_solver** out) {
    if (!b || !out || i < 0 || i >= (int)b->b.sc.size()) {
        blz::set_error("blz_batch_solver: bad argument");
        return BLZ_E_INVALID;
    }
    *out = b->b.sc[i];
    return BLZ_OK;
}

int blz_batch_set_warm_start(blz_batch* b, blz_net* net) {
    if (!b) {
        blz::set_error("null batch");
        return BLZ_E_INVALID;
    }
    for (blz_solver* h : b->b.sc) {
        BLZ_HIP(hipStreamSynchronize(h->s.stream));
        if (int e = blz::warm_start_attach(h->s, net ? &net->n : nullptr)) return e;
    }
    b->b.drop_graphs();
    return BLZ_OK;
}

int blz_batch_autotune(blz_batch* b, int sweeps, int* variant) {
    if (!b) {
        blz::set_error("null batch");
        return BLZ_E_INVALID;
    }
    return b->b.autotune(sweeps, variant);
}

int blz_batch_set_variant(blz_batch* b, int variant) {
    if (!b) {
        blz::set_error("null batch");
        return BLZ_E_INVALID;
    }
    const int* end = blz::kVariants + blz::NTUNE;
    if (std::find(blz::kVariants, end, variant) == end) {
        blz::set_error("blz_batch_set_variant: not a batch sweep variant (see blz_batch_stats.tune_variant)");
        return BLZ_E_INVALID;
    }
    for (blz_solver* h : b->b.sc) BLZ_HIP(hipStreamSynchronize(h->s.stream));
    BLZ_HIP(hipStreamSynchronize(b->b.root));
    if (int e = b->b.build_tables(variant)) return e;
    b->b.drop_graphs();
    return BLZ_OK;
}

int blz_batch_step(blz_batch* b, float time) {
    if (!b) {
        blz::set_error("null batch");
        return BLZ_E_INVALID;
    }
    return b->b.step(time);
}

int blz_batch_get_stats(blz_batch* b, blz_batch_stats* out) {
    if (!b || !out) {
        blz::set_error("null argument");
        return BLZ_E_INVALID;
    }
    Batch& B = b->b;
    out->scenes = (int)B.sc.size();
    out->variant = B.variant;
    out->graph_nodes = B.nodes;
    out->graphs_run = B.graphs_run;
    out->capture_ms = B.capture_ms;
    out->step_ms = B.step_ms;
    long long vox = 0;
    for (blz_solver* h : B.sc) vox += (long long)h->s.W * h->s.H * h->s.D;
    out->voxels = vox;
    for (int c = 0; c < blz::NTUNE; ++c) {
        out->tune_variant[c] = blz::kVariants[c];
        out->tune_ms[c] = B.tune_ms[c];
    }
    return BLZ_OK;
}

} // extern "C"
