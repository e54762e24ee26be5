// solver.h — internal state of one solver handle and the launchers of its
// kernels. Not part of the C ABI (include/blaze.h is).
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "common.h"
#include "dev_math.h"
#include "unet.h"

namespace blz {

#define BLZ_HIP(expr)                                                                  \
    do {                                                                               \
        hipError_t e_ = (expr);                                                        \
        if (e_ != hipSuccess) {                                                        \
            ::blz::set_error(std::string(#expr " failed: ") + hipGetErrorString(e_)); \
            return BLZ_E_HIP;                                                          \
        }                                                                              \
    } while (0)

// One staggered field: front/back device buffers plus its geometry.
struct Field {
    float* buf[2] = {nullptr, nullptr};
    int cur = 0; // buf[cur] is the front
    int nx = 0, ny = 0, nz = 0;
    float ox = 0, oy = 0, oz = 0;
    bool flip = false;
    float* front() const { return buf[cur]; }
    float* back() const { return buf[cur ^ 1]; }
    void swap() { cur ^= 1; }
    size_t count() const { return (size_t)nx * ny * nz; }
    Gv view() const { return Gv{front(), nx, ny, nz, ox, oy, oz}; }
    Gv view_back() const { return Gv{back(), nx, ny, nz, ox, oy, oz}; }
};

// Pressure and right-hand side live in a red/black split, padded layout
// that interleaves four consecutive z-slabs per column: colour c = (i+j)&1,
// m = i>>1, slab group g = k>>2; plane c stores float4 elements indexed
// ((g+1)*Hp + (j+1))*Wp + (m+1) (Wp = W/2+2, Hp = H+2, Dg = ceil(D/4) groups
// plus one ghost group on each side), component k&3. A wave reading one
// group of 64 neighbouring columns reads 1 KiB contiguous, and one 16-byte
// load per lane covers four Gauss-Seidel steps. The ghost shell (m, j or g
// outside the grid, and slabs k >= D of the last group) of the pressure
// planes holds -0.0f, the additive identity of IEEE arithmetic
// (x + -0 == x for every x), so an update adds all neighbour slots
// unconditionally.
struct RBL {
    int Wh, H, D, Wp, Hp, Dg;
    __host__ __device__ int at(int i, int j, int k) const {
        return ((((k >> 2) + 1) * Hp + (j + 1)) * Wp + ((i >> 1) + 1)) * 4 + (k & 3);
    }
};
struct RB {
    float* c[2] = {nullptr, nullptr};
};

enum { F_T = 0, F_U, F_V, F_W, F_S, F_N };

// Per-solve control block of the residual-driven Gauss-Seidel loop, kept on
// the device: the residual kernel updates it and the sweep and residual
// kernels of the same solve do nothing once `done` is set, so the host
// enqueues sweeps without waiting for residual values.
struct SolveCtl {
    int done;   // 1 once the residual <= tol or the sweep cap is reached
    int iters;  // sweeps counted so far
    int checks; // residual evaluations so far
    int ws;     // last learned correction: 0 none, 1 kept, 2 rejected (pressure restored)
    double res, res0; // last and first residual
    double res_ws;    // residual right after the last learned correction
    int ws_kept, ws_rejected; // learned corrections kept / rejected in this solve
    double res_prev;          // residual of the check before the last one (-1: none)
    int skip_spare;           // 1: the predicted judge kept the correction (no uncorrected sweeps)
    int ntrace;               // entries written to trace
    blz_trace_entry* trace;   // residual history of this solve (nullptr: not recorded)
    int trace_cap;            // capacity of trace
    int swept;                // sweeps executed, including the judge's batches on the path not kept
    // adaptive check spacing (residual_check_adaptive; the oracle's rbgs()
    // follows the same rule): a batch's check is evaluated only when the
    // sweep count reaches next_check (or the cap); `gap` is the interval in
    // use, a multiple of `every`
    int adaptive, every, gap, next_check;
    int stop_stall;           // end the solve at an FP32 fixed point
    int stall_it;             // sweep count at which the current residual value was first seen
};
constexpr int STALL_SWEEPS = 60;   // the same residual at every check over this many sweeps: fixed point
constexpr int GAP_MAX_CHECKS = 64; // longest check interval, in residual_check_every units

// interval to the next check of an adaptive solve from the contraction over
// the last one (restated in oracle/blaze_oracle.hpp, Solver::next_gap)
__host__ __device__ inline int next_check_gap(int gap, int every, double r, double prev, double tol) {
    const double rho = r / prev;
    if (rho < 1.0 && r * rho <= tol) return max(every, gap / 4 / every * every);
    if (rho < 1.0 && r * rho * rho <= tol) return max(every, gap / 2 / every * every);
    if (rho < 1.0 && r * rho * rho * rho * rho <= tol) return gap;
    return min(2 * gap, GAP_MAX_CHECKS * every);
}

// one residual evaluation of a solve, appended to its history (when enabled)
__device__ __forceinline__ void trace_push(SolveCtl* c, int kind, double res) {
    if (c->trace && c->ntrace < c->trace_cap) {
        blz_trace_entry& t = c->trace[c->ntrace];
        t.iters = c->iters;
        t.kind = kind;
        t.res = res;
        c->ntrace += 1;
    }
}
constexpr int WS_GMAX = 4096;      // cells of the warm start's global-correction grid (16^3)
constexpr int MAX_SOLVES = 64;     // pressure solves per step (bounds max_substeps)
constexpr int SOLVE_LOOKAHEAD = 2; // sweep batches the host may enqueue past the last one it has seen finish

constexpr int RESIDUAL_BLOCKS = 1024;

// one grid of a multi-grid fused sweep launch (byte offsets into base)
struct FusedScene {
    float* base;
    unsigned nbytes, srcR, srcB, rhsR, rhsB, dstR, dstB;
    int W, H, D, Wp, Hp;
    float scale;
    const int* stop;
    int first_block, nwaves;
};

// one grid of a multi-grid residual launch
struct ResScene {
    RB p, rhs;
    RBL L;
    int W, H, D;
    double c;
    int cb[6];
    double* partial;
    SolveCtl* ctl;
    int first_block, nblocks;
    double tol;
    int cap, seq;
    int* host_done;
    const int* gate; // the launch skips this grid while *gate != 0 (the solve's done flag, or a judge's skip flag)
};

// one scene of a batch's warm-start judge launch: the FP64 residual partials
// of the corrected (pa) and uncorrected (pb) paths, the solve's control
// block, and the restore copy (judge planes -> home planes, n floats) with
// its blocks of the restore launch
struct JudgeScene {
    const double* pa;
    const double* pb;
    int na, nb;
    SolveCtl* ctl;
    int* host_done;
    double tol;
    int cap;
    const float* backup;
    float* p;
    size_t n;
    int first_block, nblocks;
};

struct Solver {
    blz_settings s{};
    int device = 0;           // HIP device current at creation (owns the stream and all buffers)
    int W = 0, H = 0, D = 0;
    float dx = 0;
    bool smoke = false;
    hipStream_t stream = nullptr;
    Field f[F_N];
    RB p, rhs;              // planes of one allocation: p red, p black, rhs red, rhs black,
                            // spare p red, spare p black (fused sweeps), judge p red, judge p
                            // black (the warm start's uncorrected copy), then a tail of slack
    float* prb = nullptr;
    size_t prb_floats = 0;  // floats in prb
    int judge_set = 4;      // first plane of the warm start's uncorrected copy: 6 when the judge
                            // planes exist (the judge's sweeps may then be fused), else 4
    RBL L{};
    size_t Np = 0;          // floats per padded plane
    float* tmp = nullptr;     // scratch for cell-centred intermediates (7 * W*H*D floats) and the
                              // autotuners' saved pressure planes (2 * Np floats): max of the two
    size_t tmp_floats = 0;
    double* red = nullptr;    // reduction slots (device): 8 results + RESIDUAL_BLOCKS partials
    double* red_host = nullptr; // pinned mirror
    SolveCtl* ctl = nullptr;      // MAX_SOLVES control blocks (device)
    SolveCtl* ctl_host = nullptr; // pinned copy, filled at the end of a step
    blz_trace_entry* trace = nullptr; // residual histories: MAX_SOLVES x trace_cap entries (device)
    int trace_cap = 0;                // 0: histories off
    int last_nsolves = 0;             // solves of the last finalized step (their ctl_host entries are valid)
    int* zero_flag = nullptr;     // device int that stays 0: stop flag of fixed-count solves
    volatile int* done_host = nullptr; // mapped pinned int: sequence number of the last converged solve
    int* done_host_dev = nullptr;      // its device address
    int solve_seq = 0;
    std::vector<int> solves; // solves enqueued since the last finalize: sweep cap, or -1 (fixed count)
    hipEvent_t batch_ev[SOLVE_LOOKAHEAD + 1] = {};
    // particles
    float* part = nullptr;
    int npart = 0, part_cap = 0;
    // staged particle uploads (double-buffered pinned host memory, a copy
    // stream at default priority, one event per slot)
    float* stage_host[2] = {nullptr, nullptr};
    float* stage_dev[2] = {nullptr, nullptr};
    int stage_cap[2] = {0, 0};
    int stage_count[2] = {0, 0};
    hipEvent_t stage_ev[2] = {nullptr, nullptr};
    hipStream_t copy_stream = nullptr;
    int stage_next = 0;   // slot the next stage_source fills
    int stage_ready = -1; // slot staged last (-1: none)
    // particle-to-grid scratch (grow-only)
    void* p2g_buf = nullptr;
    size_t p2g_bytes = 0;
    size_t p2g_bound = 0;              // upper bound of the splat's entries (host, from the particle set)
    size_t stage_bound[2] = {0, 0};    // the same for each staged slot
    const unsigned* p2g_overflow = nullptr; // device flag of the last splat: its entries exceeded the bound
    // stats / timing: spans of device work bracketed by pooled events,
    // resolved after the stream has been synchronised (no extra waits)
    blz_stats st{};
    bool timing = false;
    enum { SPAN_SOURCE = BLZ_NUM_STAGES, SPAN_STEP, SPAN_RBGS };
    struct Span {
        int kind, e0, e1, launches;
        int solve, batch, sweeps; // RBGS batches: owning solve, batch index, sweeps
    };
    std::vector<hipEvent_t> evpool;
    size_t ev_used = 0;
    std::vector<Span> spans;
    int rbgs_variant = 0; // requested variant (0 = autotuned)
    int rbgs_auto = 0;    // autotuned variant when rbgs_variant == 0
    // learned warm start (attached network; activations of its own)
    const Net* ws_src = nullptr;
    Net* ws_net = nullptr;     // view: shares ws_src's weights, owns activations
    int ws_f = 4;               // coarsening factor of the warm start
    int ws_parts = 3;           // bit 0: network correction, bit 1: global correction
    int ws_g = 1;               // block size of the global correction's grid
    int Gw = 1, Gh = 1, Gd = 1; // that grid (at most WS_GMAX cells)
    int Wc = 0, Hc = 0, Dc = 0; // coarse grid, real cells
    int Wn = 0, Hn = 0, Dn = 0; // network grid (coarse rounded up to multiples of 4)
    float* ws_buf = nullptr;   // coarse residual, mask (Wn*Hn*Dn each), scale bits
    float* rnd_buf = nullptr;  // renderer: two scatter volumes and the frame buffer
    size_t rnd_bytes = 0;
    // captured (batched) steps: noise offsets of wind and turbulence read
    // from device memory instead of kernel arguments
    const float* dev_off = nullptr;

    int span_begin(int* e0);
    int span_end(int kind, int e0, int launches, int solve = -1, int batch = -1, int sweeps = 0);
    int resolve_spans();
    int finalize(); // after a stream synchronisation: solve statistics and spans

    ~Solver();
    int init(const blz_settings& st);
    int add_source(const float* part8, int n);
    int stage_source(const float* part8, int n);
    int add_source_staged();
    int step_enqueue(float time, bool sync);
    int substep(float time, float dt);
    int run_stage(int stage, float time, float dt);
    int project_solve(float dt);
    int autotune(int sweeps, int* best, float* ms_per_sweep);
    int enqueue_solve_stats();
    int max_abs_velocity(float* out);
};

// ---- launchers (each enqueues on s.stream; return BLZ_OK or error) ----
int launch_cooldown(Solver& s, float dt);
int launch_drag(Solver& s, float dt);
int launch_buoyancy(Solver& s, float dt);
int launch_vorticity(Solver& s, float dt);
int launch_wind(Solver& s, float time, float dt);
void noise_offsets(const Solver& s, float time, float out[2]);
int launch_turbulence(Solver& s, float time, float dt);
int launch_divergence(Solver& s);
// enqueues `sweeps` sweeps; *launches = kernels launched, *kind = 1 (two
// colour passes per sweep) or 2 (fused sweeps)
int launch_rbgs_sweeps(Solver& s, float dt, int sweeps, const int* stop, int* launches, int* kind);
bool fused_available(const Solver& s);
bool rbgs_variant_valid(int v);
void clear_step_stats(blz_stats& st);
int launch_residual_check(Solver& s, float dt, SolveCtl* c, int* host_done, int seq, double tol, int cap, int add,
                          int first);
// adaptive: check spacing from the contraction; stop_stall: end at an FP32 fixed point
int launch_ctl_reset(Solver& s, SolveCtl* c, bool adaptive = false, bool stop_stall = false);
int launch_residual_ws(Solver& s, float dt, SolveCtl* c, int* host_done, int seq, double tol, int cap);
int render_frame(Solver& s, const blz_render_settings& r, float* host_rgb);
// the same with the library's static per-device scratch (no allocation)
int render_frame_noalloc(Solver& s, const blz_render_settings& r, float* host_rgb);
int warm_start_set_parts(Solver& s, int parts);
int launch_residual_ws_final(Solver& s, SolveCtl* c, int* host_done, int seq, double tol, int cap, int nb_partial);
int launch_rbgs_spare(Solver& s, float dt, int sweeps, const int* stop);
int launch_residual_partial(Solver& s, float dt, bool spare, double* partial, const int* stop, int* nb_out);
int residual_blocks(const Solver& s);
int warm_start_attach(Solver& s, const Net* net);
// learned correction + `judge` sweeps on both paths (counted once)
int warm_start_apply(Solver& s, float dt, SolveCtl* c, int* host_done, int seq, double tol, int cap, int judge,
                     bool predicted = false);
int warm_start_input(Solver& s, float dt, float* host_in2);
int warm_start_input_global(Solver& s, float dt, float* host_in2);
// warm_start_apply in two parts: the corrections, then the judge's sweeps,
// residuals and decision (a batch runs the second part for all scenes at once)
int warm_start_prepare(Solver& s, float dt, SolveCtl* c);
// warm_start_prepare of the default order (warm_start_splits) in two parts
// around the network's forward: up to its packed input, and from its output
// (a batch runs the forwards of all scenes together in between)
bool warm_start_splits(const Solver& s);
int warm_start_prepare_input(Solver& s, float dt, SolveCtl* c);
int warm_start_prepare_finish(Solver& s, SolveCtl* c);
// The same two halves for several grids at once (a batch), each kernel one
// launch over a job table of the grids, for the grids warm_start_multi_ok
// admits; every grid gets what its own launches give it. The tables hold
// the grids' buffers: rebuild after a grid's warm-start state changes.
constexpr int WS_MULTI_LAUNCHES = 11;
bool warm_start_multi_ok(const Solver& s);
struct WsMulti {
    void* tab = nullptr;
    size_t bytes = 0;
    std::vector<unsigned char> cur; // the table's content
    int n = 0;
    unsigned blocks[WS_MULTI_LAUNCHES] = {};
    ~WsMulti();
    int build(const std::vector<Solver*>& grids, const std::vector<float>& dts);
    int input(hipStream_t st) const;  // warm_start_prepare_input of every grid
    int finish(hipStream_t st) const; // warm_start_prepare_finish of every grid
};
int warm_start_judge(Solver& s, float dt, SolveCtl* c, int* host_done, int seq, double tol, int cap, int judge,
                     bool predicted);
// a batch's judges: decisions (one block per scene) and restores
int launch_ws_judge_multi(hipStream_t st, const JudgeScene* js, int n, int restore_blocks, int add);
// a batch's predicted judges (re-applications), in three launches: the
// decision on the corrected path, then (after the uncorrected path's
// sweeps and residuals, gated on the decision) its residual, and restores
int launch_ws_judge_pred_multi(hipStream_t st, const JudgeScene* js, int n, int add);
int launch_ws_judge_fin_multi(hipStream_t st, const JudgeScene* js, int n, int restore_blocks, int add);
// residual partials of every grid of a table (no control-block update)
int launch_residual_partials_multi(hipStream_t st, const ResScene* tab, int n, int blocks);
int warm_start_global(Solver& s, float dt); // the global correction alone, unjudged (training data)
int warm_start_apply_given(Solver& s, float dt, const float* host_pred, double* before, double* after, int* kept);
int launch_pressure_update(Solver& s, float dt);
int launch_advect(Solver& s, float dt);

// The pressure update and the advection of many grids as multi-grid
// launches (a batch's step): one job per (grid, face axis) or (grid,
// advected field or pair), blocks mapped to jobs through ascending first
// blocks. Each cell gets what the single-grid launch gives it.
struct PupJob {
    const float* in;
    float* out;
    RB p;
    RBL L;
    int W, H, D;
    float sc;
    int closedLo, closedHi;
    int first_block, n; // n faces
};
enum { ADV_FREE = 0, ADV_CLAMPED = 1, ADV_PAIR = 2 }; // velocity component; T or S alone; T with S
struct AdvJob {
    Gv U, V, W; // front velocities
    Gv G1, G2;  // advected field (G2: the pair's second field)
    float* out1;
    float* out2;
    float dx, s; // cell size; the scale's parameter (dx, or 1/dx when dx is a power of two)
    float a, b, dt;
    int first_block, n; // n cells
};
// append grid s's jobs (pupdate: jobs[axis]; advection: jobs[2 kind + (1/dx
// exact)]) and make the field swaps the single-grid launchers make
void pupdate_jobs(Solver& s, float dt, std::vector<PupJob>* jobs);
void advect_jobs(Solver& s, float dt, std::vector<AdvJob>* jobs);
int launch_pupdate_multi(hipStream_t st, int axis, const PupJob* tab, int n, int blocks);
int launch_advect_multi(hipStream_t st, int kind, bool mul, const AdvJob* tab, int n, int blocks);
// the same for the stages before the solve: buoyancy (jobs[1/dx exact]),
// divergence, and the control block's reset (one thread per job)
struct BuoyJob {
    Gv T;
    float* V;
    int vnx, vny, vnz;
    float vox, voy, voz;
    float s, dx, beta, g, dt; // s: the scale's parameter, as AdvJob's
    int first_block, n;
};
struct DivJob {
    const float *U, *V, *Wf;
    RB rhs;
    RBL L;
    int W, H, D;
    float sc;
    int cb[6]; // closed_bounds
    int first_block, n;
};
struct CtlJob {
    SolveCtl* c;
    blz_trace_entry* trace;
    int trace_cap, adaptive, stop_stall, every;
};
void buoyancy_jobs(Solver& s, float dt, std::vector<BuoyJob>* jobs);
void divergence_jobs(Solver& s, std::vector<DivJob>& jobs);
void ctl_reset_jobs(Solver& s, SolveCtl* c, std::vector<CtlJob>& jobs, bool adaptive = false,
                    bool stop_stall = false);
int launch_buoyancy_multi(hipStream_t st, bool mul, const BuoyJob* tab, int n, int blocks);
int launch_divergence_multi(hipStream_t st, const DivJob* tab, int n, int blocks);
int launch_ctl_reset_multi(hipStream_t st, const CtlJob* tab, int n);
int launch_p2g(Solver& s);
size_t p2g_entry_bound(const float* part8, int n, float dx, int nch);
int launch_max_abs_velocity(Solver& s, double* dev_out);
int launch_clear_doubles(Solver& s, double* dev, int n);
int launch_p_to_natural(Solver& s, float* out, const RB& src);
int launch_natural_to_rb(Solver& s, RB& dst, const float* in);
int launch_init_pressure(Solver& s);
int launch_fill_pressure(Solver& s, float v, hipStream_t st = nullptr); // on st, else the solver's stream
int selftest_division(hipStream_t st, int d, unsigned long long* dev_bad);

inline unsigned blocks_for(size_t n, unsigned bs) { return (unsigned)((n + bs - 1) / bs); }

} // namespace blz

struct blz_solver {
    blz::Solver s;
};
