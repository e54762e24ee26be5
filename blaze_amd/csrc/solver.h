// solver.h — internal state of one solver handle and the launchers of its
// kernels. Not part of the C ABI (include/blaze.h is).
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "common.h"
#include "dev_math.h"

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

// Pressure and right-hand side live in a red/black split layout: colour
// c = (i+j)&1, element (i>>1) + j*(W/2) + k*(W/2)*H of plane c.
struct RB {
    float* c[2] = {nullptr, nullptr};
};

enum { F_T = 0, F_U, F_V, F_W, F_S, F_N };

struct Solver {
    blz_settings s{};
    int W = 0, H = 0, D = 0;
    float dx = 0;
    bool smoke = false;
    hipStream_t stream = nullptr;
    Field f[F_N];
    RB p, rhs;
    float* tmp = nullptr;     // 7 * W*H*D scratch for cell-centred intermediates
    double* red = nullptr;    // reduction slots (device)
    double* red_host = nullptr; // pinned mirror
    // particles
    float* part = nullptr;
    int npart = 0, part_cap = 0;
    // particle-to-grid scratch (grow-only)
    void* p2g_buf = nullptr;
    size_t p2g_bytes = 0;
    // stats / timing
    blz_stats st{};
    bool timing = false;
    hipEvent_t ev[2 * (BLZ_NUM_STAGES + 2)] = {};
    int rbgs_variant = 0;

    ~Solver();
    int init(const blz_settings& st);
    int add_source(const float* part8, int n);
    int step_enqueue(float time, bool sync);
    int substep(float time, float dt);
    int run_stage(int stage, float time, float dt);
    int project_solve(float dt);
    int residual(float dt, double* out);
    int max_abs_velocity(float* out);
};

// ---- launchers (each enqueues on s.stream; return BLZ_OK or error) ----
int launch_cooldown(Solver& s, float dt);
int launch_drag(Solver& s, float dt);
int launch_buoyancy(Solver& s, float dt);
int launch_vorticity(Solver& s, float dt);
int launch_wind(Solver& s, float time, float dt);
int launch_turbulence(Solver& s, float time, float dt);
int launch_divergence(Solver& s);
int launch_rbgs_sweeps(Solver& s, float dt, int sweeps);
int launch_residual(Solver& s, float dt, double* dev_out);
int launch_pressure_update(Solver& s, float dt);
int launch_advect(Solver& s, float dt);
int launch_p2g(Solver& s);
int launch_max_abs_velocity(Solver& s, double* dev_out);
int launch_clear_doubles(Solver& s, double* dev, int n);
int launch_p_to_natural(Solver& s, float* out, const RB& src);
int launch_natural_to_rb(Solver& s, RB& dst, const float* in);

inline unsigned blocks_for(size_t n, unsigned bs) { return (unsigned)((n + bs - 1) / bs); }

} // namespace blz
