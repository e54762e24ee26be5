// Timer.h — frame counter and simulation time read by the solver, plus
// per-stage device times of the last frame.
//
// Reference: Timer (asiroisvigneux/Blaze src/Timer.h:32-132). The reference
// brackets stages with host clock(); here the times are HIP-event device
// times reported by the library (blz_stats).
#pragma once

#include <cstdio>
#include <string>

#include "../../include/blaze.h"

class Timer {
  public:
    int iter = 0;
    float time = 0.0f;
    blz_stats last{};
    double render_ms = 0.0; // last RenderEngine::render (wall, includes the image copy)
    double write_ms = 0.0;  // last RenderEngine::writeToDisk

    // the reference's per-frame table (src/Timer.h:80-132): seconds per
    // stage and its share of the frame, then the solver's own line. Stage
    // times here are device times (HIP events); scattering runs inside the
    // render call and is reported there.
    std::string getStats() const {
        const double src = last.source_ms * 1e-3;
        double st[BLZ_NUM_STAGES];
        for (int q = 0; q < BLZ_NUM_STAGES; ++q) st[q] = last.stage_ms[q] * 1e-3;
        const double rnd = render_ms * 1e-3, wr = write_ms * 1e-3, scat = 0.0;
        double total = src + rnd + wr + scat;
        for (int q = 0; q < BLZ_NUM_STAGES; ++q) total += st[q];
        const double den = total > 0.0 ? total : 1.0;
        std::string msg;
        char buf[256];
        std::snprintf(buf, sizeof(buf), "\nFrame: %03d\n", iter);
        msg += buf;
        auto line = [&](const char* name, double t) {
            std::snprintf(buf, sizeof(buf), "%-26s%4.6f s,   ratio: %1.2f\n", name, (float)t, (float)(t / den));
            msg += buf;
        };
        line("Sourcing:", src);
        line("Cooldown:", st[BLZ_STAGE_COOLDOWN]);
        line("Drag:", st[BLZ_STAGE_DRAG]);
        line("Buoyancy:", st[BLZ_STAGE_BUOYANCY]);
        line("Vorticity Confinement:", st[BLZ_STAGE_VORTICITY]);
        line("Wind:", st[BLZ_STAGE_WIND]);
        line("Turbulence:", st[BLZ_STAGE_TURBULENCE]);
        line("Compute Divergence:", st[BLZ_STAGE_DIVERGENCE]);
        line("Gauss-Seidel Solve:", st[BLZ_STAGE_PRESSURE]);
        line("Apply Pressure Gradient:", st[BLZ_STAGE_PROJECT_UPDATE]);
        line("Advection:", st[BLZ_STAGE_ADVECT]);
        line("Scattering:", scat);
        line("Render:", rnd);
        line("Write to Disk:", wr);
        std::snprintf(buf, sizeof(buf), "%-26s%4.6f s,   ratio: %1.2f\n", "Total Time:", (float)total, 1.0f);
        msg += buf;
        std::snprintf(buf, sizeof(buf), "Solver: substeps %d, RBGS sweeps %d, residual %.3e\n", last.substeps,
                      last.rbgs_iters, last.residual);
        msg += buf;
        return msg;
    }
};

