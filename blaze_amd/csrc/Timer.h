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

    std::string getStats() const {
        static const char* names[BLZ_NUM_STAGES] = {"cooldown", "drag", "vorticity", "buoyancy", "wind",
                                                    "turbulence", "divergence", "pressure", "update", "advect"};
        std::string msg;
        char buf[160];
        std::snprintf(buf, sizeof(buf), "frame %d  time %.4f  substeps %d  rbgs %d  residual %.3e\n", iter,
                      (double)time, last.substeps, last.rbgs_iters, last.residual);
        msg += buf;
        std::snprintf(buf, sizeof(buf), "  source %8.3f ms\n", last.source_ms);
        msg += buf;
        for (int s = 0; s < BLZ_NUM_STAGES; ++s) {
            if (last.stage_ms[s] <= 0.0) continue;
            std::snprintf(buf, sizeof(buf), "  %-10s %8.3f ms\n", names[s], last.stage_ms[s]);
            msg += buf;
        }
        std::snprintf(buf, sizeof(buf), "  step     %8.3f ms\n", last.step_ms);
        msg += buf;
        if (render_ms > 0.0) {
            std::snprintf(buf, sizeof(buf), "  render   %8.3f ms\n  write    %8.3f ms\n", render_ms, write_ms);
            msg += buf;
        }
        return msg;
    }
};
