// SceneSettings.h — scene configuration of the C++ drop-in API.
//
// Mirrors the member names of the reference struct (asiroisvigneux/Blaze
// src/SceneSettings.h:32-71) for the solver and source fields, with plain
// host vector types in place of CUDA's; the renderer's keys are kept as the
// C-ABI blz_render_settings.
// toC() yields the C-ABI blz_settings consumed by the HIP library.
#pragma once

#include "../../include/blaze.h"

namespace blz {
struct int2v { int x, y; };
struct int3v { int x, y, z; };
struct float2v { float x, y; };
struct float3v { float x, y, z; };
} // namespace blz

struct SceneSettings {
    char sceneDir[256];
    char sceneName[256];

    char particleFile[256];
    bool sourceAnimated;
    blz::int2v sourceRange;
    int sourceMaxParticleCount;

    char renderFile[256];
    blz::int2v renderRange;

    blz::float3v domainBboxMin;
    blz::float3v domainBboxMax;
    bool closedBounds[6];

    int maxIterSolve;
    float dx;
    float dt;

    blz::int3v gridRes;
    blz::int2v renderRes;

    float density;
    float gravity;
    float buoyancy;
    float coolingRate;
    float vorticityConf;
    float drag;

    int windDir; // 0:X+, 1:X-, 2:Z+, 3:Z-
    float windAmp;
    float windSpeed;
    float windTurbAmp;
    float windTurbScale;

    float turbulence_amp;
    float turbulence_scale;
    blz::float2v turbMaskTempRamp;
    blz::float2v turbMaskVelRamp;

    // extensions (defaults reproduce the reference)
    int noiseSeed;
    float residualTol;
    int residualCheckEvery;
    float cflNumber;
    int maxSubsteps;
    float smokeDensity;
    int warmStart;
    int residualCheckAdaptive;
    bool hasParticleGen;
    blz_particle_gen particleGen;
    // renderer keys (src/SceneSettings.h:73-107), parsed when present
    bool hasRender;
    blz_render_settings render;

    blz_settings toC() const {
        blz_settings s;
        blz_settings_default(&s);
        s.grid_res[0] = gridRes.x; s.grid_res[1] = gridRes.y; s.grid_res[2] = gridRes.z;
        s.dx = dx; s.dt = dt;
        s.domain_min[0] = domainBboxMin.x; s.domain_min[1] = domainBboxMin.y; s.domain_min[2] = domainBboxMin.z;
        s.domain_max[0] = domainBboxMax.x; s.domain_max[1] = domainBboxMax.y; s.domain_max[2] = domainBboxMax.z;
        for (int a = 0; a < 6; ++a) s.closed_bounds[a] = closedBounds[a] ? 1 : 0;
        s.max_iter_solve = maxIterSolve;
        s.density = density; s.gravity = gravity; s.buoyancy = buoyancy;
        s.cooling_rate = coolingRate; s.vorticity_conf = vorticityConf; s.drag = drag;
        s.wind_dir = windDir; s.wind_amp = windAmp; s.wind_speed = windSpeed;
        s.wind_turb_amp = windTurbAmp; s.wind_turb_scale = windTurbScale;
        s.turbulence_amp = turbulence_amp; s.turbulence_scale = turbulence_scale;
        s.turb_mask_temp_ramp[0] = turbMaskTempRamp.x; s.turb_mask_temp_ramp[1] = turbMaskTempRamp.y;
        s.turb_mask_vel_ramp[0] = turbMaskVelRamp.x; s.turb_mask_vel_ramp[1] = turbMaskVelRamp.y;
        s.source_animated = sourceAnimated ? 1 : 0;
        s.source_range[0] = sourceRange.x; s.source_range[1] = sourceRange.y;
        s.source_max_particle_count = sourceMaxParticleCount;
        s.noise_seed = noiseSeed; s.residual_tol = residualTol; s.residual_check_every = residualCheckEvery;
        s.cfl_number = cflNumber; s.max_substeps = maxSubsteps; s.smoke_density = smokeDensity;
        s.warm_start = warmStart;
        s.residual_check_adaptive = residualCheckAdaptive;
        return s;
    }
};
