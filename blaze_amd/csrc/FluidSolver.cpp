// FluidSolver.cpp — C++ drop-in solver and scene parser over the C ABI.
//
// Reference: FluidSolver (asiroisvigneux/Blaze src/FluidSolver.cu:671-955),
// Parser (src/Parser.h:41-219). Error convention of the reference
// (checkCudaErrors, src/utils/helper_cuda.h:576-591): print and exit.
#include "FluidSolver.h"

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>

#include "Parser.h"
#include "json.h"

namespace {
void check(int rc, const char* what) {
    if (rc != BLZ_OK) {
        std::fprintf(stderr, "blaze: %s failed (%d): %s\n", what, rc, blz_last_error());
        std::exit(EXIT_FAILURE);
    }
}
void copy_str(char* dst, size_t len, const std::string& s) {
    std::strncpy(dst, s.c_str(), len - 1);
    dst[len - 1] = 0;
}
} // namespace

std::string Parser::getDirPath(const std::string& fname) {
    size_t pos = fname.find_last_of("\\/");
    return (std::string::npos == pos) ? "" : fname.substr(0, pos);
}

void Parser::sceneParse(SceneSettings* scene, const char* sceneFile) {
    std::memset(scene, 0, sizeof(*scene));
    blz_settings s;
    char dir[256], pfile[256];
    blz_particle_gen gen;
    int hasGen = 0;
    check(blz_parse_scene(sceneFile, &s, dir, sizeof(dir), pfile, sizeof(pfile), &gen, &hasGen), "sceneParse");
    copy_str(scene->sceneDir, sizeof(scene->sceneDir), dir);
    copy_str(scene->particleFile, sizeof(scene->particleFile), pfile);
    // descriptive keys the solver does not use (optional here)
    std::ifstream f(sceneFile, std::ios::binary);
    std::stringstream ss;
    ss << f.rdbuf();
    blz::json::Value root;
    std::string err;
    if (blz::json::parse(ss.str(), root, err)) {
        if (const auto* v = root.find("sceneName"); v && v->is_string())
            copy_str(scene->sceneName, sizeof(scene->sceneName), v->str);
        if (const auto* v = root.find("renderFile"); v && v->is_string())
            copy_str(scene->renderFile, sizeof(scene->renderFile), v->str);
        if (const auto* v = root.find("renderRange"); v && v->is_array() && v->arr.size() >= 2) {
            scene->renderRange.x = (int)(float)v->arr[0].num;
            scene->renderRange.y = (int)(float)v->arr[1].num;
        }
        if (const auto* v = root.find("renderRes"); v && v->is_array() && v->arr.size() >= 2) {
            scene->renderRes.x = (int)v->arr[0].num;
            scene->renderRes.y = (int)v->arr[1].num;
        }
    }
    // renderer keys: all present, or the scene is simulated without frames
    scene->hasRender = blz_parse_render(sceneFile, &scene->render) == BLZ_OK;
    scene->sourceAnimated = s.source_animated != 0;
    scene->sourceRange = {s.source_range[0], s.source_range[1]};
    scene->sourceMaxParticleCount = s.source_max_particle_count;
    scene->domainBboxMin = {s.domain_min[0], s.domain_min[1], s.domain_min[2]};
    scene->domainBboxMax = {s.domain_max[0], s.domain_max[1], s.domain_max[2]};
    for (int a = 0; a < 6; ++a) scene->closedBounds[a] = s.closed_bounds[a] != 0;
    scene->maxIterSolve = s.max_iter_solve;
    scene->dx = s.dx;
    scene->dt = s.dt;
    scene->gridRes = {s.grid_res[0], s.grid_res[1], s.grid_res[2]};
    scene->density = s.density;
    scene->gravity = s.gravity;
    scene->buoyancy = s.buoyancy;
    scene->coolingRate = s.cooling_rate;
    scene->vorticityConf = s.vorticity_conf;
    scene->drag = s.drag;
    scene->windDir = s.wind_dir;
    scene->windAmp = s.wind_amp;
    scene->windSpeed = s.wind_speed;
    scene->windTurbAmp = s.wind_turb_amp;
    scene->windTurbScale = s.wind_turb_scale;
    scene->turbulence_amp = s.turbulence_amp;
    scene->turbulence_scale = s.turbulence_scale;
    scene->turbMaskTempRamp = {s.turb_mask_temp_ramp[0], s.turb_mask_temp_ramp[1]};
    scene->turbMaskVelRamp = {s.turb_mask_vel_ramp[0], s.turb_mask_vel_ramp[1]};
    scene->noiseSeed = s.noise_seed;
    scene->residualTol = s.residual_tol;
    scene->residualCheckEvery = s.residual_check_every;
    scene->cflNumber = s.cfl_number;
    scene->maxSubsteps = s.max_substeps;
    scene->smokeDensity = s.smoke_density;
    scene->warmStart = s.warm_start;
    scene->residualCheckAdaptive = s.residual_check_adaptive;
    scene->hasParticleGen = hasGen != 0;
    if (hasGen) scene->particleGen = gen;
}

void Parser::sourceParticleParse(const SceneSettings* scn, std::vector<float>& part8, int& pointCount,
                                 int frameNum) {
    if (scn->hasParticleGen) {
        part8.resize((size_t)scn->particleGen.count * 8);
        check(blz_generate_particles(&scn->particleGen, part8.data()), "particle generator");
        pointCount = scn->particleGen.count;
        return;
    }
    char particleFile[1024];
    if (scn->sourceAnimated) { // <prefix>_%04d.part, src/Parser.h:182-185
        std::string name(scn->particleFile);
        std::string sub = name.size() >= 10 ? name.substr(0, name.size() - 10) : name;
        std::snprintf(particleFile, sizeof(particleFile), "%s/%s_%04d.part", scn->sceneDir, sub.c_str(), frameNum);
    } else {
        std::snprintf(particleFile, sizeof(particleFile), "%s/%s", scn->sceneDir, scn->particleFile);
    }
    int n = 0;
    check(blz_count_particles(particleFile, &n), "sourceParticleParse");
    part8.resize((size_t)n * 8);
    check(blz_read_particles(particleFile, part8.data(), n, &pointCount), "sourceParticleParse");
}

FluidSolver::FluidSolver(Timer* tmr, SceneSettings* scn) : mScn(scn), mTmr(tmr) {
    const blz_settings s = scn->toC();
    check(blz_create(&s, &mH), "FluidSolver");
    check(blz_set_timing(mH, 1), "FluidSolver");
}

FluidSolver::~FluidSolver() {
    if (mPrefetch.valid()) mPrefetch.wait();
    blz_destroy(mH);
}

// addSource, src/FluidSolver.cu:770-821. Animated .part sources: the frame
// was read and its upload started by the previous call's worker (or is read
// now); the worker then reads frame + 1 and stages it while this frame's
// step runs (the reference parses and copies synchronously).
void FluidSolver::addSource() {
    const bool inRange = mTmr->iter >= mScn->sourceRange.x && mTmr->iter <= mScn->sourceRange.y;
    if (mScn->sourceAnimated && !mScn->hasParticleGen) {
        if (!inRange) return;
        if (mPrefetch.valid()) check(mPrefetch.get(), "addSource (staged upload)");
        if (mPrefetchFrame != mTmr->iter) {
            Parser::sourceParticleParse(mScn, mPart, mParticleCount, mTmr->iter);
            check(blz_stage_source(mH, mPart.data(), mParticleCount), "addSource");
        }
        check(blz_add_source_staged(mH), "addSource");
        const int next = mTmr->iter + 1;
        mPrefetchFrame = -1;
        if (next <= mScn->sourceRange.y) {
            mPrefetchFrame = next;
            mPrefetch = std::async(std::launch::async, [this, next]() -> int {
                std::vector<float> buf;
                int n = 0;
                Parser::sourceParticleParse(mScn, buf, n, next);
                return blz_stage_source(mH, buf.data(), n);
            });
        }
        return;
    }
    if (!mScn->sourceAnimated || inRange) {
        bool fresh = false;
        if (!mSingleFrameSourceInit || mScn->sourceAnimated) {
            Parser::sourceParticleParse(mScn, mPart, mParticleCount, mTmr->iter);
            mSingleFrameSourceInit = true;
            fresh = true;
        }
        if (mParticleCount > 0)
            check(blz_add_source(mH, fresh ? mPart.data() : nullptr, mParticleCount), "addSource");
    }
}

void FluidSolver::step() {
    check(blz_step(mH, mTmr->time), "step");
    blz_stats st;
    check(blz_get_stats(mH, &st), "step");
    mTmr->last = st;
}

GridHandle FluidSolver::getTempGrid() {
    void* p = nullptr;
    int dims[3];
    check(blz_get_field(mH, BLZ_FIELD_T, &p, dims), "getTempGrid");
    return GridHandle{(const float*)p, dims[0], dims[1], dims[2], 0.5f, 0.5f, 0.5f, mScn->dx};
}
