// scene.cpp — host-side scene description: settings defaults, the .blz JSON
// reader, the .part particle reader and a deterministic particle generator.
//
// Reference: Parser::sceneParse (asiroisvigneux/Blaze src/Parser.h:41-172),
// Parser::sourceParticleParse (src/Parser.h:174-211), SceneSettings
// (src/SceneSettings.h:32-71). The reference parses with RapidJSON and
// asserts on a missing key; here a missing required key returns
// BLZ_E_PARSE with the key named in blz_last_error().
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>

#include "common.h"
#include "json.h"

using blz::json::Value;

extern "C" int blz_abi_version(void) { return BLZ_ABI_VERSION; }

extern "C" int blz_settings_default(blz_settings* s) {
    if (!s) {
        blz::set_error("blz_settings_default: null settings");
        return BLZ_E_INVALID;
    }
    std::memset(s, 0, sizeof(*s));
    s->grid_res[0] = s->grid_res[1] = s->grid_res[2] = 64;
    s->dx = 1.0f / 64.0f;
    s->dt = 1.0f / 24.0f;
    for (int a = 0; a < 3; ++a) {
        s->domain_min[a] = 0.0f;
        s->domain_max[a] = 1.0f;
    }
    for (int a = 0; a < 6; ++a) s->closed_bounds[a] = 1;
    s->max_iter_solve = 100;
    s->density = 1.0f;
    s->gravity = -9.8f;
    s->wind_turb_scale = 1.0f;
    s->turbulence_scale = 1.0f;
    s->turb_mask_temp_ramp[1] = 1.0f;
    s->turb_mask_vel_ramp[1] = 1.0f;
    s->source_range[0] = 1;
    s->source_range[1] = 1;
    s->noise_seed = 123; // reference: compute_turbulence_kernel, src/FluidSolver.cu:304
    s->residual_tol = 0.0f;
    s->residual_check_every = 10;
    s->cfl_number = 0.0f;
    s->max_substeps = 8;
    s->smoke_density = 0.0f;
    s->warm_start = 0;
    s->residual_check_adaptive = 0;
    return BLZ_OK;
}

namespace {

struct Reader {
    const Value& root;
    std::string missing;
    bool ok = true;
    explicit Reader(const Value& r) : root(r) {}

    const Value* get(const char* key, bool required) {
        const Value* v = root.find(key);
        if (!v && required && ok) {
            ok = false;
            missing = key;
        }
        return v;
    }
    void num(const char* key, float* out, bool required = true) {
        const Value* v = get(key, required);
        if (v && v->is_number()) *out = (float)v->num;
        else if (v && ok) { ok = false; missing = std::string(key) + " (not a number)"; }
    }
    void integer(const char* key, int* out, bool required = true) {
        const Value* v = get(key, required);
        if (v && v->is_number()) *out = (int)v->num;
        else if (v && ok) { ok = false; missing = std::string(key) + " (not a number)"; }
    }
    // array of n numbers; as_int casts each element after reading it as float
    // (src/Parser.h:58-64 reads ranges with GetFloat then casts)
    void arr(const char* key, int n, float* f, int* i, bool required = true) {
        const Value* v = get(key, required);
        if (!v) return;
        if (!v->is_array() || (int)v->arr.size() < n) {
            if (ok) { ok = false; missing = std::string(key) + " (array too short)"; }
            return;
        }
        for (int a = 0; a < n; ++a) {
            const Value& e = v->arr[a];
            if (!e.is_number()) {
                if (ok) { ok = false; missing = std::string(key) + " (element not a number)"; }
                return;
            }
            if (f) f[a] = (float)e.num;
            if (i) i[a] = (int)(float)e.num;
        }
    }
};

std::string dir_of(const std::string& path) {
    size_t pos = path.find_last_of("\\/");
    return pos == std::string::npos ? std::string("") : path.substr(0, pos);
}

void copy_str(char* dst, int len, const std::string& src) {
    if (!dst || len <= 0) return;
    std::strncpy(dst, src.c_str(), (size_t)len - 1);
    dst[len - 1] = 0;
}

} // namespace

extern "C" int blz_parse_scene(const char* path, blz_settings* out, char* scene_dir, int scene_dir_len,
                               char* particle_file, int particle_file_len, blz_particle_gen* gen_out,
                               int* has_gen) {
    if (!path || !out) {
        blz::set_error("blz_parse_scene: null argument");
        return BLZ_E_INVALID;
    }
    std::ifstream f(path, std::ios::binary);
    if (!f.good()) {
        blz::set_error(std::string("scene file does not exist: ") + path);
        return BLZ_E_IO;
    }
    std::stringstream ss;
    ss << f.rdbuf();
    Value root;
    std::string err;
    if (!blz::json::parse(ss.str(), root, err) || !root.is_object()) {
        blz::set_error(std::string("malformed .blz ") + path + ": " + (err.empty() ? "not an object" : err));
        return BLZ_E_PARSE;
    }
    blz_settings s;
    blz_settings_default(&s);
    Reader r(root);
    // solver keys, all required as in src/Parser.h:41-108
    std::string pfile;
    if (const Value* v = r.get("particleFile", true)) pfile = v->is_string() ? v->str : "";
    r.integer("sourceAnimated", &s.source_animated);
    r.arr("sourceRange", 2, nullptr, s.source_range);
    r.integer("sourceMaxParticleCount", &s.source_max_particle_count);
    r.integer("maxIterSolve", &s.max_iter_solve);
    r.num("dx", &s.dx);
    r.num("dt", &s.dt);
    r.arr("gridRes", 3, nullptr, s.grid_res);
    r.arr("domainBboxMin", 3, s.domain_min, nullptr);
    r.arr("domainBboxMax", 3, s.domain_max, nullptr);
    r.arr("closedBounds", 6, nullptr, s.closed_bounds);
    r.num("density", &s.density);
    r.num("gravity", &s.gravity);
    r.num("coolingRate", &s.cooling_rate);
    r.num("buoyancy", &s.buoyancy);
    r.num("vorticityConf", &s.vorticity_conf);
    r.num("drag", &s.drag);
    r.num("windAmp", &s.wind_amp);
    r.num("windSpeed", &s.wind_speed);
    r.num("windTurbAmp", &s.wind_turb_amp);
    r.num("windTurbScale", &s.wind_turb_scale);
    r.integer("windDir", &s.wind_dir);
    r.num("turbulenceAmp", &s.turbulence_amp);
    r.num("turbulenceScale", &s.turbulence_scale);
    r.arr("turbMaskTempRamp", 2, s.turb_mask_temp_ramp, nullptr);
    r.arr("turbMaskVelRamp", 2, s.turb_mask_vel_ramp, nullptr);
    // extensions, optional
    r.integer("noiseSeed", &s.noise_seed, false);
    r.num("residualTol", &s.residual_tol, false);
    r.integer("residualCheckEvery", &s.residual_check_every, false);
    r.num("cflNumber", &s.cfl_number, false);
    r.integer("maxSubsteps", &s.max_substeps, false);
    r.num("smokeDensity", &s.smoke_density, false);
    r.integer("warmStart", &s.warm_start, false);
    r.integer("residualCheckAdaptive", &s.residual_check_adaptive, false);
    if (!r.ok) {
        blz::set_error(std::string("missing or invalid .blz key: ") + r.missing);
        return BLZ_E_PARSE;
    }
    int hg = 0;
    if (const Value* g = root.find("particleGen")) {
        if (!g->is_object()) {
            blz::set_error("particleGen must be an object");
            return BLZ_E_PARSE;
        }
        blz_particle_gen pg;
        std::memset(&pg, 0, sizeof(pg));
        Reader gr(*g);
        gr.integer("count", &pg.count);
        if (const Value* sv = g->find("seed")) pg.seed = (unsigned long long)sv->num;
        gr.arr("center", 3, pg.center, nullptr);
        gr.num("radius", &pg.radius);
        gr.num("pscale", &pg.pscale);
        gr.num("temp", &pg.temp);
        gr.arr("vel", 3, pg.vel, nullptr, false);
        gr.num("radialVel", &pg.radial_vel, false);
        if (!gr.ok) {
            blz::set_error(std::string("missing or invalid particleGen key: ") + gr.missing);
            return BLZ_E_PARSE;
        }
        if (gen_out) *gen_out = pg;
        hg = 1;
    }
    if (has_gen) *has_gen = hg;
    // sceneDir = directory of the real path (src/Parser.h:43-45)
    char real[4096];
    std::string rp = realpath(path, real) ? std::string(real) : std::string(path);
    copy_str(scene_dir, scene_dir_len, dir_of(rp));
    copy_str(particle_file, particle_file_len, pfile);
    *out = s;
    return BLZ_OK;
}

extern "C" int blz_count_particles(const char* path, int* count) {
    if (!path || !count) {
        blz::set_error("blz_count_particles: null argument");
        return BLZ_E_INVALID;
    }
    std::ifstream f(path, std::ios::binary | std::ios::ate);
    if (!f.good()) {
        blz::set_error(std::string("Particle file does not exist on disk: ") + path);
        return BLZ_E_IO;
    }
    const std::streamoff bytes = f.tellg();
    *count = (int)(bytes / (8 * (std::streamoff)sizeof(float)));
    return BLZ_OK;
}

// Parser::sourceParticleParse (src/Parser.h:174-211): whole records only
extern "C" int blz_read_particles(const char* path, float* part8, int max_count, int* count) {
    if (!path || !count || (!part8 && max_count > 0)) {
        blz::set_error("blz_read_particles: null argument");
        return BLZ_E_INVALID;
    }
    std::ifstream f(path, std::ios::binary);
    if (!f.good()) {
        blz::set_error(std::string("Particle file does not exist on disk: ") + path);
        return BLZ_E_IO;
    }
    int n = 0;
    float rec[8];
    while (n < max_count && f.read(reinterpret_cast<char*>(rec), sizeof(rec))) {
        std::memcpy(part8 + (size_t)n * 8, rec, sizeof(rec));
        ++n;
    }
    *count = n;
    return BLZ_OK;
}

namespace {
// SplitMix64 (Steele, Lea, Flood 2014)
struct SplitMix {
    unsigned long long x;
    unsigned long long next() {
        unsigned long long z = (x += 0x9e3779b97f4a7c15ULL);
        z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
        z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
        return z ^ (z >> 31);
    }
    double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); } // [0,1)
};
} // namespace

// Uniform points in a ball by rejection from the cube (only +,-,*,sqrt in
// double, so the set is bit-identical on every IEEE host).
extern "C" int blz_generate_particles(const blz_particle_gen* g, float* part8) {
    if (!g || (!part8 && g->count > 0) || g->count < 0) {
        blz::set_error("blz_generate_particles: bad argument");
        return BLZ_E_INVALID;
    }
    SplitMix rng{g->seed};
    for (int q = 0; q < g->count; ++q) {
        double x, y, z, r2;
        do {
            x = rng.uniform() * 2.0 - 1.0;
            y = rng.uniform() * 2.0 - 1.0;
            z = rng.uniform() * 2.0 - 1.0;
            r2 = x * x + y * y + z * z;
        } while (r2 > 1.0);
        const double rad = (double)g->radius;
        float* o = part8 + (size_t)q * 8;
        o[0] = (float)((double)g->center[0] + rad * x);
        o[1] = (float)((double)g->center[1] + rad * y);
        o[2] = (float)((double)g->center[2] + rad * z);
        const double len = std::sqrt(r2);
        const double ux = len > 0.0 ? x / len : 0.0, uy = len > 0.0 ? y / len : 0.0, uz = len > 0.0 ? z / len : 0.0;
        o[3] = (float)((double)g->vel[0] + (double)g->radial_vel * ux);
        o[4] = (float)((double)g->vel[1] + (double)g->radial_vel * uy);
        o[5] = (float)((double)g->vel[2] + (double)g->radial_vel * uz);
        o[6] = g->pscale;
        o[7] = g->temp;
    }
    return BLZ_OK;
}

// ---- render settings (src/Parser.h:109-170: lights, cam, sampling, shader,
// colour ramp, multi-scatter; all required there and here) ----

extern "C" int blz_render_settings_default(blz_render_settings* r) {
    if (!r) {
        blz::set_error("blz_render_settings_default: null settings");
        return BLZ_E_INVALID;
    }
    std::memset(r, 0, sizeof(*r));
    r->width = 64;
    r->height = 64;
    r->cam_focal = 50.0f;
    r->cam_aperture = 41.4214f;
    r->cam_u[0] = 1.0f;
    r->cam_v[1] = 1.0f;
    r->cam_w[2] = 1.0f;
    r->cam_trans[0] = 0.5f;
    r->cam_trans[1] = 0.5f;
    r->cam_trans[2] = 3.0f;
    r->p_step = 1.0f / 64.0f;
    r->s_step = 2.0f / 64.0f;
    r->cutoff = 0.01f;
    r->density_scale = 1.0f;
    r->shadow_density_scale = 1.0f;
    for (int a = 0; a < 3; ++a) {
        r->volume_color[a] = 1.0f;
        r->scattering[a] = 1.0f;
    }
    r->ramp_size = 1;
    r->ramp_colors[0][0] = r->ramp_colors[0][1] = r->ramp_colors[0][2] = 1.0f;
    r->temp_range[1] = 1.0f;
    r->ms_mask[1] = 1.0f;
    return BLZ_OK;
}

extern "C" int blz_parse_render(const char* path, blz_render_settings* out) {
    if (!path || !out) {
        blz::set_error("blz_parse_render: null argument");
        return BLZ_E_INVALID;
    }
    std::ifstream f(path, std::ios::binary);
    if (!f.good()) {
        blz::set_error(std::string("scene file does not exist: ") + path);
        return BLZ_E_IO;
    }
    std::stringstream ss;
    ss << f.rdbuf();
    Value root;
    std::string err;
    if (!blz::json::parse(ss.str(), root, err) || !root.is_object()) {
        blz::set_error(std::string("malformed .blz ") + path + ": " + (err.empty() ? "not an object" : err));
        return BLZ_E_PARSE;
    }
    blz_render_settings r;
    blz_render_settings_default(&r);
    Reader rd(root);
    rd.arr("renderRes", 2, nullptr, &r.width);
    rd.integer("lightCount", &r.light_count);
    rd.num("primarySampleRate", &r.p_step);
    rd.num("shadowSamplingRate", &r.s_step);
    rd.num("cutoff", &r.cutoff);
    rd.num("densityScale", &r.density_scale);
    rd.num("shadowDensityScale", &r.shadow_density_scale);
    rd.arr("volumeColor", 3, r.volume_color, nullptr);
    rd.arr("scatteringColor", 3, r.scattering, nullptr);
    rd.num("absorption", &r.absorption);
    rd.num("gain", &r.gain);
    rd.num("emissionScale", &r.emission_scale);
    rd.arr("shaderTempRange", 2, r.temp_range, nullptr);
    rd.num("multiScatterScale", &r.ms_scale);
    rd.arr("multiScatterDensityMask", 2, r.ms_mask, nullptr);
    rd.arr("multiScatterColor", 3, r.ms_color, nullptr);
    rd.integer("multiScatterBlurIter", &r.ms_blur_iter);
    std::string bad;
    if (rd.ok && (r.light_count < 0 || r.light_count > BLZ_MAX_LIGHTS)) bad = "lightCount (0..10)";
    // lights[i]: light_exposure, light_intensity, light_samples, light_dir, light_color
    const Value* lights = rd.ok ? rd.get("lights", true) : nullptr;
    if (rd.ok && bad.empty()) {
        if (!lights || !lights->is_array() || (int)lights->arr.size() < r.light_count) {
            bad = "lights (fewer than lightCount)";
        } else {
            for (int i = 0; i < r.light_count && bad.empty(); ++i) {
                Reader l(lights->arr[i]);
                l.num("light_exposure", &r.light_exposure[i]);
                l.num("light_intensity", &r.light_intensity[i]);
                l.integer("light_samples", &r.light_samples[i]);
                l.arr("light_dir", 3, r.light_dir[i], nullptr);
                l.arr("light_color", 3, r.light_color[i], nullptr);
                if (!l.ok) bad = "lights[" + std::to_string(i) + "]." + l.missing;
            }
        }
    }
    // cam: focal, aperture, rotMat rows -> u, v, w, t
    if (rd.ok && bad.empty()) {
        const Value* cam = rd.get("cam", true);
        if (cam && cam->is_object()) {
            Reader c(*cam);
            c.num("focal", &r.cam_focal);
            c.num("aperture", &r.cam_aperture);
            c.arr("t", 3, r.cam_trans, nullptr);
            const Value* rm = c.get("rotMat", true);
            if (c.ok && (!rm || !rm->is_array() || rm->arr.size() < 3)) bad = "cam.rotMat (3 rows)";
            float* rows[3] = {r.cam_u, r.cam_v, r.cam_w};
            for (int q = 0; q < 3 && c.ok && bad.empty(); ++q) {
                const Value& row = rm->arr[q];
                if (!row.is_array() || row.arr.size() < 3) {
                    bad = "cam.rotMat row";
                    break;
                }
                for (int a = 0; a < 3; ++a) {
                    if (!row.arr[a].is_number()) {
                        bad = "cam.rotMat element";
                        break;
                    }
                    rows[q][a] = (float)row.arr[a].num;
                }
            }
            if (!c.ok && bad.empty()) bad = "cam." + c.missing;
        } else if (rd.ok) {
            bad = "cam (object)";
        }
    }
    // shaderColorRamp: count, keys, colors
    if (rd.ok && bad.empty()) {
        const Value* ramp = rd.get("shaderColorRamp", true);
        if (ramp && ramp->is_object()) {
            Reader c(*ramp);
            c.integer("count", &r.ramp_size);
            if (c.ok && (r.ramp_size < 1 || r.ramp_size > BLZ_MAX_RAMP)) bad = "shaderColorRamp.count (1..10)";
            if (c.ok && bad.empty()) {
                c.arr("keys", r.ramp_size, r.ramp_keys, nullptr);
                const Value* cols = c.get("colors", true);
                if (c.ok && (!cols || !cols->is_array() || (int)cols->arr.size() < r.ramp_size)) {
                    bad = "shaderColorRamp.colors";
                } else {
                    for (int q = 0; q < r.ramp_size && c.ok && bad.empty(); ++q) {
                        const Value& col = cols->arr[q];
                        if (!col.is_array() || col.arr.size() < 3) {
                            bad = "shaderColorRamp.colors row";
                            break;
                        }
                        for (int a = 0; a < 3; ++a) r.ramp_colors[q][a] = (float)col.arr[a].num;
                    }
                }
            }
            if (!c.ok && bad.empty()) bad = "shaderColorRamp." + c.missing;
        } else if (rd.ok) {
            bad = "shaderColorRamp (object)";
        }
    }
    if (!rd.ok) {
        blz::set_error(std::string("missing or invalid .blz render key: ") + rd.missing);
        return BLZ_E_PARSE;
    }
    if (!bad.empty()) {
        blz::set_error(std::string("missing or invalid .blz render key: ") + bad);
        return BLZ_E_PARSE;
    }
    *out = r;
    return BLZ_OK;
}
