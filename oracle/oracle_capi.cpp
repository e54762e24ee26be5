/*
 * oracle_capi.cpp — extern "C" access to the CPU oracle for tests/ and
 * bench.py's cpu_baseline leg (TEST INFRASTRUCTURE ONLY; the product never
 * links this library).
 */
#include <cstdio>
#include <cstring>
#include <new>

#include "blaze_oracle.hpp"
#include "render_oracle.hpp"

namespace {

struct Handle {
    int fp64 = 0;
    orc::Solver<float> f;
    orc::Solver<double> d;
};

template <class R> std::vector<R>* raw_field(orc::Solver<R>& s, int which, bool back) {
    if (which == BLZ_FIELD_P) return &s.p;
    if (which == BLZ_FIELD_RHS) return &s.rhs;
    orc::Grid<R>* g = s.grid(which);
    if (!g) return nullptr;
    return back ? &g->back : &g->front;
}

template <class R> int dims_of(orc::Solver<R>& s, int which, int* dims) {
    if (which == BLZ_FIELD_P || which == BLZ_FIELD_RHS) {
        dims[0] = s.W; dims[1] = s.H; dims[2] = s.D;
        return 0;
    }
    orc::Grid<R>* g = s.grid(which);
    if (!g) return -1;
    dims[0] = g->nx; dims[1] = g->ny; dims[2] = g->nz;
    return 0;
}

} // namespace

extern "C" {

void* orc_create(const blz_settings* s, int fp64) {
    Handle* h = new (std::nothrow) Handle();
    if (!h) return nullptr;
    h->fp64 = fp64;
    if (fp64) h->d.init(*s); else h->f.init(*s);
    return h;
}

void orc_destroy(void* p) { delete (Handle*)p; }

void orc_set_threads(int n) {
#ifdef _OPENMP
    omp_set_num_threads(n > 0 ? n : 1);
#else
    (void)n;
#endif
}

int orc_max_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

void orc_set_rbgs_order(void* p, int interleaved) {
    Handle* h = (Handle*)p;
    h->f.interleaved_rbgs = interleaved != 0;
    h->d.interleaved_rbgs = interleaved != 0;
}

int orc_add_source(void* p, const float* part8, int n) {
    Handle* h = (Handle*)p;
    if (h->fp64) h->d.add_source(part8, n); else h->f.add_source(part8, n);
    return 0;
}

int orc_step(void* p, float time) {
    Handle* h = (Handle*)p;
    if (h->fp64) h->d.step(time); else h->f.step(time);
    return 0;
}

int orc_run_stage(void* p, int stage, float time, float dt) {
    Handle* h = (Handle*)p;
    if (stage < 0 || stage >= BLZ_NUM_STAGES) return BLZ_E_INVALID;
    if (h->fp64) {
        h->d.run_stage(stage, (double)time, (double)dt);
    } else {
        h->f.run_stage(stage, time, dt);
    }
    return 0;
}

int orc_field_dims(void* p, int which, int* dims) {
    Handle* h = (Handle*)p;
    return h->fp64 ? dims_of(h->d, which, dims) : dims_of(h->f, which, dims);
}

/* copy a field out as double (exact for float fields) */
int orc_copy_field(void* p, int which, int back, double* out) {
    Handle* h = (Handle*)p;
    if (h->fp64) {
        auto* v = raw_field(h->d, which, back != 0);
        if (!v) return BLZ_E_INVALID;
        std::memcpy(out, v->data(), v->size() * sizeof(double));
    } else {
        auto* v = raw_field(h->f, which, back != 0);
        if (!v) return BLZ_E_INVALID;
        for (size_t i = 0; i < v->size(); ++i) out[i] = (double)(*v)[i];
    }
    return 0;
}

int orc_set_field(void* p, int which, const double* in) {
    Handle* h = (Handle*)p;
    if (h->fp64) {
        auto* v = raw_field(h->d, which, false);
        if (!v) return BLZ_E_INVALID;
        std::memcpy(v->data(), in, v->size() * sizeof(double));
    } else {
        auto* v = raw_field(h->f, which, false);
        if (!v) return BLZ_E_INVALID;
        for (size_t i = 0; i < v->size(); ++i) (*v)[i] = (float)in[i];
    }
    return 0;
}

/* set a back buffer (for pinning stages that read them) */
int orc_set_field_back(void* p, int which, const double* in) {
    Handle* h = (Handle*)p;
    if (which == BLZ_FIELD_P || which == BLZ_FIELD_RHS) return BLZ_E_INVALID;
    if (h->fp64) {
        auto* v = raw_field(h->d, which, true);
        if (!v) return BLZ_E_INVALID;
        std::memcpy(v->data(), in, v->size() * sizeof(double));
    } else {
        auto* v = raw_field(h->f, which, true);
        if (!v) return BLZ_E_INVALID;
        for (size_t i = 0; i < v->size(); ++i) (*v)[i] = (float)in[i];
    }
    return 0;
}

/* the curl part of the turbulence stage alone (add_curl_noise_kernel,
   src/FluidSolver.cu:271-293), reading the back buffers as they are */
int orc_curl_noise(void* p, float dt) {
    Handle* h = (Handle*)p;
    if (h->fp64) h->d.curl_noise((double)dt); else h->f.curl_noise(dt);
    return 0;
}

int orc_get_stats(void* p, int* substeps, int* iters_total, int* iters_last, double* res, double* res0) {
    Handle* h = (Handle*)p;
    if (h->fp64) {
        *substeps = h->d.substeps; *iters_total = h->d.iters_total; *iters_last = h->d.iters_last;
        *res = h->d.res; *res0 = h->d.res0;
    } else {
        *substeps = h->f.substeps; *iters_total = h->f.iters_total; *iters_last = h->f.iters_last;
        *res = h->f.res; *res0 = h->f.res0;
    }
    return 0;
}

double orc_residual(void* p, float dt) {
    Handle* h = (Handle*)p;
    return h->fp64 ? h->d.residual((double)dt) : h->f.residual(dt);
}

/* ---- primitives, for closed-form pin tests ---- */
float orc_cerp_f32(float q0, float q1, float q2, float q3, float s) {
    return orc::Grid<float>::cerp(q0, q1, q2, q3, s);
}
double orc_cerp_f64(double q0, double q1, double q2, double q3, double s) {
    return orc::Grid<double>::cerp(q0, q1, q2, q3, s);
}
float orc_cubic_pulse_f32(float x, float w) { return orc::Grid<float>::cubic_pulse(x, w); }
float orc_perlin_f32(float x, float y, float z, float scale, int seed) {
    return orc::perlin<float>(orc::mk(x, y, z), scale, seed);
}
double orc_perlin_f64(double x, double y, double z, double scale, int seed) {
    return orc::perlin<double>(orc::mk(x, y, z), scale, seed);
}
unsigned orc_noise_hash(unsigned s) { return orc::noise_hash(s); }
unsigned orc_lattice_hash(float x, float y, float z, float seed) { return orc::lattice_hash(x, y, z, seed); }

/* sample one field of a solver at an object-space position */
double orc_sample(void* p, int which, int cubic, int clamp, double x, double y, double z) {
    Handle* h = (Handle*)p;
    if (h->fp64) {
        auto* g = h->d.grid(which);
        if (!g) return 0.0;
        auto pos = orc::mk<double>(x, y, z);
        return cubic ? g->tricubic(pos, clamp != 0) : g->trilinear(pos);
    }
    auto* g = h->f.grid(which);
    if (!g) return 0.0;
    auto pos = orc::mk<float>((float)x, (float)y, (float)z);
    return cubic ? g->tricubic(pos, clamp != 0) : g->trilinear(pos);
}

// the renderer over a caller-given temperature grid (rows row0, row0 +
// stride, ... below row1)
int orc_render(const float* T, int W, int H, int D, float dx, const float* dmin, const float* dmax,
               const blz_render_settings* r, float* rgb, int row0, int row1, int stride) {
    if (!T || !dmin || !dmax || !r || !rgb || W <= 0 || H <= 0 || D <= 0 || r->width <= 0 || r->height <= 0 ||
        row0 < 0 || row1 > r->height || row0 > row1 || stride < 1 || !(r->p_step > 0.0f) || !(r->s_step > 0.0f) ||
        r->light_count < 0 || r->light_count > BLZ_MAX_LIGHTS || r->ramp_size < 1 || r->ramp_size > BLZ_MAX_RAMP)
        return -1;
    rnd::render(T, W, H, D, dx, dmin, dmax, *r, rgb, row0, row1, stride);
    return 0;
}

float orc_bexp(float x) { return rnd::bexp(x); }

} // extern "C"
