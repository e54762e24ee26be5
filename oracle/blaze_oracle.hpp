/*
 * blaze_oracle.hpp — CPU restatement of the Blaze solver step (TEST
 * INFRASTRUCTURE ONLY).
 *
 * This is the parity checker for the HIP product in blaze_amd/. Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it. The
 * product never links, calls or falls back to it.
 *
 * It transcribes asiroisvigneux/Blaze's solver kernels (src/FluidSolver.cu,
 * src/Grid.h, thirdparty/cuda-noise/cuda_noise.cuh) as serial loops, with
 * these deterministic orders replacing the reference's races:
 *   - particle-to-grid: per cell, contributions summed in ascending particle
 *     index from +0 (reference: float atomicAdd, src/Grid.h:331);
 *   - vorticity / curl-noise face updates: (face + lower cell) + upper cell
 *     (reference: float atomicAdd, src/FluidSolver.cu:242-247,287-292);
 *   - RBGS: per sweep, all red (i+j even) columns walk k=0..D-1, then all
 *     black columns (equal to the reference's per-slab red/black order
 *     inside one block, src/FluidSolver.cu:526-598; both orders are here);
 *   - pressure boundary faces copy the post-update inward face
 *     (src/FluidSolver.cu:630-668).
 * Floating-point policy: build with -ffp-contract=off, no fast-math; the
 * evaluation order of every expression follows the reference source.
 *
 * Templated on R (float: parity oracle; double: FP64 error reference).
 * In double mode the Perlin lattice hash input stays float so that both
 * precisions see the same noise lattice.
 *
 * Parity pins: every function below is checked in tests/test_oracle_*.py
 * against closed forms, invariants or brute force (see the test docstrings).
 */
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <type_traits>
#include <vector>

#include "../include/blaze.h"

#ifdef _OPENMP
#include <omp.h>
#endif

namespace orc {

/* Constant of the reference written as a float literal: exactly that float
   in float mode, the decimal value in double mode. */
#define ORC_K(R, lit) ((R)(std::is_same<R, float>::value ? (double)(lit##f) : (lit)))

template <class R> struct V3 {
    R x, y, z;
};
template <class R> static inline V3<R> mk(R x, R y, R z) { return V3<R>{x, y, z}; }
/* helper_math.h:1267-1269 dot, :1314-1316 length, :1443-1446 cross */
template <class R> static inline R dot3(V3<R> a, V3<R> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <class R> static inline R len3(V3<R> a) { return std::sqrt(dot3(a, a)); }
template <class R> static inline V3<R> cross3(V3<R> a, V3<R> b) {
    return mk<R>(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}

/* ------------------------------------------------------------------ */
/* Perlin noise, thirdparty/cuda-noise/cuda_noise.cuh                  */
/* ------------------------------------------------------------------ */

/* hash :46-56 (32-bit integer mix) */
static inline uint32_t noise_hash(uint32_t s) {
    s = (s + 0x7ed55d16u) + (s << 12);
    s = (s ^ 0xc761c23cu) ^ (s >> 19);
    s = (s + 0x165667b1u) + (s << 5);
    s = (s + 0xd3a2646cu) ^ (s << 9);
    s = (s + 0xfd7046c5u) + (s << 3);
    s = (s ^ 0xb55a4f09u) ^ (s >> 16);
    return s;
}

/* float -> uint32 conversion, saturating (negative/NaN -> 0, >= 2^32 -> max),
   which is what the GPU conversion instruction does. */
static inline uint32_t f2u_sat(float v) {
    if (!(v > 0.0f)) return 0u;
    if (v >= 4294967296.0f) return 0xffffffffu;
    return (uint32_t)v;
}

/* randomIntGrid :120-123, float arithmetic left to right */
static inline uint32_t lattice_hash(float x, float y, float z, float seed) {
    float a = x * 1723.0f;
    a = a + y * 93241.0f;
    a = a + z * 149812.0f;
    a = a + 3824.0f;
    a = a + seed;
    return noise_hash(f2u_sat(a));
}

/* grad :179-201 (cases 0xD and 0xF as in the source) */
template <class R> static inline R noise_grad(int h, R x, R y, R z) {
    switch (h & 0xF) {
    case 0x0: return x + y;
    case 0x1: return -x + y;
    case 0x2: return x - y;
    case 0x3: return -x - y;
    case 0x4: return x + z;
    case 0x5: return -x + z;
    case 0x6: return x - z;
    case 0x7: return -x - z;
    case 0x8: return y + z;
    case 0x9: return -y + z;
    case 0xA: return y - z;
    case 0xB: return -y - z;
    case 0xC: return y + x;
    case 0xD: return -y + z;
    case 0xE: return y - x;
    default: return -y - z;
    }
}

/* fade :204-207 */
template <class R> static inline R noise_fade(R t) {
    return t * t * t * (t * (t * R(6) - R(15)) + R(10));
}
/* lerp :167-170 */
template <class R> static inline R noise_lerp(R a, R b, R r) { return a * (R(1) - r) + b * r; }

/* perlinNoise :570-615 */
template <class R> static R perlin(V3<R> pos, R scale, int seed) {
    const float fseed = (float)seed;
    R px = pos.x * scale, py = pos.y * scale, pz = pos.z * scale;
    R ix = std::floor(px), iy = std::floor(py), iz = std::floor(pz);
    px -= ix;
    py -= iy;
    pz -= iz;
    R u = noise_fade(px), v = noise_fade(py), w = noise_fade(pz);
    /* lattice corners: ix + 1.0f etc. in float */
    float fx = (float)ix, fy = (float)iy, fz = (float)iz;
    float fx1 = fx + 1.0f, fy1 = fy + 1.0f, fz1 = fz + 1.0f;
    R i000 = noise_grad<R>((int)lattice_hash(fx, fy, fz, fseed), px, py, pz);
    R i100 = noise_grad<R>((int)lattice_hash(fx1, fy, fz, fseed), px - R(1), py, pz);
    R i010 = noise_grad<R>((int)lattice_hash(fx, fy1, fz, fseed), px, py - R(1), pz);
    R i110 = noise_grad<R>((int)lattice_hash(fx1, fy1, fz, fseed), px - R(1), py - R(1), pz);
    R i001 = noise_grad<R>((int)lattice_hash(fx, fy, fz1, fseed), px, py, pz - R(1));
    R i101 = noise_grad<R>((int)lattice_hash(fx1, fy, fz1, fseed), px - R(1), py, pz - R(1));
    R i011 = noise_grad<R>((int)lattice_hash(fx, fy1, fz1, fseed), px, py - R(1), pz - R(1));
    R i111 = noise_grad<R>((int)lattice_hash(fx1, fy1, fz1, fseed), px - R(1), py - R(1), pz - R(1));
    R x00 = noise_lerp(i000, i100, u);
    R x10 = noise_lerp(i010, i110, u);
    R x01 = noise_lerp(i001, i101, u);
    R x11 = noise_lerp(i011, i111, u);
    R y0 = noise_lerp(x00, x10, v);
    R y1 = noise_lerp(x01, x11, v);
    return noise_lerp(y0, y1, w);
}

/* ------------------------------------------------------------------ */
/* Staggered grid view, src/Grid.h:40-337                              */
/* ------------------------------------------------------------------ */
template <class R> struct Grid {
    int nx = 0, ny = 0, nz = 0;
    R ox = 0, oy = 0, oz = 0; /* origin offset in cells */
    R dx = 0;
    bool flip = false;
    V3<R> dmin{0, 0, 0}, dsize{0, 0, 0};
    std::vector<R> front, back;

    void init(int w, int h, int d, R ox_, R oy_, R oz_, R dx_, V3<R> dmin_, V3<R> dmax_, bool flip_) {
        nx = w; ny = h; nz = d;
        ox = ox_; oy = oy_; oz = oz_;
        dx = dx_;
        dmin = dmin_;
        dsize = mk<R>(dmax_.x - dmin_.x, dmax_.y - dmin_.y, dmax_.z - dmin_.z);
        flip = flip_;
        front.assign((size_t)w * h * d, R(0));
        back.assign((size_t)w * h * d, R(0));
    }
    size_t size() const { return (size_t)nx * ny * nz; }
    size_t idx(int i, int j, int k) const { return (size_t)i + (size_t)j * nx + (size_t)k * nx * ny; }
    R& at(int i, int j, int k) { return front[idx(i, j, k)]; }
    R& atB(int i, int j, int k) { return back[idx(i, j, k)]; }
    void swap() { front.swap(back); }

    /* :56-90 */
    R xo2g(R x) const { return x / dx - ox; }
    R yo2g(R y) const { return y / dx - oy; }
    R zo2g(R z) const { return z / dx - oz; }
    R xw2o(R x) const { return x - dmin.x; }
    R yw2o(R y) const { return dsize.y - (y - dmin.y); }
    R zw2o(R z) const { return z - dmin.z; }
    V3<R> w2o(R x, R y, R z) const { return mk<R>(x - dmin.x, dsize.y - (y - dmin.y), z - dmin.z); }
    /* :270-290 */
    R i2o(int i) const { return (R(i) + ox) * dx; }
    R j2o(int j) const { return (R(j) + oy) * dx; }
    R k2o(int k) const { return (R(k) + oz) * dx; }
    V3<R> g2o(int i, int j, int k) const { return mk<R>(i2o(i), j2o(j), k2o(k)); }

    /* :94-108 */
    R trilinear(V3<R> p) {
        const R gX = std::fmax(std::fmin(xo2g(p.x), R(nx) - ORC_K(R, 1.001)), R(0));
        const R gY = std::fmax(std::fmin(yo2g(p.y), R(ny) - ORC_K(R, 1.001)), R(0));
        const R gZ = std::fmax(std::fmin(zo2g(p.z), R(nz) - ORC_K(R, 1.001)), R(0));
        const int i = (int)gX, j = (int)gY, k = (int)gZ;
        const R tx = gX - R(i), ty = gY - R(j), tz = gZ - R(k);
        const R c00 = at(i, j, k) * (R(1) - tx) + at(i + 1, j, k) * tx;
        const R c10 = at(i, j + 1, k) * (R(1) - tx) + at(i + 1, j + 1, k) * tx;
        const R c01 = at(i, j, k + 1) * (R(1) - tx) + at(i + 1, j, k + 1) * tx;
        const R c11 = at(i, j + 1, k + 1) * (R(1) - tx) + at(i + 1, j + 1, k + 1) * tx;
        return (R(1) - tz) * ((R(1) - ty) * c00 + ty * c10) + tz * ((R(1) - ty) * c01 + ty * c11);
    }

    /* :148-163 cubic Lagrange on nodes -1,0,1,2 */
    static R cerp(R q0, R q1, R q2, R q3, R s) {
        const R w0 = (R(-1) / R(3)) * s + ORC_K(R, 0.5) * s * s - (R(1) / R(6)) * s * s * s;
        const R w1 = R(1) - s * s + ORC_K(R, 0.5) * (s * s * s - s);
        const R w2 = s + ORC_K(R, 0.5) * (s * s - s * s * s);
        const R w3 = (R(1) / R(6)) * (s * s * s - s);
        return w0 * q0 + w1 * q1 + w2 * q2 + w3 * q3;
    }
    static R cerpc(R q0, R q1, R q2, R q3, R s) { return std::fmax(cerp(q0, q1, q2, q3, s), R(0)); }

    /* :165-222, x first, then y, then z; clamp at every stage when asked */
    R tricubic(V3<R> p, bool clamp) {
        const R gX = std::fmax(std::fmin(xo2g(p.x), R(nx) - ORC_K(R, 1.001)), R(0));
        const R gY = std::fmax(std::fmin(yo2g(p.y), R(ny) - ORC_K(R, 1.001)), R(0));
        const R gZ = std::fmax(std::fmin(zo2g(p.z), R(nz) - ORC_K(R, 1.001)), R(0));
        const int i1 = (int)gX, j1 = (int)gY, k1 = (int)gZ;
        const int is[4] = {std::max(i1 - 1, 0), i1, i1 + 1, std::min(i1 + 2, nx - 1)};
        const int js[4] = {std::max(j1 - 1, 0), j1, j1 + 1, std::min(j1 + 2, ny - 1)};
        const int ks[4] = {std::max(k1 - 1, 0), k1, k1 + 1, std::min(k1 + 2, nz - 1)};
        const R sx = gX - R(i1), sy = gY - R(j1), sz = gZ - R(k1);
        R cz[4];
        for (int c = 0; c < 4; ++c) {
            R cy[4];
            for (int b = 0; b < 4; ++b) {
                const R q0 = at(is[0], js[b], ks[c]), q1 = at(is[1], js[b], ks[c]);
                const R q2 = at(is[2], js[b], ks[c]), q3 = at(is[3], js[b], ks[c]);
                cy[b] = clamp ? cerpc(q0, q1, q2, q3, sx) : cerp(q0, q1, q2, q3, sx);
            }
            cz[c] = clamp ? cerpc(cy[0], cy[1], cy[2], cy[3], sy) : cerp(cy[0], cy[1], cy[2], cy[3], sy);
        }
        return clamp ? cerpc(cz[0], cz[1], cz[2], cz[3], sz) : cerp(cz[0], cz[1], cz[2], cz[3], sz);
    }

    /* :309-313 */
    static R cubic_pulse(R x, R w) {
        if (x > w) return R(0);
        x /= w;
        return R(1) - x * x * (R(3) - R(2) * x);
    }

    /* :315-335, accumulating into back in call order */
    void sphere_to_grid(R px, R py, R pz, R pscale, R value) {
        const R radius = pscale / R(2);
        auto trunc = [](R v) -> int {
            /* (int) conversion; clamp only values far outside any grid */
            if (!(v == v)) v = -R(1 << 30); /* NaN: empty footprint, as on the GPU */
            if (v > R(1 << 30)) v = R(1 << 30);
            if (v < -R(1 << 30)) v = -R(1 << 30);
            return (int)v;
        };
        const int imin = std::max(trunc(xo2g(xw2o(px - radius))), 0);
        const int imax = std::min(trunc(xo2g(xw2o(px + radius))) + 1, nx);
        const int jmin = std::max(trunc(yo2g(yw2o(py + radius))), 0);
        const int jmax = std::min(trunc(yo2g(yw2o(py - radius))) + 1, ny);
        const int kmin = std::max(trunc(zo2g(zw2o(pz - radius))), 0);
        const int kmax = std::min(trunc(zo2g(zw2o(pz + radius))) + 1, nz);
        const V3<R> po = w2o(px, py, pz);
        for (int i = imin; i < imax; i++)
            for (int j = jmin; j < jmax; j++)
                for (int k = kmin; k < kmax; k++) {
                    const V3<R> g = g2o(i, j, k);
                    const R mask = cubic_pulse(len3(mk<R>(po.x - g.x, po.y - g.y, po.z - g.z)), radius);
                    if (mask > R(0)) atB(i, j, k) += flip ? -value * mask : value * mask;
                }
    }
};

/* fit, src/FluidSolver.cu:250-252 */
template <class R> static inline R fit(R v, R minIn, R maxIn, R minOut, R maxOut) {
    return std::fmax(std::fmin((v - minIn) / (maxIn - minIn), R(1)), R(0)) * (maxOut - minOut) + minOut;
}

static inline bool abs_gt_1e5(double v) { return std::fabs(v) > 1e-5; }

/* ------------------------------------------------------------------ */
/* Solver, src/FluidSolver.cu:671-955                                  */
/* ------------------------------------------------------------------ */
template <class R> struct Solver {
    blz_settings s{};
    int W = 0, H = 0, D = 0;
    R dx = 0;
    bool smoke = false;
    Grid<R> T, U, V, Wg, S; /* S = smoke density (extension) */
    std::vector<R> rhs, p;
    std::vector<float> part; /* uploaded particle records, 8 floats each */
    int npart = 0;
    bool interleaved_rbgs = false; /* reference per-slab order (pin test) */

    /* stats */
    int substeps = 0, iters_total = 0, iters_last = 0;
    double res = -1.0, res0 = -1.0;

    void init(const blz_settings& st) {
        s = st;
        W = s.grid_res[0]; H = s.grid_res[1]; D = s.grid_res[2];
        dx = (R)s.dx;
        const V3<R> dmin = mk<R>(s.domain_min[0], s.domain_min[1], s.domain_min[2]);
        const V3<R> dmax = mk<R>(s.domain_max[0], s.domain_max[1], s.domain_max[2]);
        /* create_grids_kernel, src/FluidSolver.cu:30-45 */
        T.init(W, H, D, ORC_K(R, 0.5), ORC_K(R, 0.5), ORC_K(R, 0.5), dx, dmin, dmax, false);
        U.init(W + 1, H, D, R(0), ORC_K(R, 0.5), ORC_K(R, 0.5), dx, dmin, dmax, false);
        V.init(W, H + 1, D, ORC_K(R, 0.5), R(0), ORC_K(R, 0.5), dx, dmin, dmax, true);
        Wg.init(W, H, D + 1, ORC_K(R, 0.5), ORC_K(R, 0.5), R(0), dx, dmin, dmax, false);
        smoke = s.smoke_density > 0.0f;
        if (smoke) S.init(W, H, D, ORC_K(R, 0.5), ORC_K(R, 0.5), ORC_K(R, 0.5), dx, dmin, dmax, false);
        rhs.assign((size_t)W * H * D, R(0));
        p.assign((size_t)W * H * D, R(0)); /* zeroed once, :715-716 */
    }

    Grid<R>* grid(int which) {
        switch (which) {
        case BLZ_FIELD_T: return &T;
        case BLZ_FIELD_U: return &U;
        case BLZ_FIELD_V: return &V;
        case BLZ_FIELD_W: return &Wg;
        case BLZ_FIELD_DENSITY: return smoke ? &S : nullptr;
        default: return nullptr;
        }
    }

    /* ---- addSource, :770-821 (clear back, splat, overwrite front) ---- */
    void add_source(const float* part8, int n) {
        if (part8) {
            part.assign(part8, part8 + (size_t)n * 8);
            npart = n;
        }
        if (npart <= 0) return;
        Grid<R>* gs[5] = {&T, &U, &V, &Wg, smoke ? &S : nullptr};
        /* each grid's cells receive contributions in ascending particle index */
#pragma omp parallel for schedule(static, 1)
        for (int g = 0; g < 5; ++g) {
            Grid<R>* G = gs[g];
            if (!G) continue;
            std::fill(G->back.begin(), G->back.end(), R(0));
            for (int q = 0; q < npart; ++q) {
                const float* r = &part[(size_t)q * 8];
                const R px = r[0], py = r[1], pz = r[2];
                const R ps = r[6];
                if (!(ps > dx / R(2))) continue; /* :83 */
                float val;
                switch (g) {
                case 0: val = r[7]; break;
                case 1: val = r[3]; break;
                case 2: val = r[4]; break;
                case 3: val = r[5]; break;
                default: val = s.smoke_density; break;
                }
                if (!abs_gt_1e5((double)(R)val)) continue; /* :84-95 */
                G->sphere_to_grid(px, py, pz, ps, (R)val);
            }
            /* set_source_from_back_to_front_kernel :99-116 */
            for (size_t c = 0; c < G->size(); ++c)
                if (abs_gt_1e5((double)G->back[c])) G->front[c] = G->back[c];
        }
    }

    /* ---- forces ---- */
    void cooldown(R dt) { /* :118-124 */
        const R f = R(1) - dt * (R)s.cooling_rate;
        for (auto& v : T.front) v *= f;
    }
    void drag(R dt) { /* :126-140 */
        const R f = R(1) - dt * (R)s.drag;
        for (auto& v : U.front) v *= f;
        for (auto& v : V.front) v *= f;
        for (auto& v : Wg.front) v *= f;
    }
    void buoyancy(R dt) { /* :142-156 */
        const R beta = s.buoyancy, g = s.gravity;
#pragma omp parallel for
        for (int k = 0; k < D; ++k)
            for (int j = 0; j < H + 1; ++j)
                for (int i = 0; i < W; ++i) {
                    const R temp = T.trilinear(V.g2o(i, j, k));
                    const R force = beta * temp * g;
                    V.at(i, j, k) += dt * force;
                }
    }

    /* computeVorticity :174-201 */
    V3<R> vort_at(int i, int j, int k) {
        const R x0 = T.i2o(i - 1), x1 = T.i2o(i), x2 = T.i2o(i + 1);
        const R y0 = T.j2o(j - 1), y1 = T.j2o(j), y2 = T.j2o(j + 1);
        const R z0 = T.k2o(k - 1), z1 = T.k2o(k), z2 = T.k2o(k + 1);
        const R w121 = Wg.trilinear(mk(x1, y2, z1)), w101 = Wg.trilinear(mk(x1, y0, z1));
        const R v112 = V.trilinear(mk(x1, y1, z2)), v110 = V.trilinear(mk(x1, y1, z0));
        const R u112 = U.trilinear(mk(x1, y1, z2)), u110 = U.trilinear(mk(x1, y1, z0));
        const R w211 = Wg.trilinear(mk(x2, y1, z1)), w011 = Wg.trilinear(mk(x0, y1, z1));
        const R v211 = V.trilinear(mk(x2, y1, z1)), v011 = V.trilinear(mk(x0, y1, z1));
        const R u121 = U.trilinear(mk(x1, y2, z1)), u101 = U.trilinear(mk(x1, y0, z1));
        const R t = R(2) * dx;
        return mk<R>((w121 - w101 - (v112 - v110)) / t, (u112 - u110 - (w211 - w011)) / t,
                     (v211 - v011 - (u121 - u101)) / t);
    }

    /* vorticity_confinement_kernel :203-248; per-cell face increments,
       gathered per face as (face + lower cell) + upper cell */
    void vorticity(R dt) {
        const R conf = s.vorticity_conf;
        const size_t n = (size_t)W * H * D;
        std::vector<R> cx(n, R(0)), cy(n, R(0)), cz(n, R(0));
        std::vector<unsigned char> act(n, 0);
        const R eps = ORC_K(R, 1e-20) * (R(1) / (dx * dt));
#pragma omp parallel for
        for (int k = 2; k < D - 2; ++k)
            for (int j = 2; j < H - 2; ++j)
                for (int i = 2; i < W - 2; ++i) {
                    V3<R> om = vort_at(i, j, k);
                    const V3<R> o211 = vort_at(i + 1, j, k), o011 = vort_at(i - 1, j, k);
                    const V3<R> o121 = vort_at(i, j + 1, k), o101 = vort_at(i, j - 1, k);
                    const V3<R> o112 = vort_at(i, j, k + 1), o110 = vort_at(i, j, k - 1);
                    const R t = R(2) * dx;
                    const V3<R> gn = mk<R>((len3(o211) - len3(o011)) / t, (len3(o121) - len3(o101)) / t,
                                           (len3(o112) - len3(o110)) / t);
                    /* N of :226 is computed by the reference but unused */
                    const V3<R> pos = T.g2o(i, j, k);
                    const V3<R> vel = mk<R>(U.trilinear(pos), V.trilinear(pos), Wg.trilinear(pos));
                    const R speed = len3(vel);
                    const R onorm = len3(om) + eps;
                    if (onorm > speed) om = mk<R>(speed * om.x / onorm, speed * om.y / onorm, speed * om.z / onorm);
                    const R cd = conf * dx;
                    const V3<R> c = cross3(gn, om);
                    const V3<R> f = mk<R>(cd * c.x, cd * c.y, cd * c.z);
                    const R h = ORC_K(R, 0.5) * dt;
                    const size_t id = T.idx(i, j, k);
                    cx[id] = h * f.x;
                    cy[id] = h * f.y;
                    cz[id] = h * f.z;
                    act[id] = 1;
                }
        face_gather(cx, cy, cz, act);
    }

    /* (face + lower cell) + upper cell, for cells flagged in act */
    void face_gather(const std::vector<R>& cx, const std::vector<R>& cy, const std::vector<R>& cz,
                     const std::vector<unsigned char>& act) {
#pragma omp parallel for
        for (int k = 0; k < D; ++k)
            for (int j = 0; j < H; ++j)
                for (int i = 0; i < W + 1; ++i) {
                    R v = U.at(i, j, k);
                    if (i >= 1 && act[T.idx(i - 1, j, k)]) v = v + cx[T.idx(i - 1, j, k)];
                    if (i < W && act[T.idx(i, j, k)]) v = v + cx[T.idx(i, j, k)];
                    U.at(i, j, k) = v;
                }
#pragma omp parallel for
        for (int k = 0; k < D; ++k)
            for (int j = 0; j < H + 1; ++j)
                for (int i = 0; i < W; ++i) {
                    R v = V.at(i, j, k);
                    if (j >= 1 && act[T.idx(i, j - 1, k)]) v = v + cy[T.idx(i, j - 1, k)];
                    if (j < H && act[T.idx(i, j, k)]) v = v + cy[T.idx(i, j, k)];
                    V.at(i, j, k) = v;
                }
#pragma omp parallel for
        for (int k = 0; k < D + 1; ++k)
            for (int j = 0; j < H; ++j)
                for (int i = 0; i < W; ++i) {
                    R v = Wg.at(i, j, k);
                    if (k >= 1 && act[T.idx(i, j, k - 1)]) v = v + cz[T.idx(i, j, k - 1)];
                    if (k < D && act[T.idx(i, j, k)]) v = v + cz[T.idx(i, j, k)];
                    Wg.at(i, j, k) = v;
                }
    }

    /* add_wind_kernel :337-390 */
    void wind(R time, R dt) {
        const bool xwind = s.wind_dir <= 1;
        const R freq = R(1) / (R)s.wind_turb_scale;
        const R avgDom = (U.dsize.x + U.dsize.y + U.dsize.z) / R(3);
        const R offset = std::fmod(time * avgDom * (R)s.wind_speed, ORC_K(R, 25000.0));
        const R amp = s.wind_amp, tamp = s.wind_turb_amp;
        const R dirv = xwind ? (s.wind_dir == 0 ? R(1) : R(-1)) : (s.wind_dir == 2 ? R(1) : R(-1));
#pragma omp parallel for
        for (int k = 0; k < D + 1; ++k)
            for (int j = 0; j < H + 1; ++j)
                for (int i = 0; i < W + 1; ++i) {
                    if (xwind) {
                        if (i >= 2 && i < (W + 1) - 2) continue;
                    } else {
                        if (k >= 2 && k < (D + 1) - 2) continue;
                    }
                    const V3<R> pos = mk<R>(U.i2o(i), V.j2o(j), Wg.k2o(k) + offset);
                    const R nU = tamp * perlin<R>(pos, freq, 123);
                    const R nV = tamp * perlin<R>(pos, freq, 456);
                    const R nW = tamp * perlin<R>(pos, freq, 789);
                    if (j < H && k < D) {
                        R& u = U.at(i, j, k);
                        u = u * (R(1) - dt) + (amp * (xwind ? (dirv + nU) : nU)) * dt;
                    }
                    if (i < W && k < D) {
                        R& v = V.at(i, j, k);
                        v = v * (R(1) - dt) + (amp * nV) * dt;
                    }
                    if (i < W && j < H) {
                        R& w = Wg.at(i, j, k);
                        w = w * (R(1) - dt) + (amp * (xwind ? nW : (dirv + nW))) * dt;
                    }
                }
    }

    /* compute_turbulence_kernel :295-335 (noise into back buffers) */
    void turb_face(Grid<R>& G, int axis, R offset, R freq, R amp) {
        const R tr0 = s.turb_mask_temp_ramp[0], tr1 = s.turb_mask_temp_ramp[1];
        const R vr0 = s.turb_mask_vel_ramp[0], vr1 = s.turb_mask_vel_ramp[1];
        const int seed = s.noise_seed;
#pragma omp parallel for
        for (int k = 0; k < G.nz; ++k)
            for (int j = 0; j < G.ny; ++j)
                for (int i = 0; i < G.nx; ++i) {
                    const V3<R> pos = G.g2o(i, j, k);
                    const R tm = fit<R>(T.trilinear(pos), tr0, tr1, R(1), R(0));
                    const V3<R> vel = mk<R>(U.trilinear(pos), V.trilinear(pos), Wg.trilinear(pos));
                    const R vm = fit<R>(len3(vel), vr0, vr1, R(0), R(1));
                    const R mask = tm * vm;
                    V3<R> np = mk<R>((R)i, (R)j, (R)k);
                    if (axis == 0) np.x = (R)i + offset;
                    if (axis == 1) np.y = (R)j + offset;
                    if (axis == 2) np.z = (R)k + offset;
                    const R nz = perlin<R>(np, freq, seed);
                    G.atB(i, j, k) = mask * amp * nz;
                }
    }

    /* compute_turbulence_kernel + add_curl_noise_kernel :271-335 */
    void turbulence(R time, R dt) {
        turb_noise(time);
        curl_noise(dt);
    }

    /* compute_turbulence_kernel :295-335 for the three face grids */
    void turb_noise(R time) {
        const R freq = R(1) / (R)s.turbulence_scale;
        const R avgGrid = R(W + H + D) / R(3);
        const R offset = std::fmod(time * avgGrid, ORC_K(R, 25000.0));
        const R amp = s.turbulence_amp;
        turb_face(U, 0, offset, freq, amp);
        turb_face(V, 1, offset, freq, amp);
        turb_face(Wg, 2, offset, freq, amp);
    }

    /* add_curl_noise_kernel :271-293: pseudo-curl of the back buffers (with
       the reference's V-for-U/W reads kept), added to the front faces */
    void curl_noise(R dt) {
        const size_t n = (size_t)W * H * D;
        std::vector<R> cx(n), cy(n), cz(n);
        std::vector<unsigned char> act(n, 1);
        const R eps = ORC_K(R, 1e-20) * (R(1) / (dx * dt));
#pragma omp parallel for
        for (int k = 0; k < D; ++k)
            for (int j = 0; j < H; ++j)
                for (int i = 0; i < W; ++i) {
                    const R vb = V.atB(i, j, k);
                    const R dudx = (U.atB(i + 1, j, k) - vb) / dx;
                    const R dvdy = (V.atB(i, j + 1, k) - vb) / dx;
                    const R dwdz = (Wg.atB(i, j, k + 1) - vb) / dx;
                    V3<R> c = mk<R>(dvdy - dwdz, dwdz - dudx, dudx - dvdy);
                    const V3<R> a = mk<R>((U.atB(i, j, k) + U.atB(i + 1, j, k)) / R(2),
                                          (V.atB(i, j, k) + V.atB(i, j + 1, k)) / R(2),
                                          (Wg.atB(i, j, k) + Wg.atB(i, j, k + 1)) / R(2));
                    const R avgSpeed = len3(a);
                    const R den = len3(c) + eps;
                    c = mk<R>(avgSpeed * c.x / den, avgSpeed * c.y / den, avgSpeed * c.z / den);
                    const R h = ORC_K(R, 0.5) * dt;
                    const size_t id = T.idx(i, j, k);
                    cx[id] = h * c.x;
                    cy[id] = h * c.y;
                    cz[id] = h * c.z;
                }
        face_gather(cx, cy, cz, act);
    }

    /* ---- projection ---- */
    void divergence() { /* :475-515 */
        const R sc = R(1) / dx;
        const int* cb = s.closed_bounds;
#pragma omp parallel for
        for (int k = 0; k < D; ++k)
            for (int j = 0; j < H; ++j)
                for (int i = 0; i < W; ++i) {
                    const size_t id = T.idx(i, j, k);
                    R r = -sc * (U.at(i + 1, j, k) - U.at(i, j, k) + V.at(i, j + 1, k) - V.at(i, j, k) +
                                 Wg.at(i, j, k + 1) - Wg.at(i, j, k));
                    if (i == 0) {
                        if (cb[0]) r -= sc * U.at(i, j, k);
                    } else if (i == W - 1) {
                        if (cb[1]) r += sc * U.at(i + 1, j, k);
                    }
                    if (j == 0) {
                        if (cb[3]) r -= sc * V.at(i, j, k);
                    } else if (j == H - 1) {
                        if (cb[2]) r += sc * V.at(i, j + 1, k);
                    }
                    if (k == 0) {
                        if (cb[4]) r -= sc * Wg.at(i, j, k);
                    } else if (k == D - 1) {
                        if (cb[5]) r += sc * Wg.at(i, j, k + 1);
                    }
                    rhs[id] = r;
                }
    }

    /* one Gauss-Seidel update, :533-561 */
    inline void gs_cell(int i, int j, int k, R scale) {
        const size_t WH = (size_t)W * H;
        const size_t id = (size_t)i + (size_t)j * W + (size_t)k * WH;
        R denom = R(0);
        R num = scale * rhs[id];
        if (i > 0) { num += p[id - 1]; denom += R(1); }
        if (i < W - 1) { num += p[id + 1]; denom += R(1); }
        if (j > 0) { num += p[id - W]; denom += R(1); }
        if (j < H - 1) { num += p[id + W]; denom += R(1); }
        if (k > 0) { num += p[id - WH]; denom += R(1); }
        if (k < D - 1) { num += p[id + WH]; denom += R(1); }
        p[id] = num / denom;
    }

    void rbgs_sweep(R scale) {
        if (interleaved_rbgs) {
            /* reference order: per slab, red then black */
            for (int k = 0; k < D; ++k)
                for (int c = 0; c < 2; ++c)
                    for (int j = 0; j < H; ++j)
                        for (int i = 0; i < W; ++i)
                            if (((i + j) & 1) == c) gs_cell(i, j, k, scale);
            return;
        }
        for (int c = 0; c < 2; ++c) {
#pragma omp parallel for
            for (int j = 0; j < H; ++j)
                for (int i = (j + c) & 1; i < W; i += 2)
                    for (int k = 0; k < D; ++k) gs_cell(i, j, k, scale);
        }
    }

    /* ||div u||_inf proxy after the update: max over cells not adjacent to an
       open face of |rhs + c (sum_nbr p - n p)|, c = dt/(rho dx^2), in FP64 */
    double residual(R dt) {
        const double c = (double)dt / ((double)(R)s.density * (double)dx * (double)dx);
        const int* cb = s.closed_bounds;
        const size_t WH = (size_t)W * H;
        double m = 0.0;
#pragma omp parallel for reduction(max : m)
        for (int k = 0; k < D; ++k)
            for (int j = 0; j < H; ++j)
                for (int i = 0; i < W; ++i) {
                    if ((i == 0 && !cb[0]) || (i == W - 1 && !cb[1]) || (j == 0 && !cb[3]) ||
                        (j == H - 1 && !cb[2]) || (k == 0 && !cb[4]) || (k == D - 1 && !cb[5]))
                        continue;
                    const size_t id = (size_t)i + (size_t)j * W + (size_t)k * WH;
                    double sum = 0.0, n = 0.0;
                    if (i > 0) { sum += (double)p[id - 1]; n += 1.0; }
                    if (i < W - 1) { sum += (double)p[id + 1]; n += 1.0; }
                    if (j > 0) { sum += (double)p[id - W]; n += 1.0; }
                    if (j < H - 1) { sum += (double)p[id + W]; n += 1.0; }
                    if (k > 0) { sum += (double)p[id - WH]; n += 1.0; }
                    if (k < D - 1) { sum += (double)p[id + WH]; n += 1.0; }
                    double r = std::fabs((double)rhs[id] + c * (sum - n * (double)p[id]));
                    if (r != r) r = INFINITY; /* NaN counts as +inf */
                    m = std::max(m, r);
                }
        return m;
    }

    void rbgs(R dt) { /* gs_solve_kernel :517-599 */
        const R scale = ((R)s.density * dx * dx) / dt;
        const int cap = s.max_iter_solve;
        if (!(s.residual_tol > 0.0f)) {
            for (int l = 0; l < cap; ++l) rbgs_sweep(scale);
            iters_last = cap;
            res = -1.0;
            res0 = -1.0;
            return;
        }
        const int every = s.residual_check_every > 0 ? s.residual_check_every : 10;
        const double tol = (double)s.residual_tol;
        const bool adaptive = s.residual_check_adaptive != 0;
        int it = 0, gap = every, stall_it = 0;
        double r = residual(dt);
        res0 = r;
        while (r > tol && it < cap) {
            const int n = std::min(gap, cap - it);
            for (int l = 0; l < n; ++l) rbgs_sweep(scale);
            it += n;
            const double prev = r;
            r = residual(dt);
            if (!adaptive) continue;
            /* an FP32 fixed point: the same residual at every check over
               STALL_SWEEPS sweeps ends the solve */
            if (r == prev) {
                if (it - stall_it >= STALL_SWEEPS) break;
            } else {
                stall_it = it;
            }
            gap = next_gap(gap, every, r, prev, tol);
        }
        iters_last = it;
        res = r;
    }

    /* Interval to the next residual check of an adaptive solve, from the
       contraction over the last interval, rho = r / prev: a quarter of it
       when the tolerance is predicted within one more interval, half within
       two, the same within four, else double (at most GAP_MAX_CHECKS
       checks' worth); always a multiple of `every`. */
    static int next_gap(int gap, int every, double r, double prev, double tol) {
        const double rho = r / prev;
        if (rho < 1.0 && r * rho <= tol) return std::max(every, gap / 4 / every * every);
        if (rho < 1.0 && r * rho * rho <= tol) return std::max(every, gap / 2 / every * every);
        if (rho < 1.0 && r * rho * rho * rho * rho <= tol) return gap;
        return std::min(2 * gap, GAP_MAX_CHECKS * every);
    }
    static constexpr int STALL_SWEEPS = 60;
    static constexpr int GAP_MAX_CHECKS = 64;

    void pressure_update(R dt) { /* :601-669 */
        const R sc = dt / ((R)s.density * dx);
        const size_t WH = (size_t)W * H;
        const int* cb = s.closed_bounds;
#pragma omp parallel for
        for (int k = 0; k < D; ++k)
            for (int j = 0; j < H; ++j)
                for (int i = 0; i < W; ++i) {
                    const size_t id = (size_t)i + (size_t)j * W + (size_t)k * WH;
                    if (i > 0) U.at(i, j, k) -= sc * (p[id] - p[id - 1]);
                    if (j > 0) V.at(i, j, k) -= sc * (p[id] - p[id - W]);
                    if (k > 0) Wg.at(i, j, k) -= sc * (p[id] - p[id - WH]);
                }
        /* boundary faces after all interior faces */
        for (int k = 0; k < D; ++k)
            for (int j = 0; j < H; ++j) {
                U.at(0, j, k) = cb[0] ? R(0) : U.at(1, j, k);
                U.at(W, j, k) = cb[1] ? R(0) : U.at(W - 1, j, k);
            }
        for (int k = 0; k < D; ++k)
            for (int i = 0; i < W; ++i) {
                V.at(i, 0, k) = cb[3] ? R(0) : V.at(i, 1, k);
                V.at(i, H, k) = cb[2] ? R(0) : V.at(i, H - 1, k);
            }
        for (int j = 0; j < H; ++j)
            for (int i = 0; i < W; ++i) {
                Wg.at(i, j, 0) = cb[4] ? R(0) : Wg.at(i, j, 1);
                Wg.at(i, j, D) = cb[5] ? R(0) : Wg.at(i, j, D - 1);
            }
    }

    /* advect_RK3_kernel :431-456 */
    void advect_grid(Grid<R>& G, R dt, bool clamp) {
        const R a = ORC_K(R, 0.5) * dt, b = ORC_K(R, 0.75) * dt;
        const R c1 = R(2) / R(9), c2 = R(3) / R(9), c3 = R(4) / R(9);
#pragma omp parallel for
        for (int k = 0; k < G.nz; ++k)
            for (int j = 0; j < G.ny; ++j)
                for (int i = 0; i < G.nx; ++i) {
                    V3<R> pos = G.g2o(i, j, k);
                    const V3<R> k1 = mk<R>(U.trilinear(pos), V.trilinear(pos), Wg.trilinear(pos));
                    const V3<R> p1 = mk<R>(pos.x - a * k1.x, pos.y - a * k1.y, pos.z - a * k1.z);
                    const V3<R> k2 = mk<R>(U.trilinear(p1), V.trilinear(p1), Wg.trilinear(p1));
                    const V3<R> p2 = mk<R>(pos.x - b * k2.x, pos.y - b * k2.y, pos.z - b * k2.z);
                    const V3<R> k3 = mk<R>(U.trilinear(p2), V.trilinear(p2), Wg.trilinear(p2));
                    pos.x -= dt * (c1 * k1.x + c2 * k2.x + c3 * k3.x);
                    pos.y -= dt * (c1 * k1.y + c2 * k2.y + c3 * k3.y);
                    pos.z -= dt * (c1 * k1.z + c2 * k2.z + c3 * k3.z);
                    G.atB(i, j, k) = G.tricubic(pos, clamp);
                }
    }

    void advect(R dt) { /* :942-953 */
        advect_grid(T, dt, true);
        advect_grid(U, dt, false);
        advect_grid(V, dt, false);
        advect_grid(Wg, dt, false);
        if (smoke) advect_grid(S, dt, true);
        T.swap(); U.swap(); V.swap(); Wg.swap();
        if (smoke) S.swap();
    }

    R max_abs_velocity() {
        R m = R(0);
        auto upd = [&m](R v) {
            R a = std::fabs(v);
            if (a != a) a = (R)INFINITY; /* NaN counts as +inf */
            m = std::max(m, a);
        };
        for (auto v : U.front) upd(v);
        for (auto v : V.front) upd(v);
        for (auto v : Wg.front) upd(v);
        return m;
    }

    void run_stage(int st, R time, R dt) {
        switch (st) {
        case BLZ_STAGE_COOLDOWN: cooldown(dt); break;
        case BLZ_STAGE_DRAG: drag(dt); break;
        case BLZ_STAGE_VORTICITY: vorticity(dt); break;
        case BLZ_STAGE_BUOYANCY: buoyancy(dt); break;
        case BLZ_STAGE_WIND: wind(time, dt); break;
        case BLZ_STAGE_TURBULENCE: turbulence(time, dt); break;
        case BLZ_STAGE_DIVERGENCE: divergence(); break;
        case BLZ_STAGE_PRESSURE: rbgs(dt); break;
        case BLZ_STAGE_PROJECT_UPDATE: pressure_update(dt); break;
        case BLZ_STAGE_ADVECT: advect(dt); break;
        default: break;
        }
    }

    /* FluidSolver::step :858-955, gates :868-925 */
    void substep(R time, R dt) {
        if (std::fabs((R)s.cooling_rate) > R(0)) cooldown(dt);
        if ((R)s.drag > R(0)) drag(dt);
        if (std::fabs((R)s.vorticity_conf) > R(0)) vorticity(dt);
        if (std::fabs((R)s.buoyancy) > R(0)) buoyancy(dt);
        if ((R)s.wind_amp > R(0)) wind(time, dt);
        if ((R)s.turbulence_amp > R(0)) turbulence(time, dt);
        divergence();
        rbgs(dt);
        iters_total += iters_last;
        pressure_update(dt);
        advect(dt);
    }

    /* one frame; with cfl_number > 0, n = ceil(max|u| dt / (cfl dx)) substeps
       of dt/n at times time + s*(dt/n) */
    void step(float time) {
        const R dt = (R)s.dt;
        int n = 1;
        if (s.cfl_number > 0.0f) {
            const R umax = max_abs_velocity();
            const R q = std::ceil(umax * dt / ((R)s.cfl_number * dx));
            const int cap = s.max_substeps > 0 ? s.max_substeps : 8;
            n = !(q >= R(1)) ? 1 : (q > R(cap) ? cap : (int)q);
        }
        const R dts = dt / R(n);
        iters_total = 0;
        for (int q = 0; q < n; ++q) substep((R)time + R(q) * dts, dts);
        substeps = n;
    }
};

} // namespace orc
