// render_oracle.hpp — CPU restatement of the renderer (TEST INFRASTRUCTURE
// ONLY). Follows RenderEngine::render (asiroisvigneux/Blaze
// src/RenderEngine.cu:312-365) and its kernels: write_scatter_to_3d_tex
// :76-101, gaussian_blur_texture :51-74, render_kernel :149-225,
// bbox_intersect :112-147; Camera src/Camera.h:42-58, Light
// src/Light.h:43-55, Shader src/Shader.h:52-72, ColorRamp
// src/ColorRamp.h:47-78, helper_math lerp/smoothstep/clamp
// (src/utils/helper_math.h:1155-1175, 1475-1478). Same evaluation order as
// the HIP kernels, FP32 with no contraction (image rows in parallel, each
// pixel serial); the texture lookup is
// FP32 trilinear with clamped texel indices, normalize divides by sqrtf and
// exp is the Cody-Waite + degree-7 Taylor function below (pinned against
// math.exp by the tests).
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "../include/blaze.h"

namespace rnd {

struct V3 {
    float x, y, z;
};
inline V3 mk(float x, float y, float z) { return V3{x, y, z}; }
inline V3 add(V3 a, V3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
inline V3 sub(V3 a, V3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
inline V3 mul(V3 a, V3 b) { return mk(a.x * b.x, a.y * b.y, a.z * b.z); }
inline V3 muls(V3 a, float s) { return mk(a.x * s, a.y * s, a.z * s); }
inline V3 smul(float s, V3 a) { return mk(s * a.x, s * a.y, s * a.z); }
inline float sqlen(V3 v) { return v.x * v.x + v.y * v.y + v.z * v.z; }
inline V3 v3(const float* a) { return mk(a[0], a[1], a[2]); }

inline float bits_f(uint32_t u) {
    float f;
    std::memcpy(&f, &u, 4);
    return f;
}

// exp: x = k ln2 + r, ln2 = hi + lo with k*hi exact; Taylor degree 7 on r;
// 2^k as two exact powers of two; 0 below -87.3365479, inf above 88.7228394
inline float bexp(float x) {
    if (x != x) return x;
    if (x > 88.7228394f) return bits_f(0x7f800000u);
    if (x < -87.3365479f) return 0.0f;
    const float k = std::floor(x * 1.44269502f + 0.5f);
    float r = x - k * 0.693145751953125f;
    r = r - k * 1.42860677e-06f;
    float p = 1.98412701e-04f;
    p = p * r + 1.38888892e-03f;
    p = p * r + 8.33333377e-03f;
    p = p * r + 4.16666679e-02f;
    p = p * r + 1.66666672e-01f;
    p = p * r + 0.5f;
    p = p * r + 1.0f;
    p = p * r + 1.0f;
    const int ki = (int)k;
    const int k1 = ki / 2, k2 = ki - k1;
    return p * bits_f((uint32_t)(k1 + 127) << 23) * bits_f((uint32_t)(k2 + 127) << 23);
}
inline V3 bexp3(V3 a) { return mk(bexp(a.x), bexp(a.y), bexp(a.z)); }

// clamp-addressed linear filter at unnormalised coordinates
inline float tex(const float* v, int nx, int ny, int nz, float x, float y, float z) {
    const float xb = x - 0.5f, yb = y - 0.5f, zb = z - 0.5f;
    const float fx = std::floor(xb), fy = std::floor(yb), fz = std::floor(zb);
    const float a = xb - fx, b = yb - fy, c = zb - fz;
    auto clampf = [](float f, float lo, float hi) { return std::fmin(std::fmax(f, lo), hi); };
    const int i0 = (int)clampf(fx, -1.0f, (float)nx), j0 = (int)clampf(fy, -1.0f, (float)ny),
              k0 = (int)clampf(fz, -1.0f, (float)nz);
    auto ci = [](int q, int n) { return q < 0 ? 0 : (q > n - 1 ? n - 1 : q); };
    const int xa = ci(i0, nx), xc = ci(i0 + 1, nx), ya = ci(j0, ny), yc = ci(j0 + 1, ny), za = ci(k0, nz),
              zc = ci(k0 + 1, nz);
    auto t = [&](int i, int j, int k) { return v[(size_t)i + (size_t)j * nx + (size_t)k * nx * ny]; };
    const float c00 = (1.0f - a) * t(xa, ya, za) + a * t(xc, ya, za);
    const float c10 = (1.0f - a) * t(xa, yc, za) + a * t(xc, yc, za);
    const float c01 = (1.0f - a) * t(xa, ya, zc) + a * t(xc, ya, zc);
    const float c11 = (1.0f - a) * t(xa, yc, zc) + a * t(xc, yc, zc);
    return (1.0f - c) * ((1.0f - b) * c00 + b * c10) + c * ((1.0f - b) * c01 + b * c11);
}

struct Ray {
    V3 o, d, inv;
    int s[3];
    float t0, t1;
};
inline Ray ray(V3 o, V3 d) {
    Ray r;
    r.o = o;
    r.d = d;
    r.inv = mk(1.0f / d.x, 1.0f / d.y, 1.0f / d.z);
    r.s[0] = r.inv.x < 0.0f;
    r.s[1] = r.inv.y < 0.0f;
    r.s[2] = r.inv.z < 0.0f;
    r.t0 = 0.0f;
    r.t1 = bits_f(0x7f800000u);
    return r;
}
inline bool hit(Ray& r, V3 lo, V3 hi) {
    const V3 b[2] = {lo, hi};
    float tmin = (b[r.s[0]].x - r.o.x) * r.inv.x;
    float tmax = (b[1 - r.s[0]].x - r.o.x) * r.inv.x;
    const float tymin = (b[r.s[1]].y - r.o.y) * r.inv.y;
    const float tymax = (b[1 - r.s[1]].y - r.o.y) * r.inv.y;
    if ((tmin > tymax) || (tymin > tmax)) return false;
    if (tymin > tmin) tmin = tymin;
    if (tymax < tmax) tmax = tymax;
    const float tzmin = (b[r.s[2]].z - r.o.z) * r.inv.z;
    const float tzmax = (b[1 - r.s[2]].z - r.o.z) * r.inv.z;
    if ((tmin > tzmax) || (tzmin > tmax)) return false;
    if (tzmin > tmin) tmin = tzmin;
    if (tzmax < tmax) tmax = tzmax;
    r.t0 = std::fmax(tmin, 0.0f);
    r.t1 = tmax;
    return !(r.t1 < 0.0f);
}

// rows row0, row0 + stride, ... below row1 of the image; T is W x H x D (x
// fastest), domain and dx as the solver's
inline void render(const float* T, int W, int H, int D, float dx, const float* dmin, const float* dmax,
                   const blz_render_settings& r, float* rgb, int row0, int row1, int stride = 1) {
    const int MAXM = 1 << 16;
    // scatter volume: 4x downsample of T above temp_range[0], then blur passes
    const int Ws = W / 4, Hs = H / 4, Ds = D / 4;
    const bool have = Ws > 0 && Hs > 0 && Ds > 0;
    const int sw = have ? Ws : 1, sh = have ? Hs : 1, sd = have ? Ds : 1;
    std::vector<float> sc((size_t)sw * sh * sd, 0.0f), tmp(sc.size(), 0.0f);
    if (r.ms_scale > 0.0f && have) {
        for (int k = 0; k < Ds; ++k)
            for (int j = 0; j < Hs; ++j)
                for (int i = 0; i < Ws; ++i) {
                    float acc = 0.0f;
                    for (int io = 0; io < 4; ++io)
                        for (int jo = 0; jo < 4; ++jo)
                            for (int ko = 0; ko < 4; ++ko) {
                                const float v =
                                    T[(size_t)(i * 4 + io) + (size_t)(j * 4 + jo) * W + (size_t)(k * 4 + ko) * W * H];
                                acc += v > r.temp_range[0] ? v : 0.0f;
                            }
                    sc[(size_t)i + (size_t)j * Ws + (size_t)k * Ws * Hs] = acc / 64.0f;
                }
        static const float filt[27] = {0, 1, 0, 1, 2, 1, 0, 1, 0, 1, 2, 1, 2, 4, 2, 1, 2, 1, 0, 1, 0, 1, 2, 1, 0, 1, 0};
        float sum = 0.0f;
        for (float f : filt) sum += f;
        float w[27];
        for (int q = 0; q < 27; ++q) w[q] = filt[q] / sum;
        for (int it = 0; it < r.ms_blur_iter; ++it) {
            for (int k = 0; k < Ds; ++k)
                for (int j = 0; j < Hs; ++j)
                    for (int i = 0; i < Ws; ++i) {
                        float acc = 0.0f;
                        int q = 0;
                        for (int dk = -1; dk <= 1; ++dk)
                            for (int dj = -1; dj <= 1; ++dj)
                                for (int di = -1; di <= 1; ++di, ++q) {
                                    const int x = std::min(std::max(i + di, 0), Ws - 1);
                                    const int y = std::min(std::max(j + dj, 0), Hs - 1);
                                    const int z = std::min(std::max(k + dk, 0), Ds - 1);
                                    acc += sc[(size_t)x + (size_t)y * Ws + (size_t)z * Ws * Hs] * w[q];
                                }
                        tmp[(size_t)i + (size_t)j * Ws + (size_t)k * Ws * Hs] = acc;
                    }
            sc.swap(tmp);
        }
    }
    // camera
    const V3 o = v3(r.cam_trans), cu = v3(r.cam_u), cv = v3(r.cam_v), cw = v3(r.cam_w);
    const float aspect = (float)r.width / (float)r.height;
    const float hw = 0.5f * r.cam_aperture, hh = hw / aspect;
    const V3 tl = add(sub(sub(o, smul(r.cam_focal, cw)), smul(hw, cu)), smul(hh, cv));
    const V3 hor = smul(2.0f * hw, cu), ver = smul(-2.0f * hh, cv);
    const float rx = (float)(unsigned)r.width, ry = (float)(unsigned)r.height;
    // shader, lights
    const V3 col = v3(r.volume_color), sca = v3(r.scattering);
    const V3 den = mk(sca.x + r.absorption, sca.y + r.absorption, sca.z + r.absorption);
    const V3 albedo = mk((col.x * sca.x) / den.x, (col.y * sca.y) / den.y, (col.z * sca.z) / den.z);
    const V3 ext = muls(mk(sca.x + r.absorption, sca.y + r.absorption, sca.z + r.absorption), -1.0f);
    V3 ldir[BLZ_MAX_LIGHTS], lill[BLZ_MAX_LIGHTS];
    for (int l = 0; l < r.light_count; ++l) {
        ldir[l] = v3(r.light_dir[l]);
        lill[l] = muls(muls(v3(r.light_color[l]), r.light_intensity[l]), std::pow(2.0f, r.light_exposure[l]));
    }
    const V3 lo = v3(dmin), hi = v3(dmax), size = sub(hi, lo);
    auto w2g = [&](V3 p) {
        return mk((p.x - lo.x) / dx, (size.y - (p.y - lo.y)) / dx, (p.z - lo.z) / dx);
    };
    auto ramp = [&](float temp) {
        const float remap = std::fmin(std::fmax((temp - r.temp_range[0]) / (r.temp_range[1] - r.temp_range[0]), 0.0f),
                                      1.0f);
        if (r.ramp_size == 1) return v3(r.ramp_colors[0]);
        int prev = 0, next = 0;
        for (int i = 0; i < r.ramp_size; ++i) {
            if (r.ramp_keys[i] <= remap) {
                prev = i;
            } else {
                next = i;
                break;
            }
        }
        const float range = r.ramp_keys[next] - r.ramp_keys[prev];
        const float blend = (remap - r.ramp_keys[prev]) / range;
        return add(v3(r.ramp_colors[prev]), smul(blend, sub(v3(r.ramp_colors[next]), v3(r.ramp_colors[prev]))));
    };
    auto smooth = [](float a, float b, float x) {
        const float y = std::fmax(0.0f, std::fmin((x - a) / (b - a), 1.0f));
        return y * y * (3.0f - (2.0f * y));
    };
    const V3 one = mk(1.0f, 1.0f, 1.0f);
    // rows are independent: any thread count gives the same image
    const int nrows = stride > 0 ? (row1 - row0 + stride - 1) / stride : 0;
#pragma omp parallel for schedule(dynamic, 1)
    for (int jr = 0; jr < nrows; ++jr)
        for (int i = 0; i < r.width; ++i) {
            const int j = row0 + jr * stride;
            float* px = rgb + 3 * ((size_t)i + (size_t)r.width * j);
            px[0] = px[1] = px[2] = 0.0f;
            const V3 q = sub(add(add(tl, smul(((float)i + 0.5f) / rx, hor)), smul(((float)j + 0.5f) / ry, ver)), o);
            const float il = 1.0f / std::sqrt(q.x * q.x + q.y * q.y + q.z * q.z);
            Ray pr = ray(o, muls(q, il));
            if (!hit(pr, lo, hi)) continue;
            V3 lumi = mk(0.0f, 0.0f, 0.0f), trans = one;
            int n = 0;
            for (float pt = pr.t0 + dx * 0.5f; pt <= pr.t1 && n < MAXM; pt += r.p_step, ++n) {
                const V3 pos = add(pr.o, smul(pt, pr.d));
                const V3 g = w2g(pos);
                const float temp = tex(T, W, H, D, g.x, g.y, g.z);
                const float scat = tex(sc.data(), sw, sh, sd, g.x / 4.0f, g.y / 4.0f, g.z / 4.0f);
                const float dens = temp * r.density_scale;
                if (dens < r.cutoff) continue;
                V3 em = mk(0.0f, 0.0f, 0.0f);
                if (r.emission_scale > 0.0f)
                    em = add(muls(ramp(temp), r.emission_scale),
                             muls(muls(muls(v3(r.ms_color), scat), r.ms_scale), smooth(r.ms_mask[0], r.ms_mask[1], temp)));
                const V3 dT = bexp3(muls(muls(ext, dens), r.p_step));
                for (int l = 0; l < r.light_count; ++l) {
                    Ray sr = ray(pos, smul(-1.0f, ldir[l]));
                    V3 st = one;
                    if (hit(sr, lo, hi)) {
                        int m = 0;
                        for (float s = sr.t0 + dx * 0.5f; s <= sr.t1 && m < MAXM; s += r.s_step, ++m) {
                            const V3 sg = w2g(add(sr.o, smul(s, sr.d)));
                            const float sdn = tex(T, W, H, D, sg.x, sg.y, sg.z) * r.density_scale * r.shadow_density_scale;
                            if (sdn < r.cutoff) continue;
                            const V3 e = muls(muls(ext, sdn), r.s_step);
                            const float dd = 1.0f + s * r.gain;
                            st = mul(st, bexp3(mk(e.x / dd, e.y / dd, e.z / dd)));
                            if (sqlen(st) < r.cutoff) break;
                        }
                    }
                    lumi = add(lumi, mul(mul(mul(mul(albedo, st), trans), lill[l]), sub(one, dT)));
                }
                lumi = add(lumi, mul(mul(trans, em), sub(one, dT)));
                trans = mul(trans, dT);
                if (sqlen(trans) < r.cutoff) break;
            }
            px[0] = lumi.x;
            px[1] = lumi.y;
            px[2] = lumi.z;
        }
}

} // namespace rnd
