// k_render.hip — the solver's consumer: volume ray marching of the
// temperature grid.
//
// Reference: RenderEngine::render (asiroisvigneux/Blaze
// src/RenderEngine.cu:312-365): write_scatter_to_3d_tex_kernel :76-101 (4x
// downsample of T above the ramp's low end), gaussian_blur_texture_kernel
// :51-74 (3x3x3 [1 2 1]-type filter, clamped), render_kernel :149-225
// (camera ray, box clip, primary march with emission, multi-scatter and one
// shadow march per directional light), bbox_intersect_kernel :112-147,
// Camera (src/Camera.h:31-60), Light (src/Light.h:31-55), Shader
// (src/Shader.h:33-75), ColorRamp (src/ColorRamp.h:32-80).
//
// Every expression keeps the reference's evaluation order. Three things are
// defined here instead of by the hardware or CUDA's library: the texture
// lookup is an explicit FP32 trilinear filter with clamped texel indices
// (the reference's clamp-addressed linear-filtered textures, whose weights
// the hardware stores in 8-bit fixed point), normalize divides by sqrtf
// (the reference multiplies by an approximate rsqrtf), and exp is the
// Cody-Waite reduction with a degree-7 Taylor polynomial below (the CPU
// oracle restates the same steps, so the image is reproducible bit for
// bit). A shadow ray that misses the box (only possible for degenerate
// directions; the reference would march it forever) leaves full
// transmittance, and both marches stop after 65536 samples.
#include <mutex>
#include <string>

#include "solver.h"

namespace blz {

namespace {

constexpr int RT = 8; // 8x8 pixel tiles, as the reference's launch
constexpr int MAX_MARCH = 1 << 16;

struct V3 {
    float x, y, z;
};
__device__ __forceinline__ V3 mkv(float x, float y, float z) { return V3{x, y, z}; }
__device__ __forceinline__ V3 operator+(V3 a, V3 b) { return mkv(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ __forceinline__ V3 operator-(V3 a, V3 b) { return mkv(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ __forceinline__ V3 operator*(V3 a, V3 b) { return mkv(a.x * b.x, a.y * b.y, a.z * b.z); }
__device__ __forceinline__ V3 operator*(V3 a, float s) { return mkv(a.x * s, a.y * s, a.z * s); }
__device__ __forceinline__ V3 operator*(float s, V3 a) { return mkv(s * a.x, s * a.y, s * a.z); }
__device__ __forceinline__ float sqlen(V3 v) { return v.x * v.x + v.y * v.y + v.z * v.z; }

// exp(x): x = k ln2 + r with ln2 split so that k*ln2_hi is exact (|k| <=
// 128), exp(r) by Taylor to degree 7 (|r| <= ln2/2, truncation < 1e-8
// relative), times 2^k as two exact powers of two. Below -87.34 (results
// under FLT_MIN) it returns 0, above 88.72 infinity.
__device__ __forceinline__ float bexp(float x) {
    if (x != x) return x;
    if (x > 88.7228394f) return __int_as_float(0x7f800000);
    if (x < -87.3365479f) return 0.0f;
    const float k = floorf(x * 1.44269502f + 0.5f);
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
    return p * __int_as_float((k1 + 127) << 23) * __int_as_float((k2 + 127) << 23);
}
__device__ __forceinline__ V3 bexp3(V3 a) { return mkv(bexp(a.x), bexp(a.y), bexp(a.z)); }

// linear-filtered, clamp-addressed lookup at unnormalised coordinates
// (texel centres at integer + 0.5)
__device__ __forceinline__ float tex_lin(const float* __restrict__ v, int nx, int ny, int nz, float x, float y,
                                         float z) {
    const float xb = x - 0.5f, yb = y - 0.5f, zb = z - 0.5f;
    const float fx = floorf(xb), fy = floorf(yb), fz = floorf(zb);
    const float a = xb - fx, b = yb - fy, c = zb - fz;
    const int i0 = (int)fminf(fmaxf(fx, -1.0f), (float)nx), j0 = (int)fminf(fmaxf(fy, -1.0f), (float)ny),
              k0 = (int)fminf(fmaxf(fz, -1.0f), (float)nz);
    const int xa = min(max(i0, 0), nx - 1), xc = min(max(i0 + 1, 0), nx - 1);
    const int ya = min(max(j0, 0), ny - 1), yc = min(max(j0 + 1, 0), ny - 1);
    const int za = min(max(k0, 0), nz - 1), zc = min(max(k0 + 1, 0), nz - 1);
    const size_t sy = (size_t)nx, sz = (size_t)nx * ny;
    auto t = [&](int i, int j, int k) { return v[(size_t)i + (size_t)j * sy + (size_t)k * sz]; };
    const float c00 = (1.0f - a) * t(xa, ya, za) + a * t(xc, ya, za);
    const float c10 = (1.0f - a) * t(xa, yc, za) + a * t(xc, yc, za);
    const float c01 = (1.0f - a) * t(xa, ya, zc) + a * t(xc, ya, zc);
    const float c11 = (1.0f - a) * t(xa, yc, zc) + a * t(xc, yc, zc);
    return (1.0f - c) * ((1.0f - b) * c00 + b * c10) + c * ((1.0f - b) * c01 + b * c11);
}

// scene constants computed on the host in the reference's order
struct RScene {
    V3 origin, top_left, horizontal, vertical;
    float res_x, res_y;
    V3 dom_min, dom_max, dom_size;
    float dx;
    int W, H, D, Ws, Hs, Ds;
    float p_step, s_step, cutoff;
    float density_scale, shadow_density_scale, gain, emission_scale;
    V3 albedo, extinction;
    int ramp_size;
    float ramp_in, ramp_out;
    float ramp_keys[BLZ_MAX_RAMP];
    V3 ramp_colors[BLZ_MAX_RAMP];
    V3 ms_color;
    float ms_scale, ms_mask_a, ms_mask_b;
    int light_count;
    V3 light_dir[BLZ_MAX_LIGHTS], light_illum[BLZ_MAX_LIGHTS];
};

__device__ __forceinline__ V3 ramp_color(const RScene& s, float temp) {
    const float remap = fminf(fmaxf((temp - s.ramp_in) / (s.ramp_out - s.ramp_in), 0.0f), 1.0f);
    if (s.ramp_size == 1) return s.ramp_colors[0];
    int prev = 0, next = 0;
    for (int i = 0; i < s.ramp_size; ++i) {
        if (s.ramp_keys[i] <= remap) {
            prev = i;
        } else {
            next = i;
            break;
        }
    }
    const float range = s.ramp_keys[next] - s.ramp_keys[prev];
    const float blend = (remap - s.ramp_keys[prev]) / range;
    return s.ramp_colors[prev] + blend * (s.ramp_colors[next] - s.ramp_colors[prev]);
}

__device__ __forceinline__ float smoothstep_f(float a, float b, float x) {
    const float y = fmaxf(0.0f, fminf((x - a) / (b - a), 1.0f));
    return y * y * (3.0f - (2.0f * y));
}

struct Ray {
    V3 o, d, inv;
    int sx, sy, sz;
    float t0, t1;
};
__device__ __forceinline__ Ray make_ray(V3 o, V3 d) {
    Ray r;
    r.o = o;
    r.d = d;
    r.inv = mkv(1.0f / d.x, 1.0f / d.y, 1.0f / d.z);
    r.sx = r.inv.x < 0.0f;
    r.sy = r.inv.y < 0.0f;
    r.sz = r.inv.z < 0.0f;
    r.t0 = 0.0f;
    r.t1 = __int_as_float(0x7f800000);
    return r;
}

// slab test of src/RenderEngine.cu:112-147 (sets t0 = max(tmin, 0), t1)
__device__ __forceinline__ bool box_hit(Ray& r, V3 lo, V3 hi) {
    const V3 b[2] = {lo, hi};
    float tmin = (b[r.sx].x - r.o.x) * r.inv.x;
    float tmax = (b[1 - r.sx].x - r.o.x) * r.inv.x;
    const float tymin = (b[r.sy].y - r.o.y) * r.inv.y;
    const float tymax = (b[1 - r.sy].y - r.o.y) * r.inv.y;
    if ((tmin > tymax) || (tymin > tmax)) return false;
    if (tymin > tmin) tmin = tymin;
    if (tymax < tmax) tmax = tymax;
    const float tzmin = (b[r.sz].z - r.o.z) * r.inv.z;
    const float tzmax = (b[1 - r.sz].z - r.o.z) * r.inv.z;
    if ((tmin > tzmax) || (tzmin > tmax)) return false;
    if (tzmin > tmin) tmin = tzmin;
    if (tzmax < tmax) tmax = tzmax;
    r.t0 = fmaxf(tmin, 0.0f);
    r.t1 = tmax;
    if (r.t1 < 0.0f) return false;
    return true;
}

// Grid::worldToGridNoOffset (src/Grid.h:292-296): y flipped
template <class S>
__device__ __forceinline__ V3 w2g(const RScene& s, const S& sc, V3 p) {
    return mkv(sc(p.x - s.dom_min.x), sc(s.dom_size.y - (p.y - s.dom_min.y)), sc(p.z - s.dom_min.z));
}

// write_scatter_to_3d_tex_kernel: mean over the 4x4x4 block of T above tmin
__global__ void __launch_bounds__(256) k_rnd_scatter(const float* __restrict__ T, int W, int H, int Ws, int Hs,
                                                     int Ds, float tmin, float* __restrict__ out) {
    const int t = blockIdx.x * 256 + threadIdx.x;
    if (t >= Ws * Hs * Ds) return;
    const int i = t % Ws, j = (t / Ws) % Hs, k = t / (Ws * Hs);
    float acc = 0.0f;
    for (int io = 0; io < 4; ++io)
        for (int jo = 0; jo < 4; ++jo)
            for (int ko = 0; ko < 4; ++ko) {
                const float v = T[(size_t)(i * 4 + io) + (size_t)(j * 4 + jo) * W + (size_t)(k * 4 + ko) * W * H];
                acc += v > tmin ? v : 0.0f;
            }
    out[t] = acc / 64.0f;
}

// gaussian_blur_texture_kernel: 27 taps (k outer, then j, then i), samples
// at texel centres of the clamped volume, weights w[tap] (sum 1)
struct BlurW {
    float w[27];
};
__global__ void __launch_bounds__(256) k_rnd_blur(const float* __restrict__ in, float* __restrict__ out, int Ws,
                                                  int Hs, int Ds, BlurW w) {
    const int t = blockIdx.x * 256 + threadIdx.x;
    if (t >= Ws * Hs * Ds) return;
    const int i = t % Ws, j = (t / Ws) % Hs, k = t / (Ws * Hs);
    float acc = 0.0f;
    int q = 0;
    for (int dk = -1; dk <= 1; ++dk)
        for (int dj = -1; dj <= 1; ++dj)
            for (int di = -1; di <= 1; ++di, ++q) {
                const int x = min(max(i + di, 0), Ws - 1), y = min(max(j + dj, 0), Hs - 1),
                          z = min(max(k + dk, 0), Ds - 1);
                acc += in[(size_t)x + (size_t)y * Ws + (size_t)z * Ws * Hs] * w.w[q];
            }
    out[t] = acc;
}

// render_kernel: one thread per pixel, 8x8 tiles
template <class S>
__global__ void __launch_bounds__(RT * RT) k_render(const float* __restrict__ T, const float* __restrict__ scat,
                                                    float* __restrict__ fb, RScene s, S sc) {
    const int i = blockIdx.x * RT + threadIdx.x, j = blockIdx.y * RT + threadIdx.y;
    const int w = (int)s.res_x;
    float* px = fb + 3 * ((size_t)i + (size_t)w * j);
    px[0] = px[1] = px[2] = 0.0f;
    // Camera::getRay: normalize(top_left + u*horizontal + v*vertical - origin)
    const V3 q = s.top_left + (((float)i + 0.5f) / s.res_x) * s.horizontal +
                 (((float)j + 0.5f) / s.res_y) * s.vertical - s.origin;
    const float inv_len = 1.0f / sqrtf(q.x * q.x + q.y * q.y + q.z * q.z);
    Ray pr = make_ray(s.origin, q * inv_len);
    if (!box_hit(pr, s.dom_min, s.dom_max)) return;
    V3 lumi = mkv(0.0f, 0.0f, 0.0f), trans = mkv(1.0f, 1.0f, 1.0f);
    const V3 one = mkv(1.0f, 1.0f, 1.0f);
    int n = 0;
    for (float pt = pr.t0 + s.dx * 0.5f; pt <= pr.t1 && n < MAX_MARCH; pt += s.p_step, ++n) {
        const V3 pos = pr.o + pt * pr.d;
        const V3 g = w2g(s, sc, pos);
        const float temp = tex_lin(T, s.W, s.H, s.D, g.x, g.y, g.z);
        const float scatter = tex_lin(scat, s.Ws, s.Hs, s.Ds, g.x / 4.0f, g.y / 4.0f, g.z / 4.0f);
        const float dens = temp * s.density_scale;
        if (dens < s.cutoff) continue;
        V3 emission = mkv(0.0f, 0.0f, 0.0f);
        if (s.emission_scale > 0.0f)
            emission = ramp_color(s, temp) * s.emission_scale +
                       s.ms_color * scatter * s.ms_scale * smoothstep_f(s.ms_mask_a, s.ms_mask_b, temp);
        const V3 dT = bexp3(s.extinction * dens * s.p_step);
        for (int l = 0; l < s.light_count; ++l) {
            Ray sr = make_ray(pos, -1.0f * s.light_dir[l]);
            V3 strans = mkv(1.0f, 1.0f, 1.0f);
            if (box_hit(sr, s.dom_min, s.dom_max)) {
                int m = 0;
                for (float st = sr.t0 + s.dx * 0.5f; st <= sr.t1 && m < MAX_MARCH; st += s.s_step, ++m) {
                    const V3 sg = w2g(s, sc, sr.o + st * sr.d);
                    const float sd =
                        tex_lin(T, s.W, s.H, s.D, sg.x, sg.y, sg.z) * s.density_scale * s.shadow_density_scale;
                    if (sd < s.cutoff) continue;
                    const V3 e = s.extinction * sd * s.s_step;
                    const float den = 1.0f + st * s.gain;
                    strans = strans * bexp3(mkv(e.x / den, e.y / den, e.z / den));
                    if (sqlen(strans) < s.cutoff) break;
                }
            }
            lumi = lumi + s.albedo * strans * trans * s.light_illum[l] * (one - dT);
        }
        lumi = lumi + trans * emission * (one - dT);
        trans = trans * dT;
        if (sqlen(trans) < s.cutoff) break;
    }
    px[0] = lumi.x;
    px[1] = lumi.y;
    px[2] = lumi.z;
}

inline unsigned nblk(size_t n) { return (unsigned)((n + 255) / 256); }

// zero fill (a kernel, not a memset node, so a captured render replays it)
__global__ void __launch_bounds__(256) k_rnd_zero(float* __restrict__ d, size_t n) {
    const size_t t = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (t < n) d[t] = 0.0f;
}

// Scratch of blz_render_noalloc: the two multi-scatter volumes and the
// device frame buffer, in place of the handle's grow-only allocation.
// 32 MiB per device (held until the process ends): a 256^3 grid's volumes
// (2 x 64^3 floats) plus a 1920x1080 frame (6.2M floats) fit. Every call
// on a device shares its copy; calls that overlap on one device race.
constexpr size_t RND_SCRATCH_FLOATS = (size_t)8 << 20;
static __device__ __align__(256) float scratch[RND_SCRATCH_FLOATS];

V3 h3(const float* a) { return V3{a[0], a[1], a[2]}; }
V3 hadd(V3 a, V3 b) { return V3{a.x + b.x, a.y + b.y, a.z + b.z}; }
V3 hsub(V3 a, V3 b) { return V3{a.x - b.x, a.y - b.y, a.z - b.z}; }
V3 hmul(V3 a, float s) { return V3{a.x * s, a.y * s, a.z * s}; }
V3 hmulf(float s, V3 a) { return V3{s * a.x, s * a.y, s * a.z}; }

} // namespace

// get_buf(n, &p): n floats of device scratch for the scatter volumes and the
// frame, or an error
template <class GetBuf>
static int render_with(Solver& s, const blz_render_settings& r, float* host_rgb, GetBuf get_buf) {
    if (r.width <= 0 || r.height <= 0 || r.width % RT || r.height % RT || r.width > 16384 || r.height > 16384) {
        set_error("render size must be positive multiples of 8 (at most 16384)");
        return BLZ_E_INVALID;
    }
    if (!(r.p_step > 0.0f) || !(r.s_step > 0.0f) || r.light_count < 0 || r.light_count > BLZ_MAX_LIGHTS ||
        r.ramp_size < 1 || r.ramp_size > BLZ_MAX_RAMP || r.ms_blur_iter < 0) {
        set_error("invalid render settings (steps > 0, 0..10 lights, 1..10 ramp keys, blur iterations >= 0)");
        return BLZ_E_INVALID;
    }
    RScene sc{};
    // Camera (src/Camera.h:42-53), host FP32 in the same order
    {
        const V3 o = h3(r.cam_trans), u = h3(r.cam_u), v = h3(r.cam_v), w = h3(r.cam_w);
        const float aspect = (float)r.width / (float)r.height;
        const float half_w = 0.5f * r.cam_aperture;
        const float half_h = half_w / aspect;
        sc.origin = o;
        sc.top_left = hadd(hsub(hsub(o, hmulf(r.cam_focal, w)), hmulf(half_w, u)), hmulf(half_h, v));
        sc.horizontal = hmulf(2.0f * half_w, u);
        sc.vertical = hmulf(-2.0f * half_h, v);
        sc.res_x = (float)(unsigned)r.width;
        sc.res_y = (float)(unsigned)r.height;
    }
    sc.dom_min = V3{s.s.domain_min[0], s.s.domain_min[1], s.s.domain_min[2]};
    sc.dom_max = V3{s.s.domain_max[0], s.s.domain_max[1], s.s.domain_max[2]};
    sc.dom_size = hsub(sc.dom_max, sc.dom_min);
    sc.dx = s.dx;
    sc.W = s.W;
    sc.H = s.H;
    sc.D = s.D;
    sc.Ws = s.W / 4;
    sc.Hs = s.H / 4;
    sc.Ds = s.D / 4;
    sc.p_step = r.p_step;
    sc.s_step = r.s_step;
    sc.cutoff = r.cutoff;
    sc.density_scale = r.density_scale;
    sc.shadow_density_scale = r.shadow_density_scale;
    sc.gain = r.gain;
    sc.emission_scale = r.emission_scale;
    // Shader (src/Shader.h:52-63)
    {
        const V3 col = h3(r.volume_color), sca = h3(r.scattering);
        const V3 den = V3{sca.x + r.absorption, sca.y + r.absorption, sca.z + r.absorption};
        const V3 num = V3{col.x * sca.x, col.y * sca.y, col.z * sca.z};
        sc.albedo = V3{num.x / den.x, num.y / den.y, num.z / den.z};
        sc.extinction = hmul(V3{sca.x + r.absorption, sca.y + r.absorption, sca.z + r.absorption}, -1.0f);
    }
    sc.ramp_size = r.ramp_size;
    sc.ramp_in = r.temp_range[0];
    sc.ramp_out = r.temp_range[1];
    for (int q = 0; q < r.ramp_size; ++q) {
        sc.ramp_keys[q] = r.ramp_keys[q];
        sc.ramp_colors[q] = h3(r.ramp_colors[q]);
    }
    sc.ms_color = h3(r.ms_color);
    sc.ms_scale = r.ms_scale;
    sc.ms_mask_a = r.ms_mask[0];
    sc.ms_mask_b = r.ms_mask[1];
    // Light (src/Light.h:43-50): colour * intensity * 2^exposure
    sc.light_count = r.light_count;
    for (int l = 0; l < r.light_count; ++l) {
        sc.light_dir[l] = h3(r.light_dir[l]);
        sc.light_illum[l] = hmul(hmul(h3(r.light_color[l]), r.light_intensity[l]), std::pow(2.0f, r.light_exposure[l]));
    }
    // buffers: two scatter volumes (at least one cell each) and the frame
    const size_t ns = (size_t)std::max(1, sc.Ws) * std::max(1, sc.Hs) * std::max(1, sc.Ds);
    const size_t nf = (size_t)r.width * r.height * 3;
    float* buf = nullptr;
    if (int e = get_buf(2 * ns + nf, &buf)) return e;
    float* scat[2] = {buf, buf + ns};
    float* fb = buf + 2 * ns;
    const float* T = s.f[F_T].front();
    const bool have_scatter = sc.Ws > 0 && sc.Hs > 0 && sc.Ds > 0;
    if (!have_scatter) {
        sc.Ws = sc.Hs = sc.Ds = 1;
    }
    if (r.ms_scale > 0.0f && have_scatter) {
        hipLaunchKernelGGL(k_rnd_scatter, dim3(nblk(ns)), dim3(256), 0, s.stream, T, s.W, s.H, sc.Ws, sc.Hs, sc.Ds,
                           r.temp_range[0], scat[0]);
        // blur weights: filter / its sum (src/RenderEngine.cu:33-47, :262-276)
        static const float filt[27] = {0, 1, 0, 1, 2, 1, 0, 1, 0, 1, 2, 1, 2, 4, 2, 1, 2, 1, 0, 1, 0, 1, 2, 1, 0, 1, 0};
        float sum = 0.0f;
        for (int q = 0; q < 27; ++q) sum += filt[q];
        BlurW bw;
        for (int q = 0; q < 27; ++q) bw.w[q] = filt[q] / sum;
        for (int it = 0; it < r.ms_blur_iter; ++it) {
            hipLaunchKernelGGL(k_rnd_blur, dim3(nblk(ns)), dim3(256), 0, s.stream, scat[0], scat[1], sc.Ws, sc.Hs,
                               sc.Ds, bw);
            std::swap(scat[0], scat[1]);
        }
    } else {
        hipLaunchKernelGGL(k_rnd_zero, dim3(nblk(ns)), dim3(256), 0, s.stream, scat[0], ns);
    }
    const dim3 grid((unsigned)(r.width / RT), (unsigned)(r.height / RT)), block(RT, RT);
    float inv = 0.0f;
    if (dx_pow2(s.dx, &inv))
        hipLaunchKernelGGL(k_render<MulDx>, grid, block, 0, s.stream, T, scat[0], fb, sc, MulDx{inv});
    else
        hipLaunchKernelGGL(k_render<DivDx>, grid, block, 0, s.stream, T, scat[0], fb, sc, DivDx{s.dx});
    BLZ_HIP(hipGetLastError());
    BLZ_HIP(hipMemcpyAsync(host_rgb, fb, nf * sizeof(float), hipMemcpyDeviceToHost, s.stream));
    BLZ_HIP(hipStreamSynchronize(s.stream));
    return BLZ_OK;
}

int render_frame(Solver& s, const blz_render_settings& r, float* host_rgb) {
    return render_with(s, r, host_rgb, [&s](size_t n, float** out) -> int {
        const size_t bytes = n * sizeof(float);
        if (bytes > s.rnd_bytes) {
            if (s.rnd_buf) BLZ_HIP(hipFree(s.rnd_buf));
            s.rnd_buf = nullptr;
            s.rnd_bytes = 0;
            if (hipMalloc(&s.rnd_buf, bytes) != hipSuccess) {
                set_error("hipMalloc failed for render buffers");
                return BLZ_E_NOMEM;
            }
            s.rnd_bytes = bytes;
        }
        *out = s.rnd_buf;
        return BLZ_OK;
    });
}

// Device address of `scratch` on each device, looked up once per device
// (with that device current) by hipGetSymbolAddress; host code uses these
// addresses, never the array's name (a host-side stand-in).
static std::mutex g_scratch_mu;
static float* g_scratch_dev[64] = {};

static int scratch_on(int device, float** out) {
    if (device < 0 || device >= 64) {
        set_error("blz_render_noalloc: device index out of range");
        return BLZ_E_INVALID;
    }
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_dev[device]) {
        int cur = -1;
        BLZ_HIP(hipGetDevice(&cur));
        if (cur != device) {
            set_error("blz_render_noalloc: the solver's device must be current");
            return BLZ_E_INVALID;
        }
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            set_error(std::string("hipGetSymbolAddress(scratch) failed: ") + hipGetErrorString(e));
            return BLZ_E_HIP;
        }
        g_scratch_dev[device] = (float*)p;
    }
    *out = g_scratch_dev[device];
    return BLZ_OK;
}

int render_frame_noalloc(Solver& s, const blz_render_settings& r, float* host_rgb) {
    float* base = nullptr;
    if (int e = scratch_on(s.device, &base)) return e;
    return render_with(s, r, host_rgb, [base](size_t n, float** out) -> int {
        if (n > RND_SCRATCH_FLOATS) {
            set_error("blz_render_noalloc: scatter volumes plus frame need " + std::to_string(n) +
                      " floats, above the scratch limit of " + std::to_string(RND_SCRATCH_FLOATS));
            return BLZ_E_INVALID;
        }
        *out = base;
        return BLZ_OK;
    });
}

} // namespace blz
