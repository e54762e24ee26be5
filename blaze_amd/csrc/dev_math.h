// dev_math.h — device-side sampling, noise and vector helpers of the solver.
//
// Every expression keeps the evaluation order of the reference source
// (asiroisvigneux/Blaze src/Grid.h, src/FluidSolver.cu,
// thirdparty/cuda-noise/cuda_noise.cuh) so that, built with
// -ffp-contract=off and IEEE division/sqrt, results are bit-identical to
// a serial CPU evaluation of the same formulas.
#pragma once
#include <cmath>

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace blz {

struct f3 {
    float x, y, z;
};
__device__ __forceinline__ f3 mk3(float x, float y, float z) { return f3{x, y, z}; }
// helper_math.h:1267-1269 (dot), :1314-1316 (length), :1443-1446 (cross)
__device__ __forceinline__ float dot3(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ float len3(f3 a) { return sqrtf(dot3(a, a)); }
__device__ __forceinline__ f3 cross3(f3 a, f3 b) {
    return mk3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}

// Read-only view of one staggered grid: dims, origin offset (cells), dx.
// Index is x-fastest: i + j*nx + k*nx*ny (src/Grid.h:256-268).
struct Gv {
    const float* __restrict__ a;
    int nx, ny, nz;
    float ox, oy, oz;
};

__device__ __forceinline__ float ld(const Gv& g, int i, int j, int k) {
    return g.a[i + j * g.nx + k * (g.nx * g.ny)];
}

// A pointer known to address global (device) memory, typed so: loads and
// stores through it are global instructions rather than flat ones (a pointer
// read from memory, or picked from an array by a computed index, is
// otherwise generic to the compiler). The address is the same.
#define BLZ_GLOBAL __attribute__((address_space(1)))
template <class T>
__device__ __forceinline__ BLZ_GLOBAL T* as_global(T* p) {
    return (BLZ_GLOBAL T*)p;
}

// grid -> object position of cell (i,j,k): (i + o) * dx  (src/Grid.h:270-290)
__device__ __forceinline__ f3 g2o(const Gv& g, float dx, int i, int j, int k) {
    return mk3(((float)i + g.ox) * dx, ((float)j + g.oy) * dx, ((float)k + g.oz) * dx);
}

// object -> grid units: p / dx (IEEE division), or p * 2^k when dx = 2^-k,
// which gives the same bits (scaling by a power of two is exact, and
// overflow, zeros, infinities and NaNs come out alike)
struct DivDx {
    float dx;
    __device__ __forceinline__ float operator()(float v) const { return v / dx; }
};
struct MulDx {
    float inv;
    __device__ __forceinline__ float operator()(float v) const { return v * inv; }
};
// true (and inv = 1/dx, exact) when dx is a power of two
inline bool dx_pow2(float dx, float* inv) {
    int e = 0;
    const float m = std::frexp(dx, &e);
    if (!(dx > 0.0f) || m != 0.5f) return false;
    *inv = std::ldexp(1.0f, 1 - e);
    return std::isfinite(*inv) && *inv * dx == 1.0f;
}

// Gathers of the samplers below. A field of fewer than 2^29 floats is read
// with buffer loads: a 32-bit byte offset against a buffer resource built
// from the field's (uniform) base, so the 64-bit address arithmetic of a
// pointer gather disappears; the values read are the same. Larger fields
// take the pointer path (a wave-uniform branch).
struct GatherB {
    __amdgpu_buffer_rsrc_t r;
    __device__ __forceinline__ float operator()(long long idx) const {
        // the builtin returns the 32 bits as an integer: reinterpret them
        return __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(r, (int)idx * 4, 0, 0));
    }
};
struct GatherP {
    const float* a;
    __device__ __forceinline__ float operator()(long long idx) const { return a[idx]; }
};
__device__ __forceinline__ bool gather_buffered(const Gv& g) {
    return (long long)g.nx * g.ny * g.nz < (1ll << 29);
}
__device__ __forceinline__ GatherB gather_b(const Gv& g) {
    return GatherB{__builtin_amdgcn_make_buffer_rsrc((void*)g.a, (short)0, g.nx * g.ny * g.nz * 4, 0x00020000)};
}

// trilinear_interpolate, src/Grid.h:94-108
template <class S, class L>
__device__ __forceinline__ float trilinear_g(const Gv& g, const L& ldg, const S& sc, f3 p) {
    const float gX = fmaxf(fminf(sc(p.x) - g.ox, (float)g.nx - 1.001f), 0.0f);
    const float gY = fmaxf(fminf(sc(p.y) - g.oy, (float)g.ny - 1.001f), 0.0f);
    const float gZ = fmaxf(fminf(sc(p.z) - g.oz, (float)g.nz - 1.001f), 0.0f);
    const int i = (int)gX, j = (int)gY, k = (int)gZ;
    const float tx = gX - (float)i, ty = gY - (float)j, tz = gZ - (float)k;
    const int sy = g.nx, sz = g.nx * g.ny;
    const int b = i + j * sy + k * sz;
    const float c00 = ldg(b) * (1.0f - tx) + ldg(b + 1) * tx;
    const float c10 = ldg(b + sy) * (1.0f - tx) + ldg(b + sy + 1) * tx;
    const float c01 = ldg(b + sz) * (1.0f - tx) + ldg(b + sz + 1) * tx;
    const float c11 = ldg(b + sy + sz) * (1.0f - tx) + ldg(b + sy + sz + 1) * tx;
    return (1.0f - tz) * ((1.0f - ty) * c00 + ty * c10) + tz * ((1.0f - ty) * c01 + ty * c11);
}
template <class S>
__device__ __forceinline__ float trilinear_s(const Gv& g, const S& sc, f3 p) {
    if (gather_buffered(g)) return trilinear_g(g, gather_b(g), sc, p);
    return trilinear_g(g, GatherP{g.a}, sc, p);
}
__device__ __forceinline__ float trilinear(const Gv& g, float dx, f3 p) { return trilinear_s(g, DivDx{dx}, p); }

// cerp / cerpc, src/Grid.h:148-163 (cubic Lagrange on nodes -1,0,1,2)
__device__ __forceinline__ float cerp(float q0, float q1, float q2, float q3, float s) {
    const float w0 = (-1.0f / 3.0f) * s + 0.5f * s * s - (1.0f / 6.0f) * s * s * s;
    const float w1 = 1.0f - s * s + 0.5f * (s * s * s - s);
    const float w2 = s + 0.5f * (s * s - s * s * s);
    const float w3 = (1.0f / 6.0f) * (s * s * s - s);
    return w0 * q0 + w1 * q1 + w2 * q2 + w3 * q3;
}

// cubic weights once per axis; cerp(q, s) == cw_apply(cw(s), q) bit for bit
struct CW {
    float w0, w1, w2, w3;
};
__device__ __forceinline__ CW cw(float s) {
    CW c;
    c.w0 = (-1.0f / 3.0f) * s + 0.5f * s * s - (1.0f / 6.0f) * s * s * s;
    c.w1 = 1.0f - s * s + 0.5f * (s * s * s - s);
    c.w2 = s + 0.5f * (s * s - s * s * s);
    c.w3 = (1.0f / 6.0f) * (s * s * s - s);
    return c;
}
__device__ __forceinline__ float cw_apply(const CW& c, float q0, float q1, float q2, float q3) {
    return c.w0 * q0 + c.w1 * q1 + c.w2 * q2 + c.w3 * q3;
}

// tricubic_interpolate, src/Grid.h:165-222: x cerps, then y, then z; with
// clamp, every 1-D stage is clamped at 0 (cerpc).
template <bool CLAMP, class S, class L>
__device__ __forceinline__ float tricubic_g(const Gv& g, const L& ldg, const S& sc, f3 p) {
    const float gX = fmaxf(fminf(sc(p.x) - g.ox, (float)g.nx - 1.001f), 0.0f);
    const float gY = fmaxf(fminf(sc(p.y) - g.oy, (float)g.ny - 1.001f), 0.0f);
    const float gZ = fmaxf(fminf(sc(p.z) - g.oz, (float)g.nz - 1.001f), 0.0f);
    const int i1 = (int)gX, j1 = (int)gY, k1 = (int)gZ;
    const int sy = g.nx, sz = g.nx * g.ny;
    const int ix[4] = {max(i1 - 1, 0), i1, i1 + 1, min(i1 + 2, g.nx - 1)};
    const int jy[4] = {max(j1 - 1, 0) * sy, j1 * sy, (j1 + 1) * sy, min(j1 + 2, g.ny - 1) * sy};
    const int kz[4] = {max(k1 - 1, 0) * sz, k1 * sz, (k1 + 1) * sz, min(k1 + 2, g.nz - 1) * sz};
    const CW cx = cw(gX - (float)i1), cy = cw(gY - (float)j1), cz = cw(gZ - (float)k1);
    float rz[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        float ry[4];
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            const int row = jy[b] + kz[c];
            float v = cw_apply(cx, ldg(row + ix[0]), ldg(row + ix[1]), ldg(row + ix[2]), ldg(row + ix[3]));
            ry[b] = CLAMP ? fmaxf(v, 0.0f) : v;
        }
        float v = cw_apply(cy, ry[0], ry[1], ry[2], ry[3]);
        rz[c] = CLAMP ? fmaxf(v, 0.0f) : v;
    }
    float v = cw_apply(cz, rz[0], rz[1], rz[2], rz[3]);
    return CLAMP ? fmaxf(v, 0.0f) : v;
}
template <bool CLAMP, class S>
__device__ __forceinline__ float tricubic_s(const Gv& g, const S& sc, f3 p) {
    if (gather_buffered(g)) return tricubic_g<CLAMP>(g, gather_b(g), sc, p);
    return tricubic_g<CLAMP>(g, GatherP{g.a}, sc, p);
}
template <bool CLAMP>
__device__ __forceinline__ float tricubic(const Gv& g, float dx, f3 p) {
    return tricubic_s<CLAMP>(g, DivDx{dx}, p);
}

// fit, src/FluidSolver.cu:250-252
__device__ __forceinline__ float fitf(float v, float minIn, float maxIn, float minOut, float maxOut) {
    return fmaxf(fminf((v - minIn) / (maxIn - minIn), 1.0f), 0.0f) * (maxOut - minOut) + minOut;
}

// ---- Perlin noise, thirdparty/cuda-noise/cuda_noise.cuh ----

// 32-bit integer mix, :46-56
__device__ __forceinline__ uint32_t nhash(uint32_t s) {
    s = (s + 0x7ed55d16u) + (s << 12);
    s = (s ^ 0xc761c23cu) ^ (s >> 19);
    s = (s + 0x165667b1u) + (s << 5);
    s = (s + 0xd3a2646cu) ^ (s << 9);
    s = (s + 0xfd7046c5u) + (s << 3);
    s = (s ^ 0xb55a4f09u) ^ (s >> 16);
    return s;
}

// saturating float -> uint32 (negative and NaN -> 0, >= 2^32 -> max)
__device__ __forceinline__ uint32_t f2u_sat(float v) {
    if (!(v > 0.0f)) return 0u;
    if (v >= 4294967296.0f) return 0xffffffffu;
    return (uint32_t)v;
}

// randomIntGrid :120-123, float sum left to right
__device__ __forceinline__ uint32_t lattice_hash(float x, float y, float z, float seed) {
    float a = x * 1723.0f;
    a = a + y * 93241.0f;
    a = a + z * 149812.0f;
    a = a + 3824.0f;
    a = a + seed;
    return nhash(f2u_sat(a));
}

// grad :179-201 (0xD = -y+z and 0xF = -y-z as in the source)
__device__ __forceinline__ float ngrad(uint32_t h, float x, float y, float z) {
    switch (h & 0xFu) {
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

__device__ __forceinline__ float nfade(float t) { return t * t * t * (t * (t * 6.0f - 15.0f) + 10.0f); }
__device__ __forceinline__ float nlerp(float a, float b, float r) { return a * (1.0f - r) + b * r; }

// perlinNoise :570-615
__device__ __forceinline__ float perlin(f3 pos, float scale, int seed) {
    const float fseed = (float)seed;
    float px = pos.x * scale, py = pos.y * scale, pz = pos.z * scale;
    const float ix = floorf(px), iy = floorf(py), iz = floorf(pz);
    px -= ix;
    py -= iy;
    pz -= iz;
    const float u = nfade(px), v = nfade(py), w = nfade(pz);
    const float ix1 = ix + 1.0f, iy1 = iy + 1.0f, iz1 = iz + 1.0f;
    const float i000 = ngrad(lattice_hash(ix, iy, iz, fseed), px, py, pz);
    const float i100 = ngrad(lattice_hash(ix1, iy, iz, fseed), px - 1.0f, py, pz);
    const float i010 = ngrad(lattice_hash(ix, iy1, iz, fseed), px, py - 1.0f, pz);
    const float i110 = ngrad(lattice_hash(ix1, iy1, iz, fseed), px - 1.0f, py - 1.0f, pz);
    const float i001 = ngrad(lattice_hash(ix, iy, iz1, fseed), px, py, pz - 1.0f);
    const float i101 = ngrad(lattice_hash(ix1, iy, iz1, fseed), px - 1.0f, py, pz - 1.0f);
    const float i011 = ngrad(lattice_hash(ix, iy1, iz1, fseed), px, py - 1.0f, pz - 1.0f);
    const float i111 = ngrad(lattice_hash(ix1, iy1, iz1, fseed), px - 1.0f, py - 1.0f, pz - 1.0f);
    const float x00 = nlerp(i000, i100, u);
    const float x10 = nlerp(i010, i110, u);
    const float x01 = nlerp(i001, i101, u);
    const float x11 = nlerp(i011, i111, u);
    const float y0 = nlerp(x00, x10, v);
    const float y1 = nlerp(x01, x11, v);
    return nlerp(y0, y1, w);
}

__device__ __forceinline__ bool abs_gt_1e5(float v) { return (double)fabsf(v) > 1e-5; }

// a / d for d in {3, 4, 5, 6} with y = RN(1/d): q0 = a*y, r = a - d*q0
// (exact by FMA), q = q0 + r*y. Checked against IEEE division for every
// float: identical whenever 2^-125 <= |a| <= FLT_MAX (d = 3, 4, 5 also
// below, except a = -0; d = 6 differs only where a/d is subnormal). Outside
// that range the IEEE division is taken.
__device__ __forceinline__ float div_small(float a, float d, float y) {
    const float q0 = a * y;
    const float r = fmaf(-q0, d, a);
    float q = fmaf(r, y, q0);
    const float m = fabsf(a);
    if (!(m >= 0x1p-125f && m <= 3.40282347e38f)) q = a / d;
    return q;
}

} // namespace blz
