// k_project.hip — pressure projection: divergence, red/black Gauss-Seidel
// pressure solve, FP64 residual, pressure-gradient velocity update.
//
// Reference kernels (asiroisvigneux/Blaze src/FluidSolver.cu):
//   compute_divergence_kernel :475-515, gs_solve_kernel :517-599,
//   pressure_gradient_update_kernel :601-669, project() :823-856.
//
// Pressure and rhs are stored split by colour (red: (i+j) even, black: odd)
// in padded planes (see RBL in solver.h). A sweep is a red pass then a black
// pass; in a pass every column of that colour walks k = 0..D-1 updating in
// place, reading the other colour (fixed during the pass) and its own k-1
// value just computed. This equals the reference's per-slab red/black order
// inside a thread block, without its inter-block races, and every access of
// a pass is unit-stride across the wave.
#include "solver.h"

namespace blz {

namespace {

constexpr unsigned BS = 256;

__device__ __forceinline__ float p_at(const RB& p, const RBL& L, int i, int j, int k) {
    return as_global(((i + j) & 1) ? p.c[1] : p.c[0])[L.at(i, j, k)];
}

// compute_divergence_kernel :475-515, written into the split layout
__device__ __forceinline__ void divergence_cell(const float* __restrict__ U, const float* __restrict__ V,
                                                const float* __restrict__ Wf, const RB& rhs, const RBL& L, int W,
                                                int H, int D, float sc, int4 cbA, int2 cbB, int t) {
    if (t >= W * H * D) return;
    const int i = t % W, j = (t / W) % H, k = t / (W * H);
    const int u0 = i + j * (W + 1) + k * (W + 1) * H;
    const int v0 = i + j * W + k * W * (H + 1);
    const int w0 = t;
    const BLZ_GLOBAL float *Ug = as_global(U), *Vg = as_global(V), *Wg = as_global(Wf);
    const float u_lo = Ug[u0], u_hi = Ug[u0 + 1];
    const float v_lo = Vg[v0], v_hi = Vg[v0 + W];
    const float w_lo = Wg[w0], w_hi = Wg[w0 + W * H];
    float r = -sc * (u_hi - u_lo + v_hi - v_lo + w_hi - w_lo);
    // closedBounds: cbA = (xMin, xMax, yMin, yMax), cbB = (zMin, zMax)
    if (i == 0) {
        if (cbA.x) r -= sc * u_lo;
    } else if (i == W - 1) {
        if (cbA.y) r += sc * u_hi;
    }
    if (j == 0) {
        if (cbA.w) r -= sc * v_lo;
    } else if (j == H - 1) {
        if (cbA.z) r += sc * v_hi;
    }
    if (k == 0) {
        if (cbB.x) r -= sc * w_lo;
    } else if (k == D - 1) {
        if (cbB.y) r += sc * w_hi;
    }
    as_global(((i + j) & 1) ? rhs.c[1] : rhs.c[0])[L.at(i, j, k)] = r;
}

__global__ void __launch_bounds__(BS) k_divergence(const float* __restrict__ U, const float* __restrict__ V,
                                                   const float* __restrict__ Wf, RB rhs, RBL L, int W, int H, int D,
                                                   float sc, int4 cbA, int2 cbB) {
    divergence_cell(U, V, Wf, rhs, L, W, H, D, sc, cbA, cbB, (int)(blockIdx.x * BS + threadIdx.x));
}

// XCD-aware block order: the dispatcher hands consecutive workgroups to the
// 8 XCDs in turn, so logical block (b % 8) * (nb / 8) + b / 8 keeps runs of
// neighbouring rows (which share their j +- 1 loads) on one XCD's L2.
__device__ __forceinline__ int xcd_block(int b, int nb) {
    return (nb % 8 == 0) ? (b % 8) * (nb / 8) + b / 8 : b;
}

struct DivTab {
    float y3, y4, y5, y6; // RN(1/d)
};

__device__ __forceinline__ float rcp_of(int d, const DivTab& t) {
    return d == 3 ? t.y3 : (d == 4 ? t.y4 : (d == 5 ? t.y5 : t.y6));
}

// One colour pass of the Gauss-Seidel sweep (gs_solve_kernel :526-598).
// Thread per column of the colour, walking k = 0..D-1; per cell
//   num = scale*rhs + p(i-1) + p(i+1) + p(j-1) + p(j+1) + p(k-1) + p(k+1)
// in that order, p = num / (number of in-domain neighbours). Missing
// neighbours read the -0 ghost shell (exact identity) and the k-1 term of
// slab 0 is a -0 register, so there are no per-cell conditions.
//
// Memory: all four planes are one buffer resource; each lane keeps fixed
// 32-bit byte offsets and the slab-group offset is a wave-uniform scalar. A
// group of four slabs costs six 16-byte loads (rhs, four neighbours, own old
// values) and one 16-byte store per lane, so PG groups (PG <= 8, within the
// 63-deep memory counter) are in flight. The loop runs a multiple of PG
// groups: groups past the end read and rewrite the ghost group with -0, and
// slabs k >= D of the last group are stored as -0, so every group costs the
// same memory operations. PG >= 2 (group g+1 is in flight while g is used).
//
// Division: FAST runs use q = q0 + r*y (q0 = num*y, r = num - d*q0 by FMA)
// with the sign of num copied onto q, which equals IEEE num / d for every
// finite num except nonzero |num| < 2^-125; if any lane of the wave meets
// such a numerator (or a non-finite one) in a group, the run returns that
// group before storing it and the pass resumes there with the EXACT run
// (div_small, IEEE fallback per cell). Both runs read only values the pass
// has not yet written, so the resumed run sees the same inputs.
typedef float f4v __attribute__((ext_vector_type(4)));
typedef unsigned u4v __attribute__((ext_vector_type(4)));
typedef int i4v __attribute__((ext_vector_type(4)));

// Slab-group loads are issued as inline-asm buffer loads and consumed after
// explicit s_waitcnt vmcnt(N) statements that take the destination registers
// as in/out operands: uses cannot move above the wait, and the registers stay
// allocated while a load is in flight. (Left to itself the compiler copies
// the ring registers at the loop back-edge, which forces a full wait once per
// iteration.) Every memory instruction of the loop is counted in N: six
// loads and one store per group, nothing else (the kernel has no scratch).
__device__ __forceinline__ f4v ld4_async(int voff, const i4v& rs, int soff) {
    f4v r;
    asm volatile("buffer_load_dwordx4 %0, %1, %2, %3 offen"
                 : "=v"(r)
                 : "v"(voff), "s"(rs), "s"(__builtin_amdgcn_readfirstlane(soff)));
    return r;
}
__device__ __forceinline__ void st4_async(const f4v& v, int voff, const i4v& rs, int soff) {
    asm volatile("buffer_store_dwordx4 %0, %1, %2, %3 offen"
                 :
                 : "v"(v), "v"(voff), "s"(rs), "s"(__builtin_amdgcn_readfirstlane(soff))
                 : "memory");
}
template <int N>
__device__ __forceinline__ void wait_group(f4v& a, f4v& b, f4v& c, f4v& d, f4v& e, f4v& f, f4v& g) {
    asm volatile("s_waitcnt vmcnt(%7)" : "+v"(a), "+v"(b), "+v"(c), "+v"(d), "+v"(e), "+v"(f), "+v"(g) : "n"(N));
}
__device__ __forceinline__ void hold6(f4v& a, f4v& b, f4v& c, f4v& d, f4v& e, f4v& f) {
    asm volatile("s_waitcnt vmcnt(0)" : "+v"(a), "+v"(b), "+v"(c), "+v"(d), "+v"(e), "+v"(f));
}

struct PassArgs {
    i4v rs; // buffer resource: base, stride 0, size, raw dword access
    int bC, bR, bL, bRt, bN, bS;
    int D, Dg, gstride;
    float scale, d_mid, d_edge, y_mid, y_edge;
};

// vmcnt bound before group u of a ring of PG groups (see column_run)
template <int PG, int U>
struct GroupWait {
    static constexpr int N = (U < PG - 1) ? 6 * PG - 12 + U : 7 * (PG - 2);
};

template <int PG, int U, bool EXACT>
__device__ __forceinline__ bool group_step(const PassArgs& a, f4v (&sl)[PG][6], int g, float& prev) {
    constexpr int UN = (U + 1) % PG;
    // slot U holds group g, slot UN group g+1. Ops issued after slot UN's
    // last load: in the first pass over the ring 6*(PG-2-U) prologue loads
    // plus 7 per finished group (= 6*PG - 12 + U); later, 7*(PG-2).
    wait_group<GroupWait<PG, U>::N>(sl[U][0], sl[U][1], sl[U][2], sl[U][3], sl[U][4], sl[U][5], sl[UN][5]);
    const int D = a.D;
    const int k0 = 4 * g;
    const float up[4] = {sl[U][5][1], sl[U][5][2], sl[U][5][3], sl[UN][5][0]};
    float X[4];
#pragma unroll
    for (int s = 0; s < 4; ++s) {
        float x = a.scale * sl[U][0][s];
        x = x + sl[U][1][s];
        x = x + sl[U][2][s];
        x = x + sl[U][3][s];
        X[s] = x + sl[U][4][s];
    }
    f4v out;
    unsigned bad = 0u;
    float pv = prev;
#pragma unroll
    for (int s = 0; s < 4; ++s) {
        const int k = k0 + s;
        float num = X[s] + pv;
        num = num + up[s];
        const bool edge = (k == 0) | (k == D - 1);
        const float dd = edge ? a.d_edge : a.d_mid, yy = edge ? a.y_edge : a.y_mid;
        float pn;
        if (EXACT) {
            pn = div_small(num, dd, yy);
        } else {
            const float q0 = num * yy;
            const float rr = fmaf(-q0, dd, num);
            pn = copysignf(fmaf(rr, yy, q0), num);
            const float mag = fabsf(num);
            const unsigned ok = (unsigned)(mag >= 0x1p-125f) & (unsigned)(mag <= 3.40282347e38f);
            const unsigned zero = (unsigned)(mag == 0.0f);
            bad |= (k < D) ? ((ok | zero) ^ 1u) : 0u;
        }
        out[s] = (k < D) ? pn : -0.0f; // slabs k >= D keep the -0 ghost value
        pv = pn;
    }
    if (!EXACT && __builtin_amdgcn_ballot_w64(bad != 0u)) return false;
    prev = pv;
    st4_async(out, a.bC, a.rs, min(g, a.Dg) * a.gstride);
    const int so = min(g + PG, a.Dg) * a.gstride; // group Dg is the ghost group
    sl[U][0] = ld4_async(a.bR, a.rs, so);
    sl[U][1] = ld4_async(a.bL, a.rs, so);
    sl[U][2] = ld4_async(a.bRt, a.rs, so);
    sl[U][3] = ld4_async(a.bN, a.rs, so);
    sl[U][4] = ld4_async(a.bS, a.rs, so);
    sl[U][5] = ld4_async(a.bC, a.rs, so);
    return true;
}

template <int PG, bool EXACT, int U>
__device__ __forceinline__ bool ring_steps(const PassArgs& a, f4v (&sl)[PG][6], int gb, float& prev, int* failed) {
    if constexpr (U == PG) {
        return true;
    } else {
        if (!group_step<PG, U, EXACT>(a, sl, gb + U, prev)) {
            *failed = gb + U;
            return false;
        }
        return ring_steps<PG, EXACT, U + 1>(a, sl, gb, prev, failed);
    }
}

// Walk groups g_start.. of this thread's column. Returns the number of
// groups (all done) or, for a fast run, the first group where some lane of
// the wave needs the exact division (nothing of that group was stored).
template <int PG, bool EXACT>
__device__ __forceinline__ int column_run(const PassArgs& a, int g_start, float& prev) {
    f4v sl[PG][6];
#pragma unroll
    for (int u = 0; u < PG; ++u) {
        const int so = min(g_start + u, a.Dg) * a.gstride;
        sl[u][0] = ld4_async(a.bR, a.rs, so);
        sl[u][1] = ld4_async(a.bL, a.rs, so);
        sl[u][2] = ld4_async(a.bRt, a.rs, so);
        sl[u][3] = ld4_async(a.bN, a.rs, so);
        sl[u][4] = ld4_async(a.bS, a.rs, so);
        sl[u][5] = ld4_async(a.bC, a.rs, so);
    }
    int result = a.Dg;
    for (int gb = g_start; gb < a.Dg; gb += PG) {
        int failed = a.Dg;
        if (!ring_steps<PG, EXACT, 0>(a, sl, gb, prev, &failed)) {
            result = failed;
            break;
        }
    }
    // drain: no load may land after its registers are reused
#pragma unroll
    for (int u = 0; u < PG; ++u) hold6(sl[u][0], sl[u][1], sl[u][2], sl[u][3], sl[u][4], sl[u][5]);
    return result;
}

template <int PG>
__global__ void __launch_bounds__(64) k_rbgs_pass(float* __restrict__ base, unsigned nbytes, unsigned ownP,
                                                  unsigned othP, unsigned ownR, int W, int H, int D, int Wp, int Hp,
                                                  int colour, float scale, DivTab tab, const int* __restrict__ stop) {
    if (*stop) return; // the solve has converged
    const int Wh = W >> 1;
    const int t = xcd_block(blockIdx.x, gridDim.x) * 64 + threadIdx.x;
    if (t >= Wh * H) return;
    const int m = t % Wh, j = t / Wh;
    const int q = (j + colour) & 1;
    const int i = 2 * m + q;
    const unsigned e0 = (unsigned)((Hp + (j + 1)) * Wp + (m + 1)); // element of group 0
    const int cnt = (int)(i > 0) + (int)(i < W - 1) + (int)(j > 0) + (int)(j < H - 1);
    PassArgs a;
    const unsigned long long b64 = (unsigned long long)base;
    a.rs[0] = (int)(unsigned)(b64 & 0xffffffffull);
    a.rs[1] = (int)(unsigned)(b64 >> 32); // stride 0
    a.rs[2] = (int)nbytes;
    a.rs[3] = 0x00020000; // raw 32-bit access
    a.bC = (int)(16u * e0 + ownP);
    a.bR = (int)(16u * e0 + ownR);
    a.bL = (int)(16u * (e0 + q - 1) + othP);
    a.bRt = (int)(16u * (e0 + q) + othP);
    a.bN = (int)(16u * (e0 - (unsigned)Wp) + othP);
    a.bS = (int)(16u * (e0 + (unsigned)Wp) + othP);
    a.D = D;
    a.Dg = (D + 3) >> 2;
    a.gstride = Hp * Wp * 16;
    a.scale = scale;
    a.d_mid = (float)(cnt + 2);
    a.d_edge = (float)(cnt + 1);
    a.y_mid = rcp_of(cnt + 2, tab);
    a.y_edge = rcp_of(cnt + 1, tab);
    float prev = -0.0f;
    const int g = column_run<PG, false>(a, 0, prev);
    if (g < a.Dg) column_run<PG, true>(a, g, prev);
}

// ---- fused sweep ----------------------------------------------------------
// One whole sweep (the red pass and the black pass) per launch, reading the
// old pressure from one plane set (src) and writing the new one to the other
// (dst), at 12 bytes per cell instead of the 16 of two in-place passes.
//
// A wave covers 64 consecutive cells i0..i0+63 of row j, lane = cell column,
// and walks slab groups g = 0..Dg-1. Per slab each lane evaluates three
// update chains, each with the reference operation order of its cell:
//   red lane   c1 = its own red cell;  c2 = the red cell just outside the
//              wave that the wave's edge black lane needs (lane 63 when lane
//              0 is black: (i0-1, j); lane 0 when lane 63 is black:
//              (i0+64, j)), otherwise a copy of c1;  c3 unused.
//   black lane c1 = red (i, j-1);  c2 = red (i, j+1);  c3 = its own black
//              cell, from the new reds (i-1, j), (i+1, j) of the neighbour
//              lanes (ds_bpermute; wave edges take the other edge's c2),
//              c1, c2, its own new k-1 and old k+1 values.
// Red values of a neighbouring row are evaluated again by the lanes that
// need them, from the same old black values in the same order, so every
// value equals the two-pass sweep's (bitwise); only a lane's own cell is
// stored. The old values stay untouched in src, so no lane can see a new
// value early, whatever the schedule. Lanes with i >= W write -0 into the
// ghost pair of their row (keeping it -0).
//
// Loads: 14 per group (6 per red-type chain: rhs, own old value for k+1,
// x-1, x+1, y-1, y+1 black neighbours; the black chain's rhs and own old
// value), one store; PG groups in flight (vmcnt bound per slot below).
// Division: q0 = num*y, r = fma(-q0, d, num), q = fma(r, y, q0) equals IEEE
// num / d for num = +0 and 2^-125 <= |num| <= FLT_MAX (exhaustively checked,
// blz_selftest_division); any other numerator of a chain whose value is used
// sends the wave to the exact run from that group (ping-pong: the inputs
// are still unmodified).
constexpr int FNL = 14; // loads per group
constexpr int FNU = 3;  // of which the k+1 sources, issued first
enum { F_PR1 = 0, F_PR2, F_PB3, F_RR1, F_XL1, F_XR1, F_YN1, F_YS1, F_RR2, F_XL2, F_XR2, F_YN2, F_YS2, F_RB3 };

template <int PG, int U, bool FIRST>
struct FusedWait {
    static constexpr int N = FIRST && U < PG - 1 ? (FNL - FNU) + (PG - 2 - U) * FNL + U * (1 + FNL)
                                                 : (FNL - FNU) + (PG - 2) * (FNL + 1);
};

struct FusedArgs {
    i4v rs;
    int vo[FNL];        // per-lane byte offsets of each load in group 0 (plane excluded)
    int vst;            // per-lane byte offset of the store (plane included)
    unsigned pl[FNL];   // plane byte offset of each load
    int D, Dg, gstride;
    float scale;
    float dm[3], ym[3], de[3], ye[3]; // denominators/reciprocals per chain, mid and k-edge
    int lane;
    unsigned active, isRed;
    unsigned need1, need2, need3;        // chain results that are used (0/1)
    unsigned okL, okR, okN, okS;         // neighbour cells of the black update inside the grid
};

template <int PG>
__device__ __forceinline__ void fused_load(const FusedArgs& a, f4v (&sl)[FNL], int g) {
    const int so = min(g, a.Dg) * a.gstride;
#pragma unroll
    for (int q = 0; q < FNL; ++q) sl[q] = ld4_async(a.vo[q], a.rs, (int)(a.pl[q] + (unsigned)so));
}

// wait for a slot's loads and the next slot's k+1 sources (first FNU loads);
// the registers are in/out operands of the wait itself
template <int N>
__device__ __forceinline__ void fused_wait(f4v (&cur)[FNL], f4v (&nxt)[FNL]) {
    asm volatile("s_waitcnt vmcnt(%17)"
                 : "+v"(cur[0]), "+v"(cur[1]), "+v"(cur[2]), "+v"(cur[3]), "+v"(cur[4]), "+v"(cur[5]), "+v"(cur[6]),
                   "+v"(cur[7]), "+v"(cur[8]), "+v"(cur[9]), "+v"(cur[10]), "+v"(cur[11]), "+v"(cur[12]),
                   "+v"(cur[13]), "+v"(nxt[0]), "+v"(nxt[1]), "+v"(nxt[2])
                 : "n"(N));
}

__device__ __forceinline__ float fdiv_fast(float num, float d, float y) {
    const float q0 = num * y;
    const float r = fmaf(-q0, d, num);
    return fmaf(r, y, q0);
}
__device__ __forceinline__ unsigned fdiv_unsafe(float num) {
    const float mag = fabsf(num);
    const unsigned ok = (unsigned)(mag >= 0x1p-125f) & (unsigned)(mag <= 3.40282347e38f);
    const unsigned pz = (unsigned)(__float_as_uint(num) == 0u);
    return (ok | pz) ^ 1u;
}

__device__ __forceinline__ float lane_pull(float v, int src_lane) {
    return __int_as_float(__builtin_amdgcn_ds_bpermute(src_lane << 2, __float_as_int(v)));
}

template <int PG, int U, bool EXACT, bool FIRST>
__device__ __forceinline__ bool fused_group(const FusedArgs& a, f4v (&sl)[PG][FNL], int g, float (&prev)[3]) {
    constexpr int UN = (U + 1) % PG;
    fused_wait<FusedWait<PG, U, FIRST>::N>(sl[U], sl[UN]);
    const int D = a.D;
    const int k0 = 4 * g;
    const f4v* c = sl[U];
    f4v out;
    unsigned bad = 0u;
    float p1 = prev[0], p2 = prev[1], p3 = prev[2];
    const int lm = (a.lane + 63) & 63, lp = (a.lane + 1) & 63;
    // phase 1: the two red-type chains over the group's four slabs
    float r1[4], r2[4], n1[4], n2[4];
#pragma unroll
    for (int s = 0; s < 4; ++s) {
        const int k = k0 + s;
        const bool edge = (k == 0) | (k == D - 1);
        const float up1 = s < 3 ? c[F_PR1][s + 1] : sl[UN][F_PR1][0];
        const float up2 = s < 3 ? c[F_PR2][s + 1] : sl[UN][F_PR2][0];
        float x = a.scale * c[F_RR1][s];
        x = x + c[F_XL1][s];
        x = x + c[F_XR1][s];
        x = x + c[F_YN1][s];
        x = x + c[F_YS1][s];
        n1[s] = x + p1;
        n1[s] = n1[s] + up1;
        x = a.scale * c[F_RR2][s];
        x = x + c[F_XL2][s];
        x = x + c[F_XR2][s];
        x = x + c[F_YN2][s];
        x = x + c[F_YS2][s];
        n2[s] = x + p2;
        n2[s] = n2[s] + up2;
        if (EXACT) {
            r1[s] = div_small(n1[s], edge ? a.de[0] : a.dm[0], edge ? a.ye[0] : a.ym[0]);
            r2[s] = div_small(n2[s], edge ? a.de[1] : a.dm[1], edge ? a.ye[1] : a.ym[1]);
        } else {
            r1[s] = fdiv_fast(n1[s], edge ? a.de[0] : a.dm[0], edge ? a.ye[0] : a.ym[0]);
            r2[s] = fdiv_fast(n2[s], edge ? a.de[1] : a.dm[1], edge ? a.ye[1] : a.ym[1]);
        }
        p1 = r1[s];
        p2 = r2[s];
    }
    // phase 2: new reds beside the black cell: lane -+ 1's chain 1, across
    // the wave edge the opposite edge lane's chain 2
    float Lv[4], Rv[4];
#pragma unroll
    for (int s = 0; s < 4; ++s) {
        Lv[s] = lane_pull(a.lane == 63 ? r2[s] : r1[s], lm);
        Rv[s] = lane_pull(a.lane == 0 ? r2[s] : r1[s], lp);
    }
    // one wait for all eight exchanges instead of one per slab
    asm volatile("" : "+v"(Lv[0]), "+v"(Lv[1]), "+v"(Lv[2]), "+v"(Lv[3]), "+v"(Rv[0]), "+v"(Rv[1]), "+v"(Rv[2]),
                 "+v"(Rv[3]));
    // phase 3: the black chain
#pragma unroll
    for (int s = 0; s < 4; ++s) {
        const int k = k0 + s;
        const bool edge = (k == 0) | (k == D - 1);
        const float up3 = s < 3 ? c[F_PB3][s + 1] : sl[UN][F_PB3][0];
        const float vL = a.okL ? Lv[s] : -0.0f, vR = a.okR ? Rv[s] : -0.0f;
        const float vN = a.okN ? r1[s] : -0.0f, vS = a.okS ? r2[s] : -0.0f;
        float x = a.scale * c[F_RB3][s];
        x = x + vL;
        x = x + vR;
        x = x + vN;
        x = x + vS;
        float n3 = x + p3;
        n3 = n3 + up3;
        float r3;
        if (EXACT) {
            r3 = div_small(n3, edge ? a.de[2] : a.dm[2], edge ? a.ye[2] : a.ym[2]);
        } else {
            r3 = fdiv_fast(n3, edge ? a.de[2] : a.dm[2], edge ? a.ye[2] : a.ym[2]);
            const unsigned b = (a.need1 & fdiv_unsafe(n1[s])) | (a.need2 & fdiv_unsafe(n2[s])) |
                               (a.need3 & fdiv_unsafe(n3));
            bad |= (k < D) ? b : 0u;
        }
        const float mine = a.isRed ? r1[s] : r3;
        out[s] = (k < D && a.active) ? mine : -0.0f;
        p3 = r3;
    }
    if (!EXACT && __builtin_amdgcn_ballot_w64(bad != 0u)) return false;
    prev[0] = p1;
    prev[1] = p2;
    prev[2] = p3;
    st4_async(out, a.vst, a.rs, min(g, a.Dg) * a.gstride);
    fused_load<PG>(a, sl[U], g + PG);
    return true;
}

template <int PG, bool EXACT, bool FIRST, int U>
__device__ __forceinline__ bool fused_ring(const FusedArgs& a, f4v (&sl)[PG][FNL], int gb, float (&prev)[3],
                                           int* failed) {
    if constexpr (U == PG) {
        return true;
    } else {
        if (!fused_group<PG, U, EXACT, FIRST>(a, sl, gb + U, prev)) {
            *failed = gb + U;
            return false;
        }
        return fused_ring<PG, EXACT, FIRST, U + 1>(a, sl, gb, prev, failed);
    }
}

__device__ __forceinline__ void hold_slot(f4v (&r)[FNL]) {
    asm volatile("s_waitcnt vmcnt(0)"
                 : "+v"(r[0]), "+v"(r[1]), "+v"(r[2]), "+v"(r[3]), "+v"(r[4]), "+v"(r[5]), "+v"(r[6]), "+v"(r[7]),
                   "+v"(r[8]), "+v"(r[9]), "+v"(r[10]), "+v"(r[11]), "+v"(r[12]), "+v"(r[13]));
}

// Walk groups g_start.. of the wave. Returns Dg, or for a fast run the first
// group where an exact division is needed (that group was not stored).
template <int PG, bool EXACT>
__device__ __forceinline__ int fused_run(const FusedArgs& a, int g_start, float (&prev)[3]) {
    f4v sl[PG][FNL];
#pragma unroll
    for (int u = 0; u < PG; ++u) fused_load<PG>(a, sl[u], g_start + u);
    int result = a.Dg;
    int failed = a.Dg;
    bool ok = fused_ring<PG, EXACT, true, 0>(a, sl, g_start, prev, &failed);
    for (int gb = g_start + PG; ok && gb < a.Dg; gb += PG) ok = fused_ring<PG, EXACT, false, 0>(a, sl, gb, prev, &failed);
    if (!ok) result = failed;
#pragma unroll
    for (int u = 0; u < PG; ++u) hold_slot(sl[u]);
    return result;
}

__device__ __forceinline__ int cell_cnt(int i, int j, int W, int H) {
    return (int)(i > 0) + (int)(i < W - 1) + (int)(j > 0) + (int)(j < H - 1);
}

// one wave (block) wv of a fused sweep of one grid
template <int PG>
__device__ __forceinline__ void fused_wave(float* __restrict__ base, unsigned nbytes, unsigned srcR, unsigned srcB,
                                           unsigned rhsR, unsigned rhsB, unsigned dstR, unsigned dstB, int W, int H,
                                           int D, int Wp, int Hp, float scale, const DivTab& tab, int wv) {
    const int nwr = (W + 63) >> 6; // waves per row
    const int j = wv / nwr;
    const int i0 = (wv - j * nwr) * 64;
    const int lane = threadIdx.x & 63;
    const int i = i0 + lane;
    const bool active = i < W;
    const bool red = ((i + j) & 1) == 0;
    // byte offset of cell (ii, jj) in padded group 1 of a plane
    auto off = [&](int ii, int jj) -> int { return 16 * (((Hp + (jj + 1)) * Wp) + ((ii >> 1) + 1)); };
    FusedArgs a;
    const unsigned long long b64 = (unsigned long long)base;
    a.rs[0] = (int)(unsigned)(b64 & 0xffffffffull);
    a.rs[1] = (int)(unsigned)(b64 >> 32);
    a.rs[2] = (int)nbytes;
    a.rs[3] = 0x00020000;
    a.D = D;
    a.Dg = (D + 3) >> 2;
    a.gstride = Hp * Wp * 16;
    a.scale = scale;
    a.lane = lane;
    // chain cells
    const bool haloL = red && lane == 63; // lane 0 black: needs red (i0-1, j)
    const bool haloR = red && lane == 0;  // lane 63 black: needs red (i0+64, j)
    const int ic = active ? i : W - 1;    // address column of an inactive lane
    int c1i, c1j, c2i, c2j;
    if (red) {
        c1i = ic;
        c1j = j;
        // a halo cell outside the grid is not used: read the lane's own
        c2i = haloL && i0 >= 1 ? i0 - 1 : (haloR && i0 + 64 < W ? i0 + 64 : ic);
        c2j = j;
    } else {
        c1i = ic;
        c1j = j - 1;
        c2i = ic;
        c2j = j + 1;
    }
    const int ci[2] = {c1i, c2i}, cj[2] = {c1j, c2j};
#pragma unroll
    for (int q = 0; q < 2; ++q) {
        const int ii = ci[q], jj = cj[q];
        const int base_slot = q == 0 ? F_RR1 : F_RR2;
        a.vo[q == 0 ? F_PR1 : F_PR2] = off(ii, jj);
        a.pl[q == 0 ? F_PR1 : F_PR2] = srcR;
        a.vo[base_slot] = off(ii, jj);
        a.pl[base_slot] = rhsR;
        a.vo[base_slot + 1] = off(ii - 1, jj); // x-1
        a.vo[base_slot + 2] = off(ii + 1, jj); // x+1
        a.vo[base_slot + 3] = off(ii, jj - 1); // y-1
        a.vo[base_slot + 4] = off(ii, jj + 1); // y+1
        for (int t = 1; t <= 4; ++t) a.pl[base_slot + t] = srcB;
        const int cnt = cell_cnt(ii, jj, W, H);
        a.dm[q] = (float)(cnt + 2);
        a.de[q] = (float)(cnt + 1);
        a.ym[q] = rcp_of(cnt + 2, tab);
        a.ye[q] = rcp_of(cnt + 1, tab);
    }
    a.vo[F_PB3] = off(ic, j);
    a.pl[F_PB3] = srcB;
    a.vo[F_RB3] = off(ic, j);
    a.pl[F_RB3] = rhsB;
    {
        const int cnt = cell_cnt(ic, j, W, H);
        a.dm[2] = (float)(cnt + 2);
        a.de[2] = (float)(cnt + 1);
        a.ym[2] = rcp_of(cnt + 2, tab);
        a.ye[2] = rcp_of(cnt + 1, tab);
    }
    // store: own cell in its colour's dst plane; inactive lanes write -0
    // into the row's ghost pair
    a.vst = (int)((red ? dstR : dstB) + (unsigned)(active ? off(i, j) : off(W, j)));
    a.isRed = red ? 1u : 0u;
    a.active = active ? 1u : 0u;
    a.need1 = active && (red || j > 0) ? 1u : 0u;
    a.need2 = (red ? ((haloL && i0 >= 1) || (haloR && i0 + 64 < W)) : (active && j < H - 1)) ? 1u : 0u;
    a.need3 = active && !red ? 1u : 0u;
    a.okL = i >= 1 ? 1u : 0u;
    a.okR = i + 1 < W ? 1u : 0u;
    a.okN = j > 0 ? 1u : 0u;
    a.okS = j < H - 1 ? 1u : 0u;
    float prev[3] = {-0.0f, -0.0f, -0.0f};
    const int g = fused_run<PG, false>(a, 0, prev);
    if (g < a.Dg) fused_run<PG, true>(a, g, prev);
}

// WPB waves per block (independent waves; no block-level synchronisation)
template <int PG, int WPB>
__global__ void __launch_bounds__(64 * WPB) k_rbgs_fused(float* __restrict__ base, unsigned nbytes, unsigned srcR,
                                                         unsigned srcB, unsigned rhsR, unsigned rhsB, unsigned dstR,
                                                         unsigned dstB, int W, int H, int D, int Wp, int Hp,
                                                         float scale, DivTab tab, const int* __restrict__ stop,
                                                         int nwaves) {
    if (*stop) return; // the solve has converged
    const int wv = xcd_block(blockIdx.x, gridDim.x) * WPB + (int)(threadIdx.x >> 6);
    if (wv >= nwaves) return;
    fused_wave<PG>(base, nbytes, srcR, srcB, rhsR, rhsB, dstR, dstB, W, H, D, Wp, Hp, scale, tab, wv);
}

// the same sweep over several grids in one launch: block b belongs to the
// last table entry whose first_block <= b (binary search, block-uniform)
template <int PG, int WPB>
__global__ void __launch_bounds__(64 * WPB) k_rbgs_fused_multi(const FusedScene* __restrict__ sc, int n,
                                                               DivTab tab) {
    const int b = xcd_block(blockIdx.x, gridDim.x);
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (sc[mid].first_block <= b) lo = mid;
        else hi = mid - 1;
    }
    const FusedScene& e = sc[lo];
    if (*e.stop) return;
    const int wv = (b - e.first_block) * WPB + (int)(threadIdx.x >> 6);
    if (wv >= e.nwaves) return;
    fused_wave<PG>(e.base, e.nbytes, e.srcR, e.srcB, e.rhsR, e.rhsB, e.dstR, e.dstB, e.W, e.H, e.D, e.Wp, e.Hp,
                   e.scale, tab, wv);
}

// ---- row-pair fused sweep ------------------------------------------------
// One whole sweep per launch (ping-pong as above), with the rows of a pair
// (j, j+1), j even, in one wave: lanes 0-30 hold cells 30s .. 30s+30 of
// row j, lanes 32-62 cells 30s-1 .. 30s+29 of row j+1 (lanes 31 and 63
// carry no cell). In both halves the end lanes are red, so every red
// neighbour of a black cell in x is a lane of the same half, and its red
// neighbour in the other row of the pair is a lane of the other half (lane
// +-33). Segments overlap by one red cell, computed by both waves from the
// same inputs in the same order and stored by both with the same bits.
// Per slab group every lane runs two update chains:
//   chain A (red type): a red lane's own cell; a black lane's red neighbour
//           outside the pair (row j-1 for row j, row j+2 for row j+1);
//   chain C (black type): a black lane's own cell, from the new reds of
//           lanes -+1 (x), the other half and its own chain A (y), all by
//           ds_bpermute, and its own new k-1 and old k+1 values.
// Each chain follows its cell's reference operation order, so every value
// equals the two-pass sweep's bits. A chain A whose cell lies outside the
// grid publishes -0, the ghost value, so no neighbour needs a mask. 8 loads
// per group (chain A: own old value for k+1, rhs, four black neighbours;
// chain C: own old value, rhs), one store; lanes without a cell write -0
// into a ghost slot of their row. Groups away from slab 0 and slab D-1 skip
// the k-edge selects. Division and the exact fallback as in the fused
// sweep above.
constexpr int PNL = 8; // loads per group
constexpr int PNU = 2; // of which the k+1 sources, issued first
enum { P_PA = 0, P_PC, P_RA, P_XL, P_XR, P_YN, P_YS, P_RC };

template <int PG, int U, bool FIRST>
struct PairWait {
    static constexpr int N = FIRST && U < PG - 1 ? (PNL - PNU) + (PG - 2 - U) * PNL + U * (1 + PNL)
                                                 : (PNL - PNU) + (PG - 2) * (PNL + 1);
};

struct PairArgs {
    i4v rs;
    int vo[PNL];      // per-lane byte offsets of each load in group 0 (plane excluded)
    unsigned pl[PNL]; // plane byte offset of each load
    int vst;          // per-lane byte offset of the store (plane included)
    int D, Dg, gstride;
    float scale;
    float dmA, ymA, deA, yeA, dmC, ymC, deC, yeC; // denominators/reciprocals, mid and k-edge
    int srcL, srcR, srcN, srcS; // bpermute addresses of the lanes holding the x-1, x+1, y-1, y+1 reds
    int pubA;                   // chain A's cell is in the grid (else it publishes -0)
    int ownRed;                 // the lane stores chain A's published value (red cells, lanes without a cell)
    int needA, needC;           // chain results that are used
};

template <int PG>
__device__ __forceinline__ void pair_load(const PairArgs& a, f4v (&sl)[PNL], int g) {
    const int so = min(g, a.Dg) * a.gstride;
#pragma unroll
    for (int q = 0; q < PNL; ++q) sl[q] = ld4_async(a.vo[q], a.rs, (int)(a.pl[q] + (unsigned)so));
}

template <int N>
__device__ __forceinline__ void pair_wait(f4v (&cur)[PNL], f4v (&nxt)[PNL]) {
    asm volatile("s_waitcnt vmcnt(%10)"
                 : "+v"(cur[0]), "+v"(cur[1]), "+v"(cur[2]), "+v"(cur[3]), "+v"(cur[4]), "+v"(cur[5]), "+v"(cur[6]),
                   "+v"(cur[7]), "+v"(nxt[0]), "+v"(nxt[1])
                 : "n"(N));
}

__device__ __forceinline__ float lane_get(float v, int src_addr) {
    return __int_as_float(__builtin_amdgcn_ds_bpermute(src_addr, __float_as_int(v)));
}

// the fast run's division is exact for this numerator (+0 or 2^-125 <=
// |num| <= FLT_MAX; see fdiv_unsafe). A -0 numerator needs every term to be
// -0, which a solver state meets only when degenerate, so it takes the
// exact run rather than a sign copy on every division. Bitwise, not
// short-circuit, operators: lane masks without branches.
__device__ __forceinline__ bool fdiv_ok(float num) {
    const float mag = fabsf(num);
    const bool lo = mag >= 0x1p-125f, hi = mag <= 3.40282347e38f, pz = __float_as_uint(num) == 0u;
    return (lo & hi) | pz;
}

// The arithmetic of one group. BOUND: the group holds slab 0, slab D-1 or
// slabs k >= D (k-edge denominators; -0 stored for k >= D); other groups
// skip those selects. Returns false (fast run) when some lane needs the
// exact division.
template <bool EXACT, bool BOUND>
__device__ __forceinline__ bool pair_math(const PairArgs& a, const f4v* c, const f4v* n, int k0, float& prevA,
                                          float& prevC, f4v& out) {
    const int D = a.D;
    float pA = prevA, pC = prevC;
    float rA[4], nA[4], pub[4];
#pragma unroll
    for (int s = 0; s < 4; ++s) {
        const int k = k0 + s;
        const bool edge = BOUND && ((k == 0) | (k == D - 1));
        const float up = s < 3 ? c[P_PA][s + 1] : n[P_PA][0];
        float x = a.scale * c[P_RA][s];
        x = x + c[P_XL][s];
        x = x + c[P_XR][s];
        x = x + c[P_YN][s];
        x = x + c[P_YS][s];
        nA[s] = x + pA;
        nA[s] = nA[s] + up;
        if (EXACT) rA[s] = div_small(nA[s], edge ? a.deA : a.dmA, edge ? a.yeA : a.ymA);
        else rA[s] = fdiv_fast(nA[s], edge ? a.deA : a.dmA, edge ? a.yeA : a.ymA);
        pA = rA[s];
        pub[s] = a.pubA != 0 ? rA[s] : -0.0f; // -0 (the ghost value) for cells outside the grid
    }
    float Lv[4], Rv[4], Nv[4], Sv[4];
#pragma unroll
    for (int s = 0; s < 4; ++s) {
        Lv[s] = lane_get(pub[s], a.srcL);
        Rv[s] = lane_get(pub[s], a.srcR);
        Nv[s] = lane_get(pub[s], a.srcN);
        Sv[s] = lane_get(pub[s], a.srcS);
    }
    asm volatile("" : "+v"(Lv[0]), "+v"(Lv[1]), "+v"(Lv[2]), "+v"(Lv[3]), "+v"(Rv[0]), "+v"(Rv[1]), "+v"(Rv[2]),
                 "+v"(Rv[3]), "+v"(Nv[0]), "+v"(Nv[1]), "+v"(Nv[2]), "+v"(Nv[3]), "+v"(Sv[0]), "+v"(Sv[1]),
                 "+v"(Sv[2]), "+v"(Sv[3]));
    const bool mA = a.needA != 0, mC = a.needC != 0; // lane masks, loop-invariant
    bool bad = false;                                 // the lane needs the exact division
#pragma unroll
    for (int s = 0; s < 4; ++s) {
        const int k = k0 + s;
        const bool edge = BOUND && ((k == 0) | (k == D - 1));
        const float up = s < 3 ? c[P_PC][s + 1] : n[P_PC][0];
        float x = a.scale * c[P_RC][s];
        x = x + Lv[s];
        x = x + Rv[s];
        x = x + Nv[s];
        x = x + Sv[s];
        float nC = x + pC;
        nC = nC + up;
        float rC;
        if (EXACT) {
            rC = div_small(nC, edge ? a.deC : a.dmC, edge ? a.yeC : a.ymC);
        } else {
            rC = fdiv_fast(nC, edge ? a.deC : a.dmC, edge ? a.yeC : a.ymC);
            const bool okA = fdiv_ok(nA[s]), okC = fdiv_ok(nC);
            if (!BOUND || k < D) bad = bad | (mA & !okA) | (mC & !okC);
        }
        const float v = a.ownRed != 0 ? pub[s] : rC;
        out[s] = (BOUND && k >= D) ? -0.0f : v;
        pC = rC;
    }
    if (!EXACT && __builtin_amdgcn_ballot_w64(bad)) return false;
    prevA = pA;
    prevC = pC;
    return true;
}

template <int PG, int U, bool EXACT, bool FIRST>
__device__ __forceinline__ bool pair_group(const PairArgs& a, f4v (&sl)[PG][PNL], int g, float& prevA, float& prevC) {
    constexpr int UN = (U + 1) % PG;
    pair_wait<PairWait<PG, U, FIRST>::N>(sl[U], sl[UN]);
    f4v out;
    const bool bound = g == 0 || 4 * g + 4 >= a.D; // wave-uniform
    const bool done = bound ? pair_math<EXACT, true>(a, sl[U], sl[UN], 4 * g, prevA, prevC, out)
                            : pair_math<EXACT, false>(a, sl[U], sl[UN], 4 * g, prevA, prevC, out);
    if (!done) return false;
    st4_async(out, a.vst, a.rs, min(g, a.Dg) * a.gstride);
    pair_load<PG>(a, sl[U], g + PG);
    return true;
}

template <int PG, bool EXACT, bool FIRST, int U>
__device__ __forceinline__ bool pair_ring(const PairArgs& a, f4v (&sl)[PG][PNL], int gb, float& pA, float& pC,
                                          int* failed) {
    if constexpr (U == PG) {
        return true;
    } else {
        if (!pair_group<PG, U, EXACT, FIRST>(a, sl, gb + U, pA, pC)) {
            *failed = gb + U;
            return false;
        }
        return pair_ring<PG, EXACT, FIRST, U + 1>(a, sl, gb, pA, pC, failed);
    }
}

__device__ __forceinline__ void hold_pair_slot(f4v (&r)[PNL]) {
    asm volatile("s_waitcnt vmcnt(0)"
                 : "+v"(r[0]), "+v"(r[1]), "+v"(r[2]), "+v"(r[3]), "+v"(r[4]), "+v"(r[5]), "+v"(r[6]), "+v"(r[7]));
}

// Walk groups g_start.. of the wave. Returns Dg, or for a fast run the first
// group where an exact division is needed (that group was not stored).
template <int PG, bool EXACT>
__device__ __forceinline__ int pair_run(const PairArgs& a, int g_start, float& pA, float& pC) {
    f4v sl[PG][PNL];
#pragma unroll
    for (int u = 0; u < PG; ++u) pair_load<PG>(a, sl[u], g_start + u);
    int result = a.Dg;
    int failed = a.Dg;
    bool ok = pair_ring<PG, EXACT, true, 0>(a, sl, g_start, pA, pC, &failed);
    for (int gb = g_start + PG; ok && gb < a.Dg; gb += PG) ok = pair_ring<PG, EXACT, false, 0>(a, sl, gb, pA, pC, &failed);
    if (!ok) result = failed;
#pragma unroll
    for (int u = 0; u < PG; ++u) hold_pair_slot(sl[u]);
    return result;
}

// waves of a row-pair sweep: 30-cell segments per row pair
__host__ __device__ __forceinline__ int pair_segments(int W) { return (W + 29) / 30; }

// one wave wv of a row-pair fused sweep
template <int PG>
__device__ __forceinline__ void pair_wave(float* __restrict__ base, unsigned nbytes, unsigned srcR, unsigned srcB,
                                          unsigned rhsR, unsigned rhsB, unsigned dstR, unsigned dstB, int W, int H,
                                          int D, int Wp, int Hp, float scale, const DivTab& tab, int wv) {
    const int nseg = pair_segments(W);
    const int jp = wv / nseg, sg = wv - jp * nseg;
    const int lane = threadIdx.x & 63;
    const int half = lane >> 5, l = lane & 31;
    const int jr = 2 * jp + half;     // the lane's row
    const int i = 30 * sg - half + l; // the lane's cell column
    const bool active = l < 31 && i >= 0 && i < W && jr < H;
    const bool red = ((i + jr) & 1) == 0;
    auto off = [&](int ii, int jj) -> int { return 16 * (((Hp + (jj + 1)) * Wp) + ((ii >> 1) + 1)); };
    PairArgs a;
    const unsigned long long b64 = (unsigned long long)base;
    a.rs[0] = (int)(unsigned)(b64 & 0xffffffffull);
    a.rs[1] = (int)(unsigned)(b64 >> 32);
    a.rs[2] = (int)nbytes;
    a.rs[3] = 0x00020000;
    a.D = D;
    a.Dg = (D + 3) >> 2;
    a.gstride = Hp * Wp * 16;
    a.scale = scale;
    // addressed cells, clamped into the grid (a clamped chain's result is
    // never used)
    const int ic = min(max(i, 0), W - 1);
    const int jc = min(jr, H - 1);
    const int ja = red ? jc : min(max(half == 0 ? jr - 1 : jr + 1, 0), H - 1);
    a.vo[P_PA] = off(ic, ja);
    a.pl[P_PA] = srcR;
    a.vo[P_RA] = off(ic, ja);
    a.pl[P_RA] = rhsR;
    a.vo[P_XL] = off(ic - 1, ja);
    a.vo[P_XR] = off(ic + 1, ja);
    a.vo[P_YN] = off(ic, ja - 1);
    a.vo[P_YS] = off(ic, ja + 1);
    a.pl[P_XL] = a.pl[P_XR] = a.pl[P_YN] = a.pl[P_YS] = srcB;
    a.vo[P_PC] = off(ic, jc);
    a.pl[P_PC] = srcB;
    a.vo[P_RC] = off(ic, jc);
    a.pl[P_RC] = rhsB;
    {
        const int cnt = cell_cnt(ic, ja, W, H);
        a.dmA = (float)(cnt + 2);
        a.deA = (float)(cnt + 1);
        a.ymA = rcp_of(cnt + 2, tab);
        a.yeA = rcp_of(cnt + 1, tab);
    }
    {
        const int cnt = cell_cnt(ic, jc, W, H);
        a.dmC = (float)(cnt + 2);
        a.deC = (float)(cnt + 1);
        a.ymC = rcp_of(cnt + 2, tab);
        a.yeC = rcp_of(cnt + 1, tab);
    }
    // neighbour reds: x from lanes -+1 of the half; in y the lane's own
    // chain A on the side outside the pair, the other half's lane across
    const int partner = half == 0 ? min(lane + 33, 63) : max(lane - 33, 0);
    a.srcL = 4 * max(lane - 1, 0);
    a.srcR = 4 * min(lane + 1, 63);
    a.srcN = 4 * (half == 0 ? lane : partner);
    a.srcS = 4 * (half == 0 ? partner : lane);
    // a neighbour outside the grid is a lane without a cell (x: cell -1 or
    // W, or lanes 31/63; y: the missing row of the pair) or the lane's own
    // chain A on a row outside the grid: all publish -0
    a.needA = active && (red || (half == 0 ? jr > 0 : jr < H - 1));
    a.pubA = a.needA;
    a.needC = active && !red;
    a.ownRed = red || !active;
    // store: the lane's cell in its colour's dst plane (the red cell of a
    // segment overlap is stored by both waves with the same bits); lanes
    // without a cell write -0 into the ghost slot past the end of their row
    a.vst = (int)(active ? (red ? dstR : dstB) + (unsigned)off(i, jr) : dstR + (unsigned)off(W, jc));
    float pA = -0.0f, pC = -0.0f;
    const int g = pair_run<PG, false>(a, 0, pA, pC);
    if (g < a.Dg) pair_run<PG, true>(a, g, pA, pC);
}

template <int PG, int WPB>
__global__ void __launch_bounds__(64 * WPB) k_rbgs_pair(float* __restrict__ base, unsigned nbytes, unsigned srcR,
                                                        unsigned srcB, unsigned rhsR, unsigned rhsB, unsigned dstR,
                                                        unsigned dstB, int W, int H, int D, int Wp, int Hp,
                                                        float scale, DivTab tab, const int* __restrict__ stop,
                                                        int nwaves) {
    if (*stop) return; // the solve has converged
    const int wv = xcd_block(blockIdx.x, gridDim.x) * WPB + (int)(threadIdx.x >> 6);
    if (wv >= nwaves) return;
    pair_wave<PG>(base, nbytes, srcR, srcB, rhsR, rhsB, dstR, dstB, W, H, D, Wp, Hp, scale, tab, wv);
}

// the row-pair sweep over several grids in one launch (table as for
// k_rbgs_fused_multi, nwaves counted in row-pair waves)
template <int PG, int WPB>
__global__ void __launch_bounds__(64 * WPB) k_rbgs_pair_multi(const FusedScene* __restrict__ sc, int n, DivTab tab) {
    const int b = xcd_block(blockIdx.x, gridDim.x);
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (sc[mid].first_block <= b) lo = mid;
        else hi = mid - 1;
    }
    const FusedScene& e = sc[lo];
    if (*e.stop) return;
    const int wv = (b - e.first_block) * WPB + (int)(threadIdx.x >> 6);
    if (wv >= e.nwaves) return;
    pair_wave<PG>(e.base, e.nbytes, e.srcR, e.srcB, e.rhsR, e.rhsB, e.dstR, e.dstB, e.W, e.H, e.D, e.Wp, e.Hp,
                  e.scale, tab, wv);
}

__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
    return v;
}

__device__ __forceinline__ double cell_res(double pc, double rc, double sum, double n, double c) {
    double r = fabs(rc + c * (sum - n * pc));
    return r != r ? __longlong_as_double(0x7ff0000000000000LL) : r;
}

// ||div u||_inf proxy: max over cells not adjacent to an open face of
// |rhs + c (sum_nbr p - n p)|, c = dt/(rho dx^2), all in FP64; NaN -> +inf.
// The neighbour sum runs in the order i-1, i+1, j-1, j+1, k-1, k+1 over all
// six slots: a missing neighbour is a -0 ghost, the identity of the sum, so
// it equals summing the existing neighbours only; n counts those.
// Thread per (pair column m, row j, segment of RES_SEG slab groups): per
// group, ten coalesced 16-byte loads cover eight cells (the red/black pair
// for four slabs); z neighbours come from the previous and next group of the
// thread's own column. Grid-stride; one partial maximum per block, reduced
// by k_res_final. `stop` is the solve's done flag: a converged solve skips.
constexpr int RES_SEG = 2;

// block blk of nblk blocks of the residual of one grid
__device__ __forceinline__ void residual_block(const RB& p, const RB& rhs, const RBL& L, int W, int H, int D,
                                               double c, int4 cbA, int2 cbB, double* __restrict__ partial, int blk,
                                               int nblk) {
    typedef float f4 __attribute__((ext_vector_type(4)));
    const int Wh = W >> 1;
    const int nseg = (L.Dg + RES_SEG - 1) / RES_SEG;
    const int total = Wh * H * nseg;
    const int gs = L.Wp * L.Hp; // float4 elements per group
    const bool openL = !cbA.x, openR = !cbA.y, openN = !cbA.w, openS = !cbA.z, openD = !cbB.x, openU = !cbB.y;
    double mx = 0.0;
    for (int t = blk * BS + threadIdx.x; t < total; t += nblk * BS) {
        const int m = t % Wh, j = (t / Wh) % H, seg = t / (Wh * H);
        const bool odd = j & 1; // colours of i0 = 2m and i1 = 2m+1: (odd, !odd)
        const BLZ_GLOBAL f4* P0 = as_global((const f4*)(odd ? p.c[1] : p.c[0]));
        const BLZ_GLOBAL f4* P1 = as_global((const f4*)(odd ? p.c[0] : p.c[1]));
        const BLZ_GLOBAL f4* R0 = as_global((const f4*)(odd ? rhs.c[1] : rhs.c[0]));
        const BLZ_GLOBAL f4* R1 = as_global((const f4*)(odd ? rhs.c[0] : rhs.c[1]));
        const int e0 = (j + 1) * L.Wp + (m + 1) + gs; // element of group 0
        const int i0 = 2 * m, i1 = 2 * m + 1;
        const double nxy0 = (double)((int)(i0 > 0) + 1 + (int)(j > 0) + (int)(j < H - 1));
        const double nxy1 = (double)(1 + (int)(i1 < W - 1) + (int)(j > 0) + (int)(j < H - 1));
        const bool skipxy0 = (i0 == 0 && openL) || (j == 0 && openN) || (j == H - 1 && openS);
        const bool skipxy1 = (i1 == W - 1 && openR) || (j == 0 && openN) || (j == H - 1 && openS);
        const int g0 = seg * RES_SEG, g1 = min(g0 + RES_SEG, L.Dg);
        f4 prv0 = P0[e0 + (g0 - 1) * gs], prv1 = P1[e0 + (g0 - 1) * gs];
        f4 cur0 = P0[e0 + g0 * gs], cur1 = P1[e0 + g0 * gs];
        for (int g = g0; g < g1; ++g) {
            const int e = e0 + g * gs;
            const f4 nxt0 = P0[e + gs], nxt1 = P1[e + gs]; // group g+1 (ghost group after the last)
            const f4 l1 = P1[e - 1], r0 = P0[e + 1];
            const f4 n1 = P1[e - L.Wp], s1 = P1[e + L.Wp];
            const f4 n0 = P0[e - L.Wp], s0 = P0[e + L.Wp];
            const f4 q0 = R0[e], q1 = R1[e];
#pragma unroll
            for (int sl = 0; sl < 4; ++sl) {
                const int k = 4 * g + sl;
                if (k >= D) break;
                const float dn0 = sl > 0 ? cur0[sl - 1] : prv0[3], up0 = sl < 3 ? cur0[sl + 1] : nxt0[0];
                const float dn1 = sl > 0 ? cur1[sl - 1] : prv1[3], up1 = sl < 3 ? cur1[sl + 1] : nxt1[0];
                const double nz = (double)((int)(k > 0) + (int)(k < D - 1));
                const bool skipz = (k == 0 && openD) || (k == D - 1 && openU);
                if (!(skipz || skipxy0)) {
                    double sum = 0.0;
                    sum += (double)l1[sl];
                    sum += (double)cur1[sl];
                    sum += (double)n1[sl];
                    sum += (double)s1[sl];
                    sum += (double)dn0;
                    sum += (double)up0;
                    mx = fmax(mx, cell_res((double)cur0[sl], (double)q0[sl], sum, nxy0 + nz, c));
                }
                if (!(skipz || skipxy1)) {
                    double sum = 0.0;
                    sum += (double)cur0[sl];
                    sum += (double)r0[sl];
                    sum += (double)n0[sl];
                    sum += (double)s0[sl];
                    sum += (double)dn1;
                    sum += (double)up1;
                    mx = fmax(mx, cell_res((double)cur1[sl], (double)q1[sl], sum, nxy1 + nz, c));
                }
            }
            prv0 = cur0;
            prv1 = cur1;
            cur0 = nxt0;
            cur1 = nxt1;
        }
    }
    __shared__ double sm[BS / 64];
    mx = wave_max(mx);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) sm[wid] = mx;
    __syncthreads();
    if (threadIdx.x == 0) {
        double w = sm[0];
        for (int q = 1; q < (int)(BS / 64); ++q) w = fmax(w, sm[q]);
        as_global(partial)[blk] = w;
    }
}

__global__ void __launch_bounds__(BS) k_residual(RB p, RB rhs, RBL L, int W, int H, int D, double c, int4 cbA,
                                                 int2 cbB, double* __restrict__ partial, const int* __restrict__ stop) {
    if (*stop) return;
    residual_block(p, rhs, L, W, H, D, c, cbA, cbB, partial, blockIdx.x, gridDim.x);
}

// a batch's check that an adaptive solve skips (not due yet)
__device__ __forceinline__ bool check_skipped(const SolveCtl* c, int add, int cap, int first) {
    return !first && c->adaptive && c->iters + add < c->next_check && c->iters + add < cap;
}

// the residual check of a solve: skipped when its solve is done or (adaptive
// spacing) the check is not due
__global__ void __launch_bounds__(BS) k_residual_chk(RB p, RB rhs, RBL L, int W, int H, int D, double c, int4 cbA,
                                                     int2 cbB, double* __restrict__ partial,
                                                     const SolveCtl* __restrict__ ctl, int add, int cap, int first) {
    if (ctl->done || check_skipped(ctl, add, cap, first)) return;
    residual_block(p, rhs, L, W, H, D, c, cbA, cbB, partial, blockIdx.x, gridDim.x);
}

// the residual of several grids in one launch (table as for the fused sweep)
__global__ void __launch_bounds__(BS) k_residual_multi(const ResScene* __restrict__ sc, int n) {
    const int b = blockIdx.x;
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (sc[mid].first_block <= b) lo = mid;
        else hi = mid - 1;
    }
    const ResScene& e = sc[lo];
    if (*e.gate) return;
    residual_block(e.p, e.rhs, e.L, e.W, e.H, e.D, e.c, make_int4(e.cb[0], e.cb[1], e.cb[2], e.cb[3]),
                   make_int2(e.cb[4], e.cb[5]), e.partial, b - e.first_block, e.nblocks);
}

// single block: the residual is the max of the partial maxima; advance the
// solve's control block as the loop `while (r > tol && it < cap)` does
// On completion the solve's sequence number goes to host_done (mapped host
// memory), which the host polls to stop enqueuing sweeps.
__device__ __forceinline__ void res_final_block(const double* __restrict__ partial, int n, SolveCtl* c,
                                                int* host_done, int seq, double tol, int cap, int add, int first) {
    if (c->done) return;
    if (check_skipped(c, add, cap, first)) { // not due: count the batch's sweeps only
        __syncthreads();                       // every thread has read c
        if (threadIdx.x == 0) {
            c->iters += add;
            c->swept += add;
        }
        return;
    }
    double mx = 0.0;
    for (int q = threadIdx.x; q < n; q += BS) mx = fmax(mx, partial[q]);
    __shared__ double sm[BS / 64];
    mx = wave_max(mx);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) sm[wid] = mx;
    __syncthreads();
    if (threadIdx.x == 0) {
        double r = sm[0];
        for (int q = 1; q < (int)(BS / 64); ++q) r = fmax(r, sm[q]);
        const double prev = c->res;
        if (first) {
            c->res0 = r;
            c->iters = 0;
            c->gap = c->every;
            c->stall_it = 0;
        } else {
            c->iters += add;
            c->swept += add;
        }
        c->res_prev = first ? -1.0 : prev;
        c->res = r;
        c->checks += 1;
        trace_push(c, BLZ_TRACE_CHECK, r);
        // an FP32 fixed point: the same residual at every check over
        // STALL_SWEEPS sweeps
        bool stalled = false;
        if (c->stop_stall && !first) {
            if (r == prev) stalled = c->iters - c->stall_it >= STALL_SWEEPS;
            else c->stall_it = c->iters;
        }
        if (c->adaptive) {
            if (!first) c->gap = next_check_gap(c->gap, c->every, r, prev, tol);
            c->next_check = c->iters + c->gap;
        }
        // a non-finite state (residual +inf) ends the solve at once: the
        // host reports it after the step instead of sweeping to the cap
        if (!(r > tol) || c->iters >= cap || stalled || r == __longlong_as_double(0x7ff0000000000000LL)) {
            c->done = 1;
            __hip_atomic_store(host_done, seq, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
            __threadfence_system();
        }
    }
}

__global__ void __launch_bounds__(BS) k_res_final(const double* __restrict__ partial, int n, SolveCtl* c,
                                                  int* host_done, int seq, double tol, int cap, int add, int first) {
    res_final_block(partial, n, c, host_done, seq, tol, cap, add, first);
}

// block q finishes grid q of a multi-grid residual launch
__global__ void __launch_bounds__(BS) k_res_final_multi(const ResScene* __restrict__ sc, int add, int first) {
    const ResScene& e = sc[blockIdx.x];
    res_final_block(e.partial, e.nblocks, e.ctl, e.host_done, e.seq, e.tol, e.cap, add, first);
}

__device__ __forceinline__ void ctl_reset(SolveCtl* c, blz_trace_entry* trace, int trace_cap, int adaptive,
                                          int stop_stall, int every) {
    c->adaptive = adaptive;
    c->stop_stall = stop_stall;
    c->every = every;
    c->gap = every;
    c->next_check = 0;
    c->stall_it = 0;
    c->trace = trace;
    c->trace_cap = trace_cap;
    c->ntrace = 0;
    c->swept = 0;
    c->done = 0;
    c->iters = 0;
    c->checks = 0;
    c->ws = 0;
    c->res = -1.0;
    c->res0 = -1.0;
    c->res_ws = -1.0;
    c->ws_kept = 0;
    c->ws_rejected = 0;
    c->res_prev = -1.0;
    c->skip_spare = 1;
}

__global__ void k_ctl_reset(SolveCtl* c, blz_trace_entry* trace, int trace_cap, int adaptive, int stop_stall,
                            int every) {
    ctl_reset(c, trace, trace_cap, adaptive, stop_stall, every);
}

// the control blocks of a table, one thread each
__global__ void __launch_bounds__(64) k_ctl_reset_multi(const CtlJob* __restrict__ tab, int n) {
    const int t = (int)(blockIdx.x * 64 + threadIdx.x);
    if (t >= n) return;
    const CtlJob& e = tab[t];
    ctl_reset(e.c, e.trace, e.trace_cap, e.adaptive, e.stop_stall, e.every);
}

// pressure_gradient_update_kernel :601-669 for the faces along AXIS,
// out-of-place: interior faces subtract s*(p_hi - p_lo); boundary faces are
// 0 when closed, else the updated value of the adjacent interior face.
template <int AXIS>
__device__ __forceinline__ void pupdate_face(const float* __restrict__ in, float* __restrict__ out, const RB& p,
                                             const RBL& L, int W, int H, int D, float sc, int closedLo, int closedHi,
                                             int t) {
    const int fx = W + (AXIS == 0), fy = H + (AXIS == 1), fz = D + (AXIS == 2);
    if (t >= fx * fy * fz) return;
    const int i = t % fx, j = (t / fx) % fy, k = t / (fx * fy);
    const int f = AXIS == 0 ? i : (AXIS == 1 ? j : k);
    const int nA = AXIS == 0 ? W : (AXIS == 1 ? H : D);
    const int sF = AXIS == 0 ? 1 : (AXIS == 1 ? fx : fx * fy); // face stride along the axis
    int g, fi = i, fj = j, fk = k; // face whose updated value is taken
    if (f == 0) {
        if (closedLo) { as_global(out)[t] = 0.0f; return; }
        g = t + sF;
        if (AXIS == 0) fi = 1; else if (AXIS == 1) fj = 1; else fk = 1;
    } else if (f == nA) {
        if (closedHi) { as_global(out)[t] = 0.0f; return; }
        g = t - sF;
        if (AXIS == 0) fi = nA - 1; else if (AXIS == 1) fj = nA - 1; else fk = nA - 1;
    } else {
        g = t;
    }
    // interior face (fi,fj,fk) lies between cells f-1 and f along AXIS
    const float p_hi = p_at(p, L, fi, fj, fk);
    const float p_lo = AXIS == 0 ? p_at(p, L, fi - 1, fj, fk)
                                 : (AXIS == 1 ? p_at(p, L, fi, fj - 1, fk) : p_at(p, L, fi, fj, fk - 1));
    as_global(out)[t] = as_global(in)[g] - sc * (p_hi - p_lo);
}

template <int AXIS>
__global__ void __launch_bounds__(BS) k_pupdate(const float* __restrict__ in, float* __restrict__ out, RB p, RBL L,
                                                int W, int H, int D, float sc, int closedLo, int closedHi) {
    pupdate_face<AXIS>(in, out, p, L, W, H, D, sc, closedLo, closedHi, (int)(blockIdx.x * BS + threadIdx.x));
}

// the job of block b in a table of first blocks (ascending)
template <class J>
__device__ __forceinline__ const J& job_of(const J* __restrict__ tab, int n, int b) {
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (tab[mid].first_block <= b) lo = mid;
        else hi = mid - 1;
    }
    return tab[lo];
}

// the divergence of every grid in a table
__global__ void __launch_bounds__(BS) k_divergence_multi(const DivJob* __restrict__ tab, int n) {
    const DivJob& e = job_of(tab, n, (int)blockIdx.x);
    divergence_cell(e.U, e.V, e.Wf, e.rhs, e.L, e.W, e.H, e.D, e.sc, make_int4(e.cb[0], e.cb[1], e.cb[2], e.cb[3]),
                    make_int2(e.cb[4], e.cb[5]), (int)((blockIdx.x - e.first_block) * BS + threadIdx.x));
}

// the pressure update on one axis of every grid in a table
template <int AXIS>
__global__ void __launch_bounds__(BS) k_pupdate_multi(const PupJob* __restrict__ tab, int n) {
    const PupJob& e = job_of(tab, n, (int)blockIdx.x);
    pupdate_face<AXIS>(e.in, e.out, e.p, e.L, e.W, e.H, e.D, e.sc, e.closedLo, e.closedHi,
                       (int)((blockIdx.x - e.first_block) * BS + threadIdx.x));
}

__global__ void __launch_bounds__(BS) k_rb_to_natural(RB p, RBL L, float* __restrict__ out, int W, int H, int D) {
    const int t = blockIdx.x * BS + threadIdx.x;
    if (t >= W * H * D) return;
    const int i = t % W, j = (t / W) % H, k = t / (W * H);
    out[t] = p_at(p, L, i, j, k);
}

__global__ void __launch_bounds__(BS) k_natural_to_rb(RB p, RBL L, const float* __restrict__ in, int W, int H,
                                                      int D) {
    const int t = blockIdx.x * BS + threadIdx.x;
    if (t >= W * H * D) return;
    const int i = t % W, j = (t / W) % H, k = t / (W * H);
    p.c[(i + j) & 1][L.at(i, j, k)] = in[t];
}

// pressure planes: +0 inside, -0 in the ghost shell; rhs planes: 0
// planes 0, 1 (home pressure set) and 4, 5 (spare set): +0 inside, -0 in
// the ghost shell; rhs planes 2, 3 and the tail: +0
__global__ void __launch_bounds__(BS) k_init_pressure(float* __restrict__ prb, size_t Np, size_t n, RBL L) {
    for (size_t t = (size_t)blockIdx.x * BS + threadIdx.x; t < n; t += (size_t)gridDim.x * BS) {
        const size_t plane = t / Np;
        const size_t e = t % Np;
        const int comp = (int)(e & 3);
        const size_t v = e >> 2;
        const int m = (int)(v % L.Wp), j = (int)((v / L.Wp) % L.Hp), g = (int)(v / ((size_t)L.Wp * L.Hp));
        const int k = (g - 1) * 4 + comp;
        const bool inside = m >= 1 && m <= L.Wh && j >= 1 && j <= L.H && g >= 1 && g <= L.Dg && k < L.D;
        const bool pset = plane != 2 && plane != 3; // pressure planes: home, spare, judge
        prb[t] = (pset && !inside) ? -0.0f : 0.0f;
    }
}

// max |a| over an array (NaN counts as +inf), into a shared double slot
__global__ void __launch_bounds__(BS) k_maxabs(const float* __restrict__ a, size_t n, double* out) {
    double m = 0.0;
    for (size_t t = (size_t)blockIdx.x * BS + threadIdx.x; t < n; t += (size_t)gridDim.x * BS) {
        double v = (double)fabsf(a[t]);
        if (v != v) v = __longlong_as_double(0x7ff0000000000000LL);
        m = fmax(m, v);
    }
    __shared__ double sm[BS / 64];
    m = wave_max(m);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) sm[wid] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        double w = sm[0];
        for (int q = 1; q < (int)(BS / 64); ++q) w = fmax(w, sm[q]);
        atomicMax((unsigned long long*)out, (unsigned long long)__double_as_longlong(w));
    }
}

__global__ void k_fill_floats(float* __restrict__ d, size_t n, float v) {
    for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (size_t)gridDim.x * blockDim.x) d[t] = v;
}

__global__ void k_fill_doubles(double* d, int n, double v) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) d[t] = v;
}

// every float a: div_small(a, d, RN(1/d)) against IEEE a / d
__global__ void __launch_bounds__(256) k_div_selftest(float d, float y, unsigned long long* bad) {
    unsigned long long local = 0;
    for (unsigned long long u = (unsigned long long)blockIdx.x * 256 + threadIdx.x; u < (1ull << 32);
         u += (unsigned long long)gridDim.x * 256) {
        const float a = __uint_as_float((unsigned)u);
        const float ref = a / d;
        const float got = div_small(a, d, y);
        const bool same = __float_as_uint(ref) == __float_as_uint(got) || (ref != ref && got != got);
        local += same ? 0 : 1;
    }
    if (local) atomicAdd(bad, local);
}

} // namespace

#define LAUNCH(kern, n, ...)                                                                  \
    do {                                                                                      \
        if ((n) > 0) {                                                                        \
            hipLaunchKernelGGL(kern, dim3(blocks_for((n), BS)), dim3(BS), 0, s.stream, __VA_ARGS__); \
            BLZ_HIP(hipGetLastError());                                                       \
        }                                                                                     \
    } while (0)

static inline int4 cb_a(const Solver& s) {
    const int* c = s.s.closed_bounds;
    return make_int4(c[0], c[1], c[2], c[3]);
}
static inline int2 cb_b(const Solver& s) {
    const int* c = s.s.closed_bounds;
    return make_int2(c[4], c[5]);
}

// the home pressure planes (ghosts included) set to v: the autotuner's
// timing state
int launch_fill_pressure(Solver& s, float v, hipStream_t st) {
    const size_t n = 2 * s.Np;
    const unsigned nb = (unsigned)std::min<size_t>(blocks_for(n, BS), 8192);
    hipLaunchKernelGGL(k_fill_floats, dim3(nb), dim3(BS), 0, st ? st : s.stream, s.prb, n, v);
    BLZ_HIP(hipGetLastError());
    return BLZ_OK;
}

int launch_init_pressure(Solver& s) {
    const unsigned nb = (unsigned)std::min<size_t>(blocks_for(s.prb_floats, BS), 8192);
    hipLaunchKernelGGL(k_init_pressure, dim3(nb), dim3(BS), 0, s.stream, s.prb, s.Np, s.prb_floats, s.L);
    BLZ_HIP(hipGetLastError());
    return BLZ_OK;
}

int launch_divergence(Solver& s) {
    const float sc = 1.0f / s.dx;
    LAUNCH(k_divergence, (size_t)s.W * s.H * s.D, s.f[F_U].front(), s.f[F_V].front(), s.f[F_W].front(), s.rhs, s.L,
           s.W, s.H, s.D, sc, cb_a(s), cb_b(s));
    return BLZ_OK;
}

void divergence_jobs(Solver& s, std::vector<DivJob>& jobs) {
    DivJob j;
    j.U = s.f[F_U].front();
    j.V = s.f[F_V].front();
    j.Wf = s.f[F_W].front();
    j.rhs = s.rhs;
    j.L = s.L;
    j.W = s.W;
    j.H = s.H;
    j.D = s.D;
    j.sc = 1.0f / s.dx;
    for (int a = 0; a < 6; ++a) j.cb[a] = s.s.closed_bounds[a];
    j.first_block = 0;
    j.n = s.W * s.H * s.D;
    jobs.push_back(j);
}

int launch_divergence_multi(hipStream_t st, const DivJob* tab, int n, int blocks) {
    if (n < 1 || blocks < 1) return BLZ_OK;
    hipLaunchKernelGGL(k_divergence_multi, dim3((unsigned)blocks), dim3(BS), 0, st, tab, n);
    BLZ_HIP(hipGetLastError());
    return BLZ_OK;
}

// set: first plane of the pressure pair swept (0 home, 4 spare, 6 judge)
template <int PG>
static int rbgs_passes(Solver& s, float scale, int sweeps, const int* stop, int set = 0) {
    const size_t cols = (size_t)(s.W / 2) * s.H;
    const dim3 grid(blocks_for(cols, 64)), block(64);
    const DivTab tab{1.0f / 3.0f, 1.0f / 4.0f, 1.0f / 5.0f, 1.0f / 6.0f};
    const unsigned plane_bytes = (unsigned)(s.Np * 4);
    const unsigned p0 = (unsigned)set * plane_bytes;
    const unsigned nbytes = (unsigned)(set ? s.prb_floats * 4 : 4 * s.Np * 4);
    for (int l = 0; l < sweeps; ++l) {
        for (int c = 0; c < 2; ++c) {
            hipLaunchKernelGGL((k_rbgs_pass<PG>), grid, block, 0, s.stream, s.prb, nbytes, p0 + c * plane_bytes,
                               p0 + (c ^ 1) * plane_bytes, (2 + c) * plane_bytes, s.W, s.H, s.D, s.L.Wp, s.L.Hp, c,
                               scale, tab, stop);
            BLZ_HIP(hipGetLastError());
        }
    }
    return BLZ_OK;
}


// residual of the spare planes into partial, for the warm start's judge
int launch_residual_partial(Solver& s, float dt, bool spare, double* partial, const int* stop, int* nb_out) {
    const double cc = (double)dt / ((double)s.s.density * (double)s.dx * (double)s.dx);
    const int nb = residual_blocks(s);
    RB pp = s.p;
    if (spare) {
        pp.c[0] = s.prb + (size_t)s.judge_set * s.Np;
        pp.c[1] = s.prb + (size_t)(s.judge_set + 1) * s.Np;
    }
    hipLaunchKernelGGL(k_residual, dim3(nb), dim3(BS), 0, s.stream, pp, s.rhs, s.L, s.W, s.H, s.D, cc, cb_a(s),
                       cb_b(s), partial, stop);
    BLZ_HIP(hipGetLastError());
    *nb_out = nb;
    return BLZ_OK;
}

// Pairs of fused sweeps: home set -> spare set -> home set, so the pressure
// is in the home planes after every pair (and at every residual check).
// home: first plane of the pressure pair swept (0, or 6 for the judge's
// copy); the spare planes (4, 5) are the other half of each pair
template <int PG, int WPB>
static int fused_pairs(Solver& s, float scale, int pairs, const int* stop, int home) {
    const int nwr = (s.W + 63) / 64;
    const int nwaves = nwr * s.H;
    const dim3 grid((unsigned)((nwaves + WPB - 1) / WPB)), block(64 * WPB);
    const DivTab tab{1.0f / 3.0f, 1.0f / 4.0f, 1.0f / 5.0f, 1.0f / 6.0f};
    const unsigned pb = (unsigned)(s.Np * 4);
    const unsigned nbytes = (unsigned)(s.prb_floats * 4);
    const unsigned hp = (unsigned)home * pb;
    for (int l = 0; l < pairs; ++l) {
        for (int h = 0; h < 2; ++h) {
            const unsigned src = h == 0 ? hp : 4u * pb, dst = h == 0 ? 4u * pb : hp;
            hipLaunchKernelGGL((k_rbgs_fused<PG, WPB>), grid, block, 0, s.stream, s.prb, nbytes, src, src + pb,
                               2u * pb, 3u * pb, dst, dst + pb, s.W, s.H, s.D, s.L.Wp, s.L.Hp, scale, tab, stop,
                               nwaves);
            BLZ_HIP(hipGetLastError());
        }
    }
    return BLZ_OK;
}

// the same pairs with the row-pair fused sweep
template <int PG, int WPB>
static int pair_pairs(Solver& s, float scale, int pairs, const int* stop, int home) {
    const int nwaves = pair_segments(s.W) * ((s.H + 1) / 2);
    const dim3 grid((unsigned)((nwaves + WPB - 1) / WPB)), block(64 * WPB);
    const DivTab tab{1.0f / 3.0f, 1.0f / 4.0f, 1.0f / 5.0f, 1.0f / 6.0f};
    const unsigned pb = (unsigned)(s.Np * 4);
    const unsigned nbytes = (unsigned)(s.prb_floats * 4);
    const unsigned hp = (unsigned)home * pb;
    for (int l = 0; l < pairs; ++l) {
        for (int h = 0; h < 2; ++h) {
            const unsigned src = h == 0 ? hp : 4u * pb, dst = h == 0 ? 4u * pb : hp;
            hipLaunchKernelGGL((k_rbgs_pair<PG, WPB>), grid, block, 0, s.stream, s.prb, nbytes, src, src + pb,
                               2u * pb, 3u * pb, dst, dst + pb, s.W, s.H, s.D, s.L.Wp, s.L.Hp, scale, tab, stop,
                               nwaves);
            BLZ_HIP(hipGetLastError());
        }
    }
    return BLZ_OK;
}

bool fused_available(const Solver& s) { return s.prb_floats >= 6 * s.Np; }

// Gauss-Seidel sweeps. variant: 2/4/8 = two in-place colour passes per
// sweep with that many slab groups in flight; 100*w + g (w = 1, 2, 4 waves
// per block, g = 2, 3 groups in flight) = fused sweeps in pairs, an odd
// last sweep as two colour passes; 1000 + 100*w + g (g = 3..6) = the same
// with the row-pair fused sweep; 0 = the autotuner's choice (8 where it
// has not run).
bool rbgs_variant_valid(int v) {
    return v == 0 || v == 2 || v == 4 || v == 8 || v == 102 || v == 103 || v == 202 || v == 203 || v == 402 ||
           v == 403 || v == 1103 || v == 1104 || v == 1105 || v == 1106 || v == 1203 || v == 1204 || v == 1403 ||
           v == 1404;
}

static int fused_dispatch(Solver& s, int v, float scale, int pairs, const int* stop, int home = 0) {
    switch (v) {
    case 1103: return pair_pairs<3, 1>(s, scale, pairs, stop, home);
    case 1104: return pair_pairs<4, 1>(s, scale, pairs, stop, home);
    case 1105: return pair_pairs<5, 1>(s, scale, pairs, stop, home);
    case 1106: return pair_pairs<6, 1>(s, scale, pairs, stop, home);
    case 1203: return pair_pairs<3, 2>(s, scale, pairs, stop, home);
    case 1204: return pair_pairs<4, 2>(s, scale, pairs, stop, home);
    case 1403: return pair_pairs<3, 4>(s, scale, pairs, stop, home);
    case 1404: return pair_pairs<4, 4>(s, scale, pairs, stop, home);
    case 102: return fused_pairs<2, 1>(s, scale, pairs, stop, home);
    case 202: return fused_pairs<2, 2>(s, scale, pairs, stop, home);
    case 203: return fused_pairs<3, 2>(s, scale, pairs, stop, home);
    case 402: return fused_pairs<2, 4>(s, scale, pairs, stop, home);
    case 403: return fused_pairs<3, 4>(s, scale, pairs, stop, home);
    default: return fused_pairs<3, 1>(s, scale, pairs, stop, home);
    }
}

// `sweeps` sweeps on the judge's pressure planes (the warm start's
// uncorrected path; same operations as on the home planes): with judge
// planes of their own, the chosen fused variant in pairs through the spare
// planes (free again once the corrected path's sweeps are done), an odd
// last sweep as two colour passes; else two-pass sweeps in place
int launch_rbgs_spare(Solver& s, float dt, int sweeps, const int* stop) {
    const float scale = (s.s.density * s.dx * s.dx) / dt;
    const int v = s.rbgs_variant ? s.rbgs_variant : s.rbgs_auto;
    if (s.judge_set == 6 && v >= 100) {
        if (int e = fused_dispatch(s, v, scale, sweeps / 2, stop, 6)) return e;
        return (sweeps & 1) ? rbgs_passes<8>(s, scale, 1, stop, 6) : BLZ_OK;
    }
    return rbgs_passes<8>(s, scale, sweeps, stop, s.judge_set);
}

int launch_rbgs_sweeps(Solver& s, float dt, int sweeps, const int* stop, int* launches, int* kind) {
    const float scale = (s.s.density * s.dx * s.dx) / dt; // Bridson p.75, :530
    int v = s.rbgs_variant ? s.rbgs_variant : s.rbgs_auto;
    if (v == 0) v = 8;
    if (v >= 100 && !fused_available(s)) v = 8;
    if (v >= 100) {
        const int pairs = sweeps / 2;
        *launches = 2 * pairs + 2 * (sweeps & 1);
        *kind = v >= 1000 ? 3 : 2;
        if (int e = fused_dispatch(s, v, scale, pairs, stop)) return e;
        return (sweeps & 1) ? rbgs_passes<8>(s, scale, 1, stop) : BLZ_OK;
    }
    *launches = 2 * sweeps;
    *kind = 1;
    switch (v) {
    case 2: return rbgs_passes<2>(s, scale, sweeps, stop);
    case 4: return rbgs_passes<4>(s, scale, sweeps, stop);
    default: return rbgs_passes<8>(s, scale, sweeps, stop);
    }
}

int fused_waves(const Solver& s) { return ((s.W + 63) / 64) * s.H; }
int pair_waves(const Solver& s) { return pair_segments(s.W) * ((s.H + 1) / 2); }

int residual_blocks(const Solver& s) {
    const size_t items = (size_t)(s.W / 2) * s.H * ((s.L.Dg + RES_SEG - 1) / RES_SEG);
    return (int)std::min<size_t>(blocks_for(items, BS), RESIDUAL_BLOCKS);
}

template <int PG, int WPB>
static int fused_multi(hipStream_t st, const FusedScene* tab, int n, int blocks) {
    const DivTab dt{1.0f / 3.0f, 1.0f / 4.0f, 1.0f / 5.0f, 1.0f / 6.0f};
    hipLaunchKernelGGL((k_rbgs_fused_multi<PG, WPB>), dim3((unsigned)blocks), dim3(64 * WPB), 0, st, tab, n, dt);
    BLZ_HIP(hipGetLastError());
    return BLZ_OK;
}

template <int PG, int WPB>
static int pair_multi(hipStream_t st, const FusedScene* tab, int n, int blocks) {
    const DivTab dt{1.0f / 3.0f, 1.0f / 4.0f, 1.0f / 5.0f, 1.0f / 6.0f};
    hipLaunchKernelGGL((k_rbgs_pair_multi<PG, WPB>), dim3((unsigned)blocks), dim3(64 * WPB), 0, st, tab, n, dt);
    BLZ_HIP(hipGetLastError());
    return BLZ_OK;
}

int launch_fused_multi(hipStream_t st, int variant, const FusedScene* tab, int n, int blocks) {
    switch (variant) {
    case 1103: return pair_multi<3, 1>(st, tab, n, blocks);
    case 1104: return pair_multi<4, 1>(st, tab, n, blocks);
    case 1203: return pair_multi<3, 2>(st, tab, n, blocks);
    case 1204: return pair_multi<4, 2>(st, tab, n, blocks);
    case 102: return fused_multi<2, 1>(st, tab, n, blocks);
    case 202: return fused_multi<2, 2>(st, tab, n, blocks);
    case 203: return fused_multi<3, 2>(st, tab, n, blocks);
    case 402: return fused_multi<2, 4>(st, tab, n, blocks);
    case 403: return fused_multi<3, 4>(st, tab, n, blocks);
    default: return fused_multi<3, 1>(st, tab, n, blocks);
    }
}

int launch_residual_partials_multi(hipStream_t st, const ResScene* tab, int n, int blocks) {
    hipLaunchKernelGGL(k_residual_multi, dim3((unsigned)blocks), dim3(BS), 0, st, tab, n);
    BLZ_HIP(hipGetLastError());
    return BLZ_OK;
}

int launch_residual_multi(hipStream_t st, const ResScene* tab, int n, int blocks, int add, int first) {
    hipLaunchKernelGGL(k_residual_multi, dim3((unsigned)blocks), dim3(BS), 0, st, tab, n);
    BLZ_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_res_final_multi, dim3((unsigned)n), dim3(BS), 0, st, tab, add, first);
    BLZ_HIP(hipGetLastError());
    return BLZ_OK;
}

// the launch arguments of a control block's reset
static CtlJob ctl_job(Solver& s, SolveCtl* c, bool adaptive, bool stop_stall) {
    // the residual history slot of this solve (control block q of s.ctl)
    const ptrdiff_t q = c - s.ctl;
    blz_trace_entry* tr = (s.trace && q >= 0 && q < MAX_SOLVES) ? s.trace + (size_t)q * s.trace_cap : nullptr;
    return CtlJob{c, tr, tr ? s.trace_cap : 0, adaptive ? 1 : 0, stop_stall ? 1 : 0,
                  std::max(1, s.s.residual_check_every)};
}

int launch_ctl_reset(Solver& s, SolveCtl* c, bool adaptive, bool stop_stall) {
    const CtlJob j = ctl_job(s, c, adaptive, stop_stall);
    hipLaunchKernelGGL(k_ctl_reset, dim3(1), dim3(1), 0, s.stream, j.c, j.trace, j.trace_cap, j.adaptive,
                       j.stop_stall, j.every);
    BLZ_HIP(hipGetLastError());
    return BLZ_OK;
}

void ctl_reset_jobs(Solver& s, SolveCtl* c, std::vector<CtlJob>& jobs, bool adaptive, bool stop_stall) {
    jobs.push_back(ctl_job(s, c, adaptive, stop_stall));
}

int launch_ctl_reset_multi(hipStream_t st, const CtlJob* tab, int n) {
    if (n < 1) return BLZ_OK;
    hipLaunchKernelGGL(k_ctl_reset_multi, dim3((unsigned)((n + 63) / 64)), dim3(64), 0, st, tab, n);
    BLZ_HIP(hipGetLastError());
    return BLZ_OK;
}

int launch_clear_doubles(Solver& s, double* dev, int n) {
    hipLaunchKernelGGL(k_fill_doubles, dim3(1), dim3(64), 0, s.stream, dev, n, 0.0);
    BLZ_HIP(hipGetLastError());
    return BLZ_OK;
}

int launch_residual_check(Solver& s, float dt, SolveCtl* c, int* host_done, int seq, double tol, int cap, int add,
                          int first) {
    const double cc = (double)dt / ((double)s.s.density * (double)s.dx * (double)s.dx);
    const size_t items = (size_t)(s.W / 2) * s.H * ((s.L.Dg + RES_SEG - 1) / RES_SEG);
    const int nb = (int)std::min<size_t>(blocks_for(items, BS), RESIDUAL_BLOCKS);
    double* partial = s.red + 8;
    hipLaunchKernelGGL(k_residual_chk, dim3(nb), dim3(BS), 0, s.stream, s.p, s.rhs, s.L, s.W, s.H, s.D, cc, cb_a(s),
                       cb_b(s), partial, c, add, cap, first);
    BLZ_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_res_final, dim3(1), dim3(BS), 0, s.stream, partial, nb, c, host_done, seq, tol, cap, add,
                       first);
    BLZ_HIP(hipGetLastError());
    return BLZ_OK;
}

int launch_residual_ws(Solver& s, float dt, SolveCtl* c, int* host_done, int seq, double tol, int cap) {
    const double cc = (double)dt / ((double)s.s.density * (double)s.dx * (double)s.dx);
    const size_t items = (size_t)(s.W / 2) * s.H * ((s.L.Dg + RES_SEG - 1) / RES_SEG);
    const int nb = (int)std::min<size_t>(blocks_for(items, BS), RESIDUAL_BLOCKS);
    hipLaunchKernelGGL(k_residual, dim3(nb), dim3(BS), 0, s.stream, s.p, s.rhs, s.L, s.W, s.H, s.D, cc, cb_a(s),
                       cb_b(s), s.red + 8, &c->done);
    BLZ_HIP(hipGetLastError());
    return launch_residual_ws_final(s, c, host_done, seq, tol, cap, nb);
}

int launch_max_abs_velocity(Solver& s, double* dev_out) {
    if (int e = launch_clear_doubles(s, dev_out, 1)) return e;
    for (int g = F_U; g <= F_W; ++g) {
        const size_t n = s.f[g].count();
        const unsigned nb = std::min<size_t>(blocks_for(n, BS), 2048);
        hipLaunchKernelGGL(k_maxabs, dim3(nb), dim3(BS), 0, s.stream, s.f[g].front(), n, dev_out);
        BLZ_HIP(hipGetLastError());
    }
    return BLZ_OK;
}

// closed_bounds slots of the low and high faces of each axis, as the
// reference assigns them (src/FluidSolver.cu:626-665: V's j = 0 face takes
// slot 3, its j = H face slot 2)
constexpr int kClosedLo[3] = {0, 3, 4}, kClosedHi[3] = {1, 2, 5};

int launch_pressure_update(Solver& s, float dt) {
    const float sc = dt / (s.s.density * s.dx);
    const int* cb = s.s.closed_bounds;
    Field& U = s.f[F_U];
    Field& V = s.f[F_V];
    Field& W = s.f[F_W];
    LAUNCH(k_pupdate<0>, U.count(), U.front(), U.back(), s.p, s.L, s.W, s.H, s.D, sc, cb[kClosedLo[0]],
           cb[kClosedHi[0]]);
    LAUNCH(k_pupdate<1>, V.count(), V.front(), V.back(), s.p, s.L, s.W, s.H, s.D, sc, cb[kClosedLo[1]],
           cb[kClosedHi[1]]);
    LAUNCH(k_pupdate<2>, W.count(), W.front(), W.back(), s.p, s.L, s.W, s.H, s.D, sc, cb[kClosedLo[2]],
           cb[kClosedHi[2]]);
    U.swap();
    V.swap();
    W.swap();
    return BLZ_OK;
}

// the jobs of launch_pressure_update (one per axis, appended to jobs[axis],
// first blocks left to the caller), with the same field swaps
void pupdate_jobs(Solver& s, float dt, std::vector<PupJob>* jobs) {
    const float sc = dt / (s.s.density * s.dx);
    const int* cb = s.s.closed_bounds;
    for (int a = 0; a < 3; ++a) {
        Field& F = s.f[F_U + a];
        PupJob j;
        j.in = F.front();
        j.out = F.back();
        j.p = s.p;
        j.L = s.L;
        j.W = s.W;
        j.H = s.H;
        j.D = s.D;
        j.sc = sc;
        j.closedLo = cb[kClosedLo[a]];
        j.closedHi = cb[kClosedHi[a]];
        j.first_block = 0;
        j.n = (int)F.count();
        jobs[a].push_back(j);
    }
    for (int a = 0; a < 3; ++a) s.f[F_U + a].swap();
}

int launch_pupdate_multi(hipStream_t st, int axis, const PupJob* tab, int n, int blocks) {
    if (n < 1 || blocks < 1) return BLZ_OK;
    if (axis == 0) hipLaunchKernelGGL(k_pupdate_multi<0>, dim3((unsigned)blocks), dim3(BS), 0, st, tab, n);
    else if (axis == 1) hipLaunchKernelGGL(k_pupdate_multi<1>, dim3((unsigned)blocks), dim3(BS), 0, st, tab, n);
    else hipLaunchKernelGGL(k_pupdate_multi<2>, dim3((unsigned)blocks), dim3(BS), 0, st, tab, n);
    BLZ_HIP(hipGetLastError());
    return BLZ_OK;
}

int launch_p_to_natural(Solver& s, float* out, const RB& src) {
    LAUNCH(k_rb_to_natural, (size_t)s.W * s.H * s.D, src, s.L, out, s.W, s.H, s.D);
    return BLZ_OK;
}

int launch_natural_to_rb(Solver& s, RB& dst, const float* in) {
    LAUNCH(k_natural_to_rb, (size_t)s.W * s.H * s.D, dst, s.L, in, s.W, s.H, s.D);
    return BLZ_OK;
}

int selftest_division(hipStream_t st, int d, unsigned long long* dev_bad) {
    const float fd = (float)d;
    hipLaunchKernelGGL(k_div_selftest, dim3(4096), dim3(256), 0, st, fd, 1.0f / fd, dev_bad);
    BLZ_HIP(hipGetLastError());
    return BLZ_OK;
}

} // namespace blz
