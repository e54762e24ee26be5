// k_source.hip — particle-to-grid cubic-pulse sourcing without atomics.
//
// Reference: addSource (asiroisvigneux/Blaze src/FluidSolver.cu:770-821):
// clear_back_buffer_kernel :57-74, add_source_to_back_buffer_kernel :76-97
// -> Grid::sphereToGrid (src/Grid.h:309-335, float atomicAdd per footprint
// cell), set_source_from_back_to_front_kernel :99-116 (front <- back where
// |back| > 1e-5).
//
// Deterministic gather: every (particle, channel) emits (cell key, value)
// entries in particle order (offsets from an exclusive scan of the counts);
// a stable radix sort by cell key (k_sort.hip) keeps ascending particle
// order inside each cell; one thread per cell then sums its run
// sequentially from +0 and overwrites the front value when |sum| > 1e-5.
// The entry buffers are sized from a host-side upper bound of the
// footprints (p2g_entry_bound), so no count is read back mid-pipeline. Cells that receive nothing keep their front value, exactly
// as the reference's copy-where-nonzero does. The reference's back buffers
// after addSource are dead (every later reader overwrites them first), so
// no full-grid clear or copy pass is needed.
#include <cmath>

#include "solver.h"

namespace blz {

namespace {

constexpr unsigned BS = 256;
constexpr int NCH = 5; // T, U, V, W, smoke density

struct PGrid {
    int nx, ny, nz;
    float ox, oy, oz;
    int flip;
    unsigned base; // first key of this grid
    float* front;
};

struct P2G {
    PGrid g[NCH];
    int nch;
    float dx;
    float dminx, dminy, dminz;
    float dsizey;
    float smoke_val;
};

__device__ __forceinline__ int trunc_i(float v) {
    v = fminf(fmaxf(v, -1073741824.0f), 1073741824.0f);
    return (int)v;
}

__device__ __forceinline__ float cubic_pulse(float x, float w) {
    if (x > w) return 0.0f;
    x /= w;
    return 1.0f - x * x * (3.0f - 2.0f * x);
}

// value splatted by particle r into channel ch, 0 when the particle or the
// channel is skipped (:83-95)
__device__ __forceinline__ bool channel_value(const float* r, int ch, const P2G& P, float* val) {
    if (!(r[6] > P.dx / 2.0f)) return false;
    float v;
    switch (ch) {
    case 0: v = r[7]; break;
    case 1: v = r[3]; break;
    case 2: v = r[4]; break;
    case 3: v = r[5]; break;
    default: v = P.smoke_val; break;
    }
    if (!abs_gt_1e5(v)) return false;
    *val = v;
    return true;
}

// Walk the footprint of sphereToGrid (src/Grid.h:315-335); call emit(cell,
// contribution) for every cell with mask > 0, in the reference loop order.
template <class F>
__device__ __forceinline__ void footprint(const float* r, const PGrid& g, const P2G& P, float val, F emit) {
    const float dx = P.dx;
    const float px = r[0], py = r[1], pz = r[2];
    const float radius = r[6] / 2.0f;
    const int imin = max(trunc_i((px - radius - P.dminx) / dx - g.ox), 0);
    const int imax = min(trunc_i((px + radius - P.dminx) / dx - g.ox) + 1, g.nx);
    const int jmin = max(trunc_i((P.dsizey - (py + radius - P.dminy)) / dx - g.oy), 0);
    const int jmax = min(trunc_i((P.dsizey - (py - radius - P.dminy)) / dx - g.oy) + 1, g.ny);
    const int kmin = max(trunc_i((pz - radius - P.dminz) / dx - g.oz), 0);
    const int kmax = min(trunc_i((pz + radius - P.dminz) / dx - g.oz) + 1, g.nz);
    const float pox = px - P.dminx, poy = P.dsizey - (py - P.dminy), poz = pz - P.dminz;
    for (int i = imin; i < imax; i++)
        for (int j = jmin; j < jmax; j++)
            for (int k = kmin; k < kmax; k++) {
                const float gx = ((float)i + g.ox) * dx, gy = ((float)j + g.oy) * dx, gz = ((float)k + g.oz) * dx;
                const float mask = cubic_pulse(len3(mk3(pox - gx, poy - gy, poz - gz)), radius);
                if (mask > 0.0f) emit(i + j * g.nx + k * g.nx * g.ny, g.flip ? -val * mask : val * mask);
            }
}

__global__ void __launch_bounds__(BS) k_p2g_count(const float* __restrict__ part, int n, P2G P,
                                                  unsigned* __restrict__ counts) {
    const int t = blockIdx.x * BS + threadIdx.x;
    if (t >= n * P.nch) return;
    const int q = t / P.nch, ch = t % P.nch;
    const float* r = part + (size_t)q * 8;
    float val;
    unsigned c = 0;
    if (channel_value(r, ch, P, &val)) footprint(r, P.g[ch], P, val, [&](int, float) { ++c; });
    counts[t] = c;
}

__global__ void __launch_bounds__(BS) k_p2g_emit(const float* __restrict__ part, int n, P2G P,
                                                 const unsigned* __restrict__ offs, unsigned cap,
                                                 unsigned* __restrict__ keys, float* __restrict__ vals) {
    const int t = blockIdx.x * BS + threadIdx.x;
    if (t >= n * P.nch) return;
    const int q = t / P.nch, ch = t % P.nch;
    const float* r = part + (size_t)q * 8;
    float val;
    if (!channel_value(r, ch, P, &val)) return;
    unsigned o = offs[t];
    const unsigned base = P.g[ch].base;
    footprint(r, P.g[ch], P, val, [&](int cell, float v) {
        if (o < cap) { // the host bound makes this always true (k_p2g_total flags the contrary)
            keys[o] = base + (unsigned)cell;
            vals[o] = v;
        }
        ++o;
    });
}

// the entry count the sort and the reduction use: the total of the scan,
// clamped to the buffers (a total above them, which the host bound rules
// out, is flagged and the splat reported as failed)
__global__ void k_p2g_total(const unsigned* __restrict__ total, unsigned cap, unsigned* __restrict__ n_out,
                            unsigned* __restrict__ overflow) {
    const unsigned t = *total;
    *n_out = t < cap ? t : cap;
    *overflow = t > cap ? 1u : 0u;
}

// one thread per run of equal keys: sequential sum in particle order
__global__ void __launch_bounds__(BS) k_p2g_reduce(const unsigned* __restrict__ keys, const float* __restrict__ vals,
                                                   const unsigned* __restrict__ total_dev, P2G P) {
    const unsigned total = *total_dev;
    const unsigned t = blockIdx.x * BS + threadIdx.x;
    if (t >= total) return;
    const unsigned key = keys[t];
    if (t > 0 && keys[t - 1] == key) return;
    float acc = 0.0f;
    for (unsigned u = t; u < total && keys[u] == key; ++u) acc += vals[u];
    if (!abs_gt_1e5(acc)) return;
    int ch = 0;
#pragma unroll
    for (int g = 1; g < NCH; ++g)
        if (g < P.nch && key >= P.g[g].base) ch = g;
    P.g[ch].front[key - P.g[ch].base] = acc;
}

int ensure(Solver& s, void** p, size_t* cap, size_t need) {
    if (*cap >= need) return BLZ_OK;
    if (*p) BLZ_HIP(hipFree(*p));
    *p = nullptr;
    *cap = 0;
    size_t n = need + need / 4 + 256;
    if (hipMalloc(p, n) != hipSuccess) {
        set_error("hipMalloc failed for particle-to-grid scratch");
        return BLZ_E_NOMEM;
    }
    *cap = n;
    (void)s;
    return BLZ_OK;
}

struct Scratch {
    void* counts = nullptr; size_t counts_cap = 0;   // counts, offsets, scan scratch, totals
    void* entries = nullptr; size_t entries_cap = 0; // keys, values and their sorted copies
    void* sort = nullptr; size_t sort_cap = 0;       // radix sort scratch
};

} // namespace

size_t scan_scratch_words(size_t n_max);
int scan_exclusive(hipStream_t st, const unsigned* in, unsigned* out, const unsigned* n_dev, size_t n_max,
                   unsigned* scratch);
size_t radix_scratch_words(size_t n_max);
int radix_sort_pairs(hipStream_t st, unsigned* keys, float* vals, unsigned* keys_alt, float* vals_alt,
                     const unsigned* n_dev, size_t n_max, int end_bit, unsigned* scratch, bool* in_alt);

// scratch buffers hang off the solver through p2g_buf
static Scratch& scratch_of(Solver& s) {
    if (!s.p2g_buf) s.p2g_buf = new Scratch();
    return *(Scratch*)s.p2g_buf;
}

void free_p2g_scratch(Solver& s) {
    if (!s.p2g_buf) return;
    Scratch* sc = (Scratch*)s.p2g_buf;
    if (sc->counts) (void)hipFree(sc->counts);
    if (sc->entries) (void)hipFree(sc->entries);
    if (sc->sort) (void)hipFree(sc->sort);
    delete sc;
    s.p2g_buf = nullptr;
}

// Upper bound of the (cell, value) entries of a particle set, on the host:
// a particle that passes the pscale > dx/2 gate (:83) covers at most
// floor(pscale/dx + 0.01) + 2 cells per axis of a channel's grid (the
// footprint box of sphereToGrid, src/Grid.h:315-324, spans pscale/dx cells
// up to float rounding, plus the truncation of both ends), in each of the
// nch channels.
size_t p2g_entry_bound(const float* part8, int n, float dx, int nch) {
    size_t b = 0;
    for (int q = 0; q < n; ++q) {
        const float ps = part8[(size_t)q * 8 + 6];
        if (!(ps > dx / 2.0f)) continue;
        const double w = std::floor((double)ps / (double)dx + 0.01) + 2.0;
        b += (size_t)(w * w * w) * (size_t)nch;
    }
    return b;
}

int launch_p2g(Solver& s) {
    const int n = s.npart;
    if (n <= 0) return BLZ_OK;
    Scratch& sc = scratch_of(s);
    P2G P{};
    P.nch = s.smoke ? 5 : 4;
    P.dx = s.dx;
    P.dminx = s.s.domain_min[0];
    P.dminy = s.s.domain_min[1];
    P.dminz = s.s.domain_min[2];
    P.dsizey = s.s.domain_max[1] - s.s.domain_min[1];
    P.smoke_val = s.s.smoke_density;
    unsigned base = 0;
    for (int g = 0; g < P.nch; ++g) {
        const Field& F = s.f[g];
        P.g[g] = PGrid{F.nx, F.ny, F.nz, F.ox, F.oy, F.oz, F.flip ? 1 : 0, base, F.front()};
        base += (unsigned)F.count();
    }
    const unsigned key_end = base;
    int end_bit = 1;
    while (end_bit < 32 && (1ull << end_bit) < (unsigned long long)key_end) ++end_bit;
    const size_t bound = s.p2g_bound;
    if (bound == 0) return BLZ_OK; // no particle passes the pscale gate
    if (bound > 0x7fffffffu) {
        set_error("particle-to-grid footprint bound exceeds 2^31 entries");
        return BLZ_E_INVALID;
    }

    // counts (m), offsets (m + 1, the total last), scan scratch, the clamped
    // total and the overflow flag
    const size_t m = (size_t)n * P.nch;
    const size_t scw = scan_scratch_words(m);
    if (int e = ensure(s, &sc.counts, &sc.counts_cap, (2 * m + 1 + scw + 2) * sizeof(unsigned))) return e;
    unsigned* counts = (unsigned*)sc.counts;
    unsigned* offs = counts + m;
    unsigned* scan_tmp = offs + m + 1;
    unsigned* n_eff = scan_tmp + scw;
    unsigned* overflow = n_eff + 1;
    if (int e = ensure(s, &sc.entries, &sc.entries_cap, bound * 16)) return e;
    if (int e = ensure(s, &sc.sort, &sc.sort_cap, radix_scratch_words(bound) * sizeof(unsigned))) return e;
    unsigned* keys_in = (unsigned*)sc.entries;
    unsigned* keys_out = keys_in + bound;
    float* vals_in = (float*)(keys_out + bound);
    float* vals_out = vals_in + bound;

    hipLaunchKernelGGL(k_p2g_count, dim3(blocks_for(m, BS)), dim3(BS), 0, s.stream, s.part, n, P, counts);
    BLZ_HIP(hipGetLastError());
    if (int e = scan_exclusive(s.stream, counts, offs, nullptr, m, scan_tmp)) return e;
    hipLaunchKernelGGL(k_p2g_total, dim3(1), dim3(1), 0, s.stream, offs + m, (unsigned)bound, n_eff, overflow);
    hipLaunchKernelGGL(k_p2g_emit, dim3(blocks_for(m, BS)), dim3(BS), 0, s.stream, s.part, n, P, offs,
                       (unsigned)bound, keys_in, vals_in);
    BLZ_HIP(hipGetLastError());
    bool alt = false;
    if (int e = radix_sort_pairs(s.stream, keys_in, vals_in, keys_out, vals_out, n_eff, bound, end_bit,
                                 (unsigned*)sc.sort, &alt))
        return e;
    hipLaunchKernelGGL(k_p2g_reduce, dim3(blocks_for(bound, BS)), dim3(BS), 0, s.stream, alt ? keys_out : keys_in,
                       alt ? vals_out : vals_in, n_eff, P);
    BLZ_HIP(hipGetLastError());
    s.p2g_overflow = overflow;
    return BLZ_OK;
}

} // namespace blz
