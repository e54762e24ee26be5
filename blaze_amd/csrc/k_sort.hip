// k_sort.hip — device-wide exclusive scan of 32-bit counts and a stable
// least-significant-digit radix sort of (key, value) pairs, for the
// particle-to-grid splat (k_source.hip). Integer arithmetic only, so every
// run gives the same order: a pass orders the pairs by one 8-bit digit and
// keeps the previous order among equal digits, so after the last pass the
// pairs are sorted by key and, inside a key, in their input order.
//
// The element count may be known on the device only (`n_dev`): kernels are
// launched for a host-side upper bound and read the count themselves.
//
// No reference counterpart (the reference accumulates with float
// atomicAdd, src/Grid.h:331, whose order is not fixed).
#include "solver.h"

namespace blz {

namespace {

constexpr int SB = 256;            // threads per block
constexpr int SCAN_ITEMS = 16;     // elements per thread of the scan
constexpr int SCAN_TILE = SB * SCAN_ITEMS;
constexpr int RADIX_BITS = 8;
constexpr int RADIX = 1 << RADIX_BITS;
constexpr int SORT_ITEMS = 8;      // rounds per tile of the sort
constexpr int SORT_TILE = SB * SORT_ITEMS;
constexpr int WAVES = SB / 64;

// block-wide exclusive scan of one value per thread (fixed order), returns
// the block total through *total
__device__ __forceinline__ unsigned block_exclusive_scan(unsigned v, unsigned* sm, unsigned* total) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    unsigned x = v;
    // inclusive scan inside the wave (Hillis-Steele over 64 lanes)
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const unsigned y = __shfl_up(x, o, 64);
        if (lane >= o) x += y;
    }
    if (lane == 63) sm[w] = x;
    __syncthreads();
    unsigned before = 0, all = 0;
#pragma unroll
    for (int q = 0; q < WAVES; ++q) {
        const unsigned t = sm[q];
        if (q < w) before += t;
        all += t;
    }
    __syncthreads(); // sm may be reused by the caller
    *total = all;
    return before + x - v;
}

// pass 1 of the scan: the sum of each tile of SCAN_TILE elements
__global__ void __launch_bounds__(SB) k_scan_sums(const unsigned* __restrict__ in, const unsigned* __restrict__ n_dev,
                                                  size_t n_host, unsigned* __restrict__ sums) {
    __shared__ unsigned sm[WAVES];
    const size_t n = n_dev ? (size_t)*n_dev : n_host;
    const size_t base = (size_t)blockIdx.x * SCAN_TILE + (size_t)threadIdx.x * SCAN_ITEMS;
    unsigned s = 0;
#pragma unroll
    for (int i = 0; i < SCAN_ITEMS; ++i)
        if (base + i < n) s += in[base + i];
    unsigned tot = 0;
    (void)block_exclusive_scan(s, sm, &tot);
    if (threadIdx.x == 0) sums[blockIdx.x] = tot;
}

// pass 2: exclusive scan of the tile sums in one block (in place)
__global__ void __launch_bounds__(SB) k_scan_top(unsigned* __restrict__ sums, int ntiles) {
    __shared__ unsigned sm[WAVES];
    unsigned carry = 0;
    for (int b0 = 0; b0 < ntiles; b0 += SB) {
        const int t = b0 + threadIdx.x;
        const unsigned v = t < ntiles ? sums[t] : 0u;
        unsigned tot = 0;
        const unsigned ex = block_exclusive_scan(v, sm, &tot);
        if (t < ntiles) sums[t] = carry + ex;
        carry += tot;
    }
}

// pass 3: each tile's exclusive scan plus its offset; out[n] receives the
// total (out holds n + 1 elements)
__global__ void __launch_bounds__(SB) k_scan_tiles(const unsigned* __restrict__ in, const unsigned* __restrict__ n_dev,
                                                   size_t n_host, const unsigned* __restrict__ sums,
                                                   unsigned* __restrict__ out) {
    __shared__ unsigned sm[WAVES];
    const size_t n = n_dev ? (size_t)*n_dev : n_host;
    const size_t base = (size_t)blockIdx.x * SCAN_TILE + (size_t)threadIdx.x * SCAN_ITEMS;
    unsigned v[SCAN_ITEMS];
    unsigned s = 0;
#pragma unroll
    for (int i = 0; i < SCAN_ITEMS; ++i) {
        v[i] = base + i < n ? in[base + i] : 0u;
        s += v[i];
    }
    unsigned tot = 0;
    unsigned run = sums[blockIdx.x] + block_exclusive_scan(s, sm, &tot);
#pragma unroll
    for (int i = 0; i < SCAN_ITEMS; ++i) {
        if (base + i < n) out[base + i] = run;
        if (base + i == n) out[n] = run; // the total (the tiles cover index n: n <= n_max)
        run += v[i];
    }
}

// the sort's digit histogram of each tile: hist[d * ntiles + tile]
__global__ void __launch_bounds__(SB) k_radix_hist(const unsigned* __restrict__ keys, const unsigned* __restrict__ n_dev,
                                                   int shift, unsigned* __restrict__ hist, int ntiles) {
    __shared__ unsigned h[RADIX];
    const size_t n = *n_dev;
    h[threadIdx.x] = 0u; // SB == RADIX
    __syncthreads();
    const size_t base = (size_t)blockIdx.x * SORT_TILE;
#pragma unroll
    for (int i = 0; i < SORT_ITEMS; ++i) {
        const size_t e = base + (size_t)i * SB + threadIdx.x;
        if (e < n) atomicAdd(&h[(keys[e] >> shift) & (RADIX - 1)], 1u);
    }
    __syncthreads();
    hist[(size_t)threadIdx.x * ntiles + blockIdx.x] = h[threadIdx.x];
}

// Stable scatter of a tile by one digit. The tile's elements are taken in
// rounds i = 0..SORT_ITEMS-1, element base + i * SB + thread, which is the
// input order (round, wave, lane). In a round each lane finds the lanes of
// its wave holding the same digit (eight ballots, one per digit bit); its
// destination is the digit's global offset (scanned histogram) + the
// elements of that digit in earlier rounds + in earlier waves of this round
// + on lower lanes of its wave.
__global__ void __launch_bounds__(SB) k_radix_scatter(const unsigned* __restrict__ keys_in,
                                                      const float* __restrict__ vals_in,
                                                      const unsigned* __restrict__ n_dev, int shift,
                                                      const unsigned* __restrict__ hist_scan, int ntiles,
                                                      unsigned* __restrict__ keys_out, float* __restrict__ vals_out) {
    __shared__ unsigned run[RADIX];        // offset of each digit for the next element
    __shared__ unsigned wc[WAVES][RADIX];  // per round: elements of each digit per wave
    const size_t n = *n_dev;
    const size_t base = (size_t)blockIdx.x * SORT_TILE;
    if (base >= n) return; // whole block: uniform
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    run[threadIdx.x] = hist_scan[(size_t)threadIdx.x * ntiles + blockIdx.x];
    const unsigned long long below = lane ? (~0ull >> (64 - lane)) : 0ull;
    for (int i = 0; i < SORT_ITEMS; ++i) {
        for (int q = 0; q < WAVES; ++q) wc[q][threadIdx.x] = 0u;
        __syncthreads();
        const size_t e = base + (size_t)i * SB + threadIdx.x;
        const bool ok = e < n;
        const unsigned key = ok ? keys_in[e] : 0u;
        const unsigned d = (key >> shift) & (RADIX - 1);
        unsigned long long peers = __ballot(ok);
#pragma unroll
        for (int b = 0; b < RADIX_BITS; ++b) {
            const unsigned long long m = __ballot((d >> b) & 1u);
            peers &= ((d >> b) & 1u) ? m : ~m;
        }
        const unsigned rank = (unsigned)__popcll(peers & below);
        if (ok && rank == 0) wc[w][d] = (unsigned)__popcll(peers);
        __syncthreads();
        if (ok) {
            unsigned off = run[d] + rank;
            for (int q = 0; q < w; ++q) off += wc[q][d];
            if (off < n) {
                keys_out[off] = key;
                vals_out[off] = vals_in[e];
            }
        }
        __syncthreads();
        unsigned add = 0;
        for (int q = 0; q < WAVES; ++q) add += wc[q][threadIdx.x];
        run[threadIdx.x] += add;
        __syncthreads();
    }
}

inline unsigned tiles_of(size_t n, size_t tile) { return (unsigned)((n + tile - 1) / tile); }

} // namespace

size_t scan_scratch_words(size_t n_max) { return tiles_of(n_max + 1, SCAN_TILE) + 1; }

// out[0..n) = exclusive prefix sums of in[0..n), out[n] = the total; n from
// n_dev when given (at most n_max), else n_max. scratch: scan_scratch_words.
int scan_exclusive(hipStream_t st, const unsigned* in, unsigned* out, const unsigned* n_dev, size_t n_max,
                   unsigned* scratch) {
    const unsigned nt = tiles_of(n_max + 1, SCAN_TILE);
    hipLaunchKernelGGL(k_scan_sums, dim3(nt), dim3(SB), 0, st, in, n_dev, n_max, scratch);
    hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(SB), 0, st, scratch, (int)nt);
    hipLaunchKernelGGL(k_scan_tiles, dim3(nt), dim3(SB), 0, st, in, n_dev, n_max, scratch, out);
    BLZ_HIP(hipGetLastError());
    return BLZ_OK;
}

size_t radix_scratch_words(size_t n_max) {
    const size_t nt = std::max<size_t>(1, tiles_of(n_max, SORT_TILE));
    return 2 * (size_t)RADIX * nt + 1 + scan_scratch_words((size_t)RADIX * nt);
}

// Stable sort of (keys, vals)[0..*n_dev) by the low end_bit bits of the
// keys, ping-ponging with (keys_alt, vals_alt); *in_alt receives whether the
// result ended in the alternate arrays. n_max bounds *n_dev.
int radix_sort_pairs(hipStream_t st, unsigned* keys, float* vals, unsigned* keys_alt, float* vals_alt,
                     const unsigned* n_dev, size_t n_max, int end_bit, unsigned* scratch, bool* in_alt) {
    const unsigned nt = std::max(1u, tiles_of(n_max, SORT_TILE));
    unsigned* hist = scratch;
    unsigned* hist_scan = scratch + (size_t)RADIX * nt;         // RADIX * nt + 1 words (the total last)
    unsigned* scan_tmp = scratch + 2 * (size_t)RADIX * nt + 1;
    unsigned *ki = keys, *ko = keys_alt;
    float *vi = vals, *vo = vals_alt;
    bool alt = false;
    for (int shift = 0; shift < end_bit; shift += RADIX_BITS) {
        hipLaunchKernelGGL(k_radix_hist, dim3(nt), dim3(SB), 0, st, ki, n_dev, shift, hist, (int)nt);
        BLZ_HIP(hipGetLastError());
        // the histograms of tiles past the count are zero, so a scan over all
        // launched tiles gives the same offsets as over the live ones
        if (int e = scan_exclusive(st, hist, hist_scan, nullptr, (size_t)RADIX * nt, scan_tmp)) return e;
        hipLaunchKernelGGL(k_radix_scatter, dim3(nt), dim3(SB), 0, st, ki, vi, n_dev, shift, hist_scan, (int)nt, ko,
                           vo);
        BLZ_HIP(hipGetLastError());
        std::swap(ki, ko);
        std::swap(vi, vo);
        alt = !alt;
    }
    *in_alt = alt;
    return BLZ_OK;
}

} // namespace blz
