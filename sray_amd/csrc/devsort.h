/*
 * devsort.h — the two data-movement primitives the path needs besides its
 * own kernels, written for gfx950: a stable selection of flagged items
 * (order kept; the bounce-level compaction, the active pixel list, photon
 * storage) and an LSD radix sort of (key, value) pairs (the photon
 * kd-tree's per-level sorts and the gather grid's cell sort).
 *
 * Both split the items into CMP_BLOCKS contiguous ranges of whole
 * 256-item tiles.  Selection: one kernel counts each block's flagged
 * items, a second has each block sum the counts before it (at most
 * CMP_BLOCKS - 1 values, no scan launch) and write its items in order; the
 * last block writes the total.  Radix sort, per digit of up to 8 bits: a
 * 256-bin histogram per block (LDS atomics), an exclusive scan of each
 * digit's block counts (one block per digit) and of the digit
 * totals (by every sorting block, 256 values), then each block writes
 * its items tile by tile, in order: a lane's rank among equal digits of
 * its wave comes from one ballot per digit bit, the waves before it in the
 * tile add their counts for the digit, and the digit's running offset
 * carries across tiles.  Items with equal digits keep their order, so the
 * sort is stable, as LSD needs.
 */
#ifndef SRAY_DEVSORT_H_
#define SRAY_DEVSORT_H_

#include <hip/hip_runtime.h>

#include <cstdint>

namespace srd {

constexpr uint32_t CMP_BLOCKS = 1024, CMP_THREADS = 256;

/* block b's items: an equal share of whole 256-item tiles, in order */
__device__ __forceinline__ void cmp_range(uint32_t items, uint32_t b, uint32_t &lo, uint32_t &hi)
{
    const uint32_t tiles = (items + CMP_THREADS - 1u) / CMP_THREADS;
    const uint32_t per = (tiles + CMP_BLOCKS - 1u) / CMP_BLOCKS;
    lo = min(b * per * CMP_THREADS, items);
    hi = min((b + 1u) * per * CMP_THREADS, items);
}

/* pass 1 of a selection: this block's flagged items into bcount[block] */
__device__ __forceinline__ void cmp_block_count(const uint8_t *__restrict__ flags, uint32_t items,
                                                uint32_t *__restrict__ bcount)
{
    __shared__ uint32_t wsum[CMP_THREADS / 64];
    uint32_t lo, hi;
    cmp_range(items, blockIdx.x, lo, hi);
    uint32_t c = 0;
    for (uint32_t j = lo + threadIdx.x; j < hi; j += CMP_THREADS) c += flags[j] ? 1u : 0u;
    for (int off = 32; off > 0; off >>= 1) c += (uint32_t)__shfl_xor((int)c, off);
    if (__lane_id() == 0) wsum[threadIdx.x / 64] = c;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t t = 0;
        for (uint32_t w = 0; w < CMP_THREADS / 64; w++) t += wsum[w];
        bcount[blockIdx.x] = t;
    }
}

/* pass 2: this block's flagged items in order from its offset (the sum of
 * the counts before it); the last block writes the total */
template <class T>
__device__ __forceinline__ void cmp_block_scatter(const T *__restrict__ in, const uint8_t *__restrict__ flags,
                                                  uint32_t items, const uint32_t *__restrict__ bcount,
                                                  uint32_t *__restrict__ total, T *__restrict__ out)
{
    __shared__ uint32_t wsum[CMP_THREADS / 64];
    const uint32_t lane = __lane_id(), w = threadIdx.x / 64u;
    uint32_t pre = 0;
    for (uint32_t k = threadIdx.x; k < blockIdx.x; k += CMP_THREADS) pre += bcount[k];
    for (int off = 32; off > 0; off >>= 1) pre += (uint32_t)__shfl_xor((int)pre, off);
    if (lane == 0u) wsum[w] = pre;
    __syncthreads();
    uint32_t base = 0;
    for (uint32_t k = 0; k < CMP_THREADS / 64; k++) base += wsum[k];
    __syncthreads();
    if (blockIdx.x == gridDim.x - 1u && threadIdx.x == 0) *total = base + bcount[blockIdx.x];
    uint32_t lo, hi;
    cmp_range(items, blockIdx.x, lo, hi);
    for (uint32_t t0 = lo; t0 < hi; t0 += CMP_THREADS) {
        const uint32_t j = t0 + threadIdx.x;
        const bool f = j < hi && flags[j] != 0;
        const unsigned long long bal = __ballot(f);
        const uint32_t r = (uint32_t)__popcll(bal & ((1ull << lane) - 1ull));
        if (lane == 0u) wsum[w] = (uint32_t)__popcll(bal);
        __syncthreads();
        uint32_t wpre = 0, tot = 0;
        for (uint32_t k = 0; k < CMP_THREADS / 64; k++) {
            const uint32_t c = wsum[k];
            wpre += k < w ? c : 0u;
            tot += c;
        }
        if (f) out[base + wpre + r] = in[j];
        base += tot;
        __syncthreads();
    }
}

/* a selection over n items (launch both with CMP_BLOCKS x CMP_THREADS) */
template <class T>
__global__ __launch_bounds__(CMP_THREADS) void k_sel_count(const uint8_t *__restrict__ flags, uint32_t n,
                                                           uint32_t *__restrict__ bcount)
{
    cmp_block_count(flags, n, bcount);
}
template <class T>
__global__ __launch_bounds__(CMP_THREADS) void k_sel_scatter(const T *__restrict__ in,
                                                             const uint8_t *__restrict__ flags, uint32_t n,
                                                             const uint32_t *__restrict__ bcount,
                                                             uint32_t *__restrict__ total, T *__restrict__ out)
{
    cmp_block_scatter(in, flags, n, bcount, total, out);
}

/* ---------------------------------------------------------------- */
/* radix sort of (key, value) pairs                                  */
/* ---------------------------------------------------------------- */
/* block counts of each digit, digit-major: bhist[d * CMP_BLOCKS + block] */
template <class K>
__global__ __launch_bounds__(CMP_THREADS) void k_rs_hist(const K *__restrict__ keys, uint32_t n, int shift, int width,
                                                         uint32_t *__restrict__ bhist)
{
    __shared__ uint32_t h[256];
    const uint32_t R = 1u << width, mask = R - 1u;
    for (uint32_t i = threadIdx.x; i < R; i += CMP_THREADS) h[i] = 0u;
    __syncthreads();
    uint32_t lo, hi;
    cmp_range(n, blockIdx.x, lo, hi);
    for (uint32_t j = lo + threadIdx.x; j < hi; j += CMP_THREADS)
        atomicAdd(&h[(uint32_t)(keys[j] >> shift) & mask], 1u);
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < R; i += CMP_THREADS) bhist[i * CMP_BLOCKS + blockIdx.x] = h[i];
}

/* per digit (one block each, one thread per sorting block): the exclusive
 * scan of the digit's block counts in place, and the digit's total */
template <int UNUSED = 0>
__global__ __launch_bounds__(CMP_BLOCKS) void k_rs_scan(uint32_t *__restrict__ bhist, uint32_t *__restrict__ dtot)
{
    __shared__ uint32_t wsum[CMP_BLOCKS / 64];
    const uint32_t lane = __lane_id(), w = threadIdx.x / 64u;
    uint32_t *row = bhist + (size_t)blockIdx.x * CMP_BLOCKS;
    const uint32_t v = row[threadIdx.x];
    uint32_t x = v;
    for (int d = 1; d < 64; d <<= 1) {
        const uint32_t y = (uint32_t)__shfl_up((int)x, d);
        if (lane >= (uint32_t)d) x += y;
    }
    if (lane == 63u) wsum[w] = x;
    __syncthreads();
    uint32_t pre = 0;
    for (uint32_t q = 0; q < w; q++) pre += wsum[q];
    row[threadIdx.x] = pre + x - v;
    if (threadIdx.x == CMP_BLOCKS - 1u) dtot[blockIdx.x] = pre + x;
}

/* each block's items to their places, tile by tile in order */
template <class K>
__global__ __launch_bounds__(CMP_THREADS) void k_rs_scatter(const K *__restrict__ keys, const uint32_t *__restrict__ vals,
                                                            uint32_t n, int shift, int width,
                                                            const uint32_t *__restrict__ boff,
                                                            const uint32_t *__restrict__ dtot, K *__restrict__ okeys,
                                                            uint32_t *__restrict__ ovals)
{
    constexpr uint32_t NW = CMP_THREADS / 64;
    __shared__ uint32_t run[256];
    __shared__ uint32_t wc[NW][256];
    const uint32_t R = 1u << width, mask = R - 1u;
    const uint32_t lane = __lane_id(), w = threadIdx.x / 64u;
    /* the digits' starts: the exclusive scan of their totals (R <= 256 =
     * CMP_THREADS, one per thread) */
    {
        const uint32_t t = threadIdx.x < R ? dtot[threadIdx.x] : 0u;
        uint32_t x = t;
        for (int d = 1; d < 64; d <<= 1) {
            const uint32_t y = (uint32_t)__shfl_up((int)x, d);
            if (lane >= (uint32_t)d) x += y;
        }
        if (lane == 63u) wc[0][w] = x;
        __syncthreads();
        uint32_t pre = 0;
        for (uint32_t q = 0; q < w; q++) pre += wc[0][q];
        if (threadIdx.x < R) run[threadIdx.x] = pre + x - t + boff[threadIdx.x * CMP_BLOCKS + blockIdx.x];
        __syncthreads();
    }
    uint32_t lo, hi;
    cmp_range(n, blockIdx.x, lo, hi);
    const unsigned long long below = (1ull << lane) - 1ull;
    for (uint32_t t0 = lo; t0 < hi; t0 += CMP_THREADS) {
        for (uint32_t i = threadIdx.x; i < NW * R; i += CMP_THREADS) wc[i / R][i % R] = 0u;
        __syncthreads();
        const uint32_t j = t0 + threadIdx.x;
        const bool valid = j < hi;
        K k = 0;
        uint32_t v = 0, d = 0;
        if (valid) {
            k = keys[j];
            v = vals[j];
            d = (uint32_t)(k >> shift) & mask;
        }
        /* the lanes of this wave with the same digit */
        unsigned long long peers = __ballot(valid);
        for (int b = 0; b < width; b++) {
            const unsigned long long on = __ballot(valid && ((d >> b) & 1u));
            peers &= ((d >> b) & 1u) ? on : ~on;
        }
        const uint32_t rank = (uint32_t)__popcll(peers & below);
        if (valid && (peers & below) == 0ull) wc[w][d] = (uint32_t)__popcll(peers);
        __syncthreads();
        if (valid) {
            uint32_t pos = run[d] + rank;
            for (uint32_t q = 0; q < w; q++) pos += wc[q][d];
            okeys[pos] = k;
            ovals[pos] = v;
        }
        __syncthreads();
        for (uint32_t i = threadIdx.x; i < R; i += CMP_THREADS) {
            uint32_t t = 0;
            for (uint32_t q = 0; q < NW; q++) t += wc[q][i];
            run[i] += t;
        }
        __syncthreads();
    }
}

} // namespace srd

#endif
