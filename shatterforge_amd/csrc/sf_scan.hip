// Device-count exclusive scan (see sf_scan.h). A tile is 256 threads x 8 consecutive items; the wave-level scan
// uses DPP-free shuffles over the 64-lane wavefront and one LDS word per wave.
#include "sf_scan.h"

namespace sf {

__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v) {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        uint32_t t = __shfl_up(v, o, 64);
        if (lane >= o) v += t;
    }
    return v;
}

// exclusive scan of one value per thread across the block; returns the block total through *tot
__device__ __forceinline__ uint32_t block_excl_scan(uint32_t v, uint32_t* tot) {
    __shared__ uint32_t wsum[SCAN_THREADS / 64];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    uint32_t inc = wave_incl_scan(v);
    if (lane == 63) wsum[wid] = inc;
    __syncthreads();
    uint32_t off = 0, all = 0;
#pragma unroll
    for (int k = 0; k < (int)(SCAN_THREADS / 64); k++) { uint32_t s = wsum[k]; if (k < wid) off += s; all += s; }
    __syncthreads();
    *tot = all;
    return off + inc - v;
}

__device__ __forceinline__ uint32_t count_of(const uint32_t* n_dev, uint32_t n_fixed, uint32_t cap) {
    uint32_t n = n_dev ? *n_dev : n_fixed;
    return n < cap ? n : cap;
}

__global__ void __launch_bounds__(SCAN_THREADS) k_tile_sums(const uint32_t* in, uint32_t* sums, const uint32_t* n_dev, uint32_t n_fixed, uint32_t cap) {
    const uint32_t n = count_of(n_dev, n_fixed, cap);
    const uint32_t base = blockIdx.x * SCAN_TILE + threadIdx.x * SCAN_ITEMS;
    uint32_t s = 0;
    if (blockIdx.x * SCAN_TILE < n) {
#pragma unroll
        for (uint32_t k = 0; k < SCAN_ITEMS; k++) if (base + k < n) s += in[base + k];
    }
    uint32_t tot;
    (void)block_excl_scan(s, &tot);
    if (threadIdx.x == 0) sums[blockIdx.x] = tot;
}

// one block: exclusive scan of the tile sums in place, total to *total
__global__ void __launch_bounds__(SCAN_THREADS) k_scan_sums(uint32_t* sums, uint32_t ntiles, uint32_t* total) {
    uint32_t carry = 0;
    for (uint32_t base = 0; base < ntiles; base += SCAN_THREADS) {
        uint32_t i = base + threadIdx.x;
        uint32_t v = i < ntiles ? sums[i] : 0u;
        uint32_t tot;
        uint32_t ex = block_excl_scan(v, &tot);
        if (i < ntiles) sums[i] = carry + ex;
        carry += tot;
    }
    if (threadIdx.x == 0 && total) *total = carry;
}

__global__ void __launch_bounds__(SCAN_THREADS) k_tile_scan(const uint32_t* in, uint32_t* out, const uint32_t* sums, const uint32_t* n_dev, uint32_t n_fixed, uint32_t cap) {
    const uint32_t n = count_of(n_dev, n_fixed, cap);
    if (blockIdx.x * SCAN_TILE >= n) return;
    const uint32_t base = blockIdx.x * SCAN_TILE + threadIdx.x * SCAN_ITEMS;
    uint32_t v[SCAN_ITEMS];
    uint32_t s = 0;
#pragma unroll
    for (uint32_t k = 0; k < SCAN_ITEMS; k++) { v[k] = base + k < n ? in[base + k] : 0u; s += v[k]; }
    uint32_t tot;
    uint32_t run = block_excl_scan(s, &tot) + sums[blockIdx.x];
#pragma unroll
    for (uint32_t k = 0; k < SCAN_ITEMS; k++) { if (base + k < n) out[base + k] = run; run += v[k]; }
}

// short arrays: one block walks the tiles in order, carrying the running total (one launch instead of three)
__global__ void __launch_bounds__(SCAN_THREADS) k_scan_one(const uint32_t* in, uint32_t* out, uint32_t* total, const uint32_t* n_dev, uint32_t n_fixed, uint32_t cap) {
    const uint32_t n = count_of(n_dev, n_fixed, cap);
    uint32_t carry = 0;
    for (uint32_t t0 = 0; t0 < n; t0 += SCAN_TILE) {
        const uint32_t base = t0 + threadIdx.x * SCAN_ITEMS;
        uint32_t v[SCAN_ITEMS];
        uint32_t sum = 0;
#pragma unroll
        for (uint32_t k = 0; k < SCAN_ITEMS; k++) { v[k] = base + k < n ? in[base + k] : 0u; sum += v[k]; }
        uint32_t tot;
        uint32_t run = carry + block_excl_scan(sum, &tot);
#pragma unroll
        for (uint32_t k = 0; k < SCAN_ITEMS; k++) { if (base + k < n) out[base + k] = run; run += v[k]; }
        carry += tot;
    }
    if (threadIdx.x == 0 && total) *total = carry;
}

void scan_u32(const uint32_t* in, uint32_t* out, uint32_t* total, const uint32_t* n_dev, uint32_t n_fixed, uint32_t cap,
              uint32_t* scratch, hipStream_t s) {
    uint32_t tiles = scan_tiles(cap);
    if (tiles <= 4) {
        hipLaunchKernelGGL(k_scan_one, dim3(1), dim3(SCAN_THREADS), 0, s, in, out, total, n_dev, n_fixed, cap);
        return;
    }
    hipLaunchKernelGGL(k_tile_sums, dim3(tiles), dim3(SCAN_THREADS), 0, s, in, scratch, n_dev, n_fixed, cap);
    hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(SCAN_THREADS), 0, s, scratch, tiles, total);
    hipLaunchKernelGGL(k_tile_scan, dim3(tiles), dim3(SCAN_THREADS), 0, s, in, out, scratch, n_dev, n_fixed, cap);
}

}  // namespace sf
