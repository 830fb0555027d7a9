// Exclusive prefix sums over u32 arrays whose length lives in device memory, so a step's launch sequence never
// waits on the host. Three passes: per-tile sums, a scan of the tile sums, per-tile scans plus the tile offset.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace sf {
constexpr uint32_t SCAN_THREADS = 256;
constexpr uint32_t SCAN_ITEMS = 8;
constexpr uint32_t SCAN_TILE = SCAN_THREADS * SCAN_ITEMS;

inline uint32_t scan_tiles(uint32_t cap) { return (cap + SCAN_TILE - 1) / SCAN_TILE; }
// scratch words needed by scan_u32 for arrays of up to `cap` elements
inline size_t scan_scratch_words(uint32_t cap) { return (size_t)scan_tiles(cap) + 64; }

// out[i] = in[0] + ... + in[i-1] for i < n, *total = the sum of all n, where n = *n_dev when n_dev is given
// (clamped to cap) and n_fixed otherwise. in and out may alias.
void scan_u32(const uint32_t* in, uint32_t* out, uint32_t* total, const uint32_t* n_dev, uint32_t n_fixed, uint32_t cap,
              uint32_t* scratch, hipStream_t s);
}  // namespace sf
