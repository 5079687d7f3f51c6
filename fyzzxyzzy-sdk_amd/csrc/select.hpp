// Order-preserving stream compaction with device-side counts.
//
// Three launches: per-tile predicate counts, one-block scan of the tile
// counts (also publishing the total to device memory), then the ordered
// write. Predicates are small functors of (index) evaluated identically in
// the count and write passes, so no flag array is materialised.
#pragma once
#include "common.hpp"

namespace fyzz_hip::sel {

constexpr int kBlock = 256;
constexpr int kItems = 16;
constexpr int kTile = kBlock * kItems;

template <class P>
__global__ __launch_bounds__(kBlock) void k_count(u32 cap, const u32* d_n, P pred, u32* tile_sums) {
    __shared__ u32 scratch[kBlock / kWave + 1];
    const u32 n = min(*d_n, cap);
    const u32 base = blockIdx.x * kTile;
    u32 c = 0;
    if (base < n)
        for (int j = 0; j < kItems; ++j) {
            const u32 i = base + j * kBlock + threadIdx.x;
            if (i < n && pred(i)) ++c;
        }
    u32 total;
    (void)block_exclusive_scan<kBlock>(c, scratch, &total);
    if (threadIdx.x == 0) tile_sums[blockIdx.x] = total;
}

// Exclusive scan of tile_sums in place (one block); *d_out = total.
__global__ __launch_bounds__(1024) void k_scan_tiles(u32 tiles, u32* tile_sums, u32* d_out) {
    __shared__ u32 scratch[1024 / kWave + 1];
    u32 carry = 0;
    for (u32 t0 = 0; t0 < tiles; t0 += 1024) {
        const u32 t = t0 + threadIdx.x;
        const u32 v = t < tiles ? tile_sums[t] : 0u;
        u32 total;
        const u32 ex = block_exclusive_scan<1024>(v, scratch, &total);
        if (t < tiles) tile_sums[t] = carry + ex;
        carry += total;
    }
    if (threadIdx.x == 0) *d_out = carry;
}

template <class P, class W>
__global__ __launch_bounds__(kBlock) void k_write(u32 cap, const u32* d_n, P pred, W write,
                                                  const u32* tile_sums) {
    __shared__ u32 scratch[kBlock / kWave + 1];
    const u32 n = min(*d_n, cap);
    const u32 base = blockIdx.x * kTile;
    if (base >= n) return;
    // Thread t owns the contiguous run [base + t*16, base + t*16 + 16).
    const u32 first = base + threadIdx.x * kItems;
    u32 c = 0;
    for (int j = 0; j < kItems; ++j) {
        const u32 i = first + j;
        if (i < n && pred(i)) ++c;
    }
    u32 total;
    u32 out = tile_sums[blockIdx.x] + block_exclusive_scan<kBlock>(c, scratch, &total);
    for (int j = 0; j < kItems; ++j) {
        const u32 i = first + j;
        if (i < n && pred(i)) write(i, out++);
    }
}

inline u32 tiles_for(u32 cap) { return (cap + kTile - 1) / kTile; }

// Compacts indices [0, *d_n) (at most cap) where pred holds; write(i, o) is
// called with consecutive o in index order; *d_out receives the count.
// tile_scratch needs tiles_for(cap) words.
template <class P, class W>
void select(hipStream_t st, u32 cap, const u32* d_n, P pred, W write, u32* tile_scratch, u32* d_out) {
    const u32 tiles = tiles_for(cap) ? tiles_for(cap) : 1;
    hipLaunchKernelGGL((k_count<P>), dim3(tiles), dim3(kBlock), 0, st, cap, d_n, pred, tile_scratch);
    hipLaunchKernelGGL(k_scan_tiles, dim3(1), dim3(1024), 0, st, tiles, tile_scratch, d_out);
    hipLaunchKernelGGL((k_write<P, W>), dim3(tiles), dim3(kBlock), 0, st, cap, d_n, pred, write, tile_scratch);
    FYZZ_HIP_CHECK(hipGetLastError());
}

}  // namespace fyzz_hip::sel
