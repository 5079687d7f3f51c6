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
constexpr int kItems = 4;
constexpr int kTile = kBlock * kItems;

__device__ inline u32 count_of(u32 cap, const u32* d_n) { return min(*d_n, cap); }

template <class P>
__global__ __launch_bounds__(kBlock) void k_count(u32 cap, const u32* d_n, P pred, u32* tile_sums,
                                                  const u32* gate) {
    __shared__ u32 scratch[kBlock / kWave + 1];
    if (gate && *gate == 0) return;
    const u32 n = count_of(cap, d_n);
    const u32 need = (n + kTile - 1) / kTile;
    for (u32 tile = blockIdx.x; tile < need; tile += gridDim.x) {
        const u32 base = tile * kTile;
        u32 c = 0;
        for (int j = 0; j < kItems; ++j) {
            const u32 i = base + j * kBlock + threadIdx.x;
            if (i < n && pred(i)) ++c;
        }
        u32 total;
        (void)block_exclusive_scan<kBlock>(c, scratch, &total);
        if (threadIdx.x == 0) tile_sums[tile] = total;
    }
}

// Exclusive scan of the needed tile sums in place (one block); *d_out = total.
static __global__ __launch_bounds__(1024) void k_scan_tiles(u32 cap, const u32* d_n, u32* tile_sums, u32* d_out,
                                                     const u32* gate) {
    __shared__ u32 scratch[1024 / kWave + 1];
    if (gate && *gate == 0) return;
    const u32 tiles = (count_of(cap, d_n) + kTile - 1) / kTile;
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

// The ordered write visits items in the count pass's layout (item j of
// thread t is index base + j*256 + t), so consecutive lanes handle
// consecutive indices and write consecutive outputs: the reads behind the
// predicate and the writer's records stay coalesced. An item's output is
// the tile's base plus the selected items before it in index order
// (rows j' < j, then earlier waves of row j, then earlier lanes).
template <class P, class W>
__global__ __launch_bounds__(kBlock) void k_write(u32 cap, const u32* d_n, P pred, W write,
                                                  const u32* tile_sums, const u32* gate) {
    constexpr int kWaves = kBlock / kWave;
    __shared__ u32 row_wave[kItems * kWaves];  // selected per (row, wave), then its exclusive prefix
    if (gate && *gate == 0) return;
    const u32 n = count_of(cap, d_n);
    const u32 need = (n + kTile - 1) / kTile;
    const u32 tid = threadIdx.x, wave = tid / kWave;
    for (u32 tile = blockIdx.x; tile < need; tile += gridDim.x) {
        const u32 base = tile * kTile;
        u32 sel = 0;  // bit j: item j selected
        u32 rank[kItems];
#pragma unroll
        for (int j = 0; j < kItems; ++j) {
            const u32 i = base + j * kBlock + tid;
            const bool f = i < n && pred(i);
            const u64 m = __ballot(f);
            rank[j] = lane_rank(m);
            sel |= f ? (1u << j) : 0u;
            if (lane_id() == 0) row_wave[j * kWaves + wave] = static_cast<u32>(__popcll(m));
        }
        __syncthreads();
        static_assert(kItems * kWaves <= kWave, "one wave scans the (row, wave) counts");
        if (tid < kWave) {
            const bool mine = tid < static_cast<u32>(kItems * kWaves);
            const u32 v = mine ? row_wave[tid] : 0u;
            const u32 ex = wave_inclusive_scan(v) - v;
            if (mine) row_wave[tid] = ex;
        }
        __syncthreads();
        const u32 out0 = tile_sums[tile];
#pragma unroll
        for (int j = 0; j < kItems; ++j)
            if (sel & (1u << j)) write(base + j * kBlock + tid, out0 + row_wave[j * kWaves + wave] + rank[j]);
        __syncthreads();
    }
}

inline u32 tiles_for(u32 cap) { return (cap + kTile - 1) / kTile; }

// Exclusive scan of value(i) over i < min(*d_n, cap) into out[i]; the
// total goes to *d_total. Same tiling as select (tile sums, one-block tile
// scan, per-tile scan and write).
template <class F>
__global__ __launch_bounds__(kBlock) void k_scan_sum(u32 cap, const u32* d_n, F value, u32* tile_sums) {
    __shared__ u32 scratch[kBlock / kWave + 1];
    const u32 n = count_of(cap, d_n);
    const u32 need = (n + kTile - 1) / kTile;
    for (u32 tile = blockIdx.x; tile < need; tile += gridDim.x) {
        const u32 base = tile * kTile;
        u32 c = 0;
        for (int j = 0; j < kItems; ++j) {
            const u32 i = base + j * kBlock + threadIdx.x;
            if (i < n) c += value(i);
        }
        u32 total;
        (void)block_exclusive_scan<kBlock>(c, scratch, &total);
        if (threadIdx.x == 0) tile_sums[tile] = total;
    }
}
template <class F>
__global__ __launch_bounds__(kBlock) void k_scan_write(u32 cap, const u32* d_n, F value, const u32* tile_sums,
                                                       u32* out) {
    __shared__ u32 scratch[kBlock / kWave + 1];
    const u32 n = count_of(cap, d_n);
    const u32 need = (n + kTile - 1) / kTile;
    for (u32 tile = blockIdx.x; tile < need; tile += gridDim.x) {
        const u32 base = tile * kTile;
        const u32 first = base + threadIdx.x * kItems;
        u32 c = 0;
        for (int j = 0; j < kItems; ++j) {
            const u32 i = first + j;
            if (i < n) c += value(i);
        }
        u32 total;
        u32 run = tile_sums[tile] + block_exclusive_scan<kBlock>(c, scratch, &total);
        for (int j = 0; j < kItems; ++j) {
            const u32 i = first + j;
            if (i < n) {
                out[i] = run;
                run += value(i);
            }
        }
    }
}
template <class F>
void exclusive_scan(hipStream_t st, u32 cap, const u32* d_n, F value, u32* out, u32* tile_scratch, u32* d_total) {
    const u32 tiles = tiles_for(cap) ? tiles_for(cap) : 1;
    const u32 grid = tiles < 1024u ? tiles : 1024u;
    hipLaunchKernelGGL((k_scan_sum<F>), dim3(grid), dim3(kBlock), 0, st, cap, d_n, value, tile_scratch);
    hipLaunchKernelGGL(k_scan_tiles, dim3(1), dim3(1024), 0, st, cap, d_n, tile_scratch, d_total,
                       static_cast<const u32*>(nullptr));
    hipLaunchKernelGGL((k_scan_write<F>), dim3(grid), dim3(kBlock), 0, st, cap, d_n, value, tile_scratch, out);
    FYZZ_HIP_CHECK(hipGetLastError());
}

// Compacts indices [0, *d_n) (at most cap) where pred holds; write(i, o) is
// called with consecutive o in index order; *d_out receives the count.
// tile_scratch needs tiles_for(cap) words. With a gate, the three launches
// do nothing (and leave *d_out alone) unless *gate != 0.
template <class P, class W>
void select(hipStream_t st, u32 cap, const u32* d_n, P pred, W write, u32* tile_scratch, u32* d_out,
            const u32* gate = nullptr) {
    const u32 tiles = tiles_for(cap) ? tiles_for(cap) : 1;
    const u32 grid = tiles < 1024u ? tiles : 1024u;
    hipLaunchKernelGGL((k_count<P>), dim3(grid), dim3(kBlock), 0, st, cap, d_n, pred, tile_scratch, gate);
    hipLaunchKernelGGL(k_scan_tiles, dim3(1), dim3(1024), 0, st, cap, d_n, tile_scratch, d_out, gate);
    hipLaunchKernelGGL((k_write<P, W>), dim3(grid), dim3(kBlock), 0, st, cap, d_n, pred, write, tile_scratch, gate);
    FYZZ_HIP_CHECK(hipGetLastError());
}

}  // namespace fyzz_hip::sel
