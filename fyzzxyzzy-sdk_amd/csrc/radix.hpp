// Stable LSD radix sort for gfx950: 8-bit digits, reduce-then-scan per
// pass (tile histogram of the keys -> per-digit scan over tiles ->
// LDS-staged scatter of keys and payloads).
//
// Tiles are 4096 elements (256 threads x 16 items). Ranking inside a tile
// uses wave64 match masks: eight __ballot()s give every lane the set of
// lanes that hold its digit, so the rank of an item among equal digits is
// a popcount, and per-wave digit counters live in LDS. Items are visited in
// original order (wave strip, item-major, lane-minor), so the sort is
// stable: equal keys keep their input order. The tile histogram reads the
// keys only; the last scatter can also write each payload's output
// position (the inverse permutation).
//
// A single-pass variant with decoupled look-back (tiles publish their
// digit counts and find their prefix from earlier tiles' words) measured
// slower here at these sizes (6M keys: 387 us against 305 us for 4
// passes): the prefix advances a few tiles per agent-scope memory round
// trip, which costs ~30 us per pass.
//
// A batch of `segs` independent arrays of equal capacity is sorted by one
// launch per kernel (blockIdx.y selects the array). Element counts are
// either fixed or read from device memory, so a count produced by an
// earlier kernel never needs a host readback.
#pragma once
#include "context.hpp"

namespace fyzz_hip::radix {

constexpr int kBlock = 256;
constexpr int kItems = 16;
constexpr int kTile = kBlock * kItems;
constexpr int kBins = 256;
constexpr int kWaves = kBlock / kWave;

// Where pass-0 keys come from.
enum class Source : int {
    Arrays = 0,     // keys_in[seg*stride + i], vals_in[seg*stride + i]
    SapPlanes = 1,  // endpoints of axis seg from six planes of nbox keys:
                    // element i -> key planes[(i&1)*3+seg][i>>1], val i
};

struct Batch {
    std::uint64_t stride;  // elements between consecutive arrays
    u32 segs;
    u32 cap;               // capacity (grid is sized for it)
    u32 n_fixed;           // used when d_count == nullptr
    const u32* d_count;    // per-array element count on device, or null
    const u32* planes;     // Source::SapPlanes
    u32 nbox;              // Source::SapPlanes
    const u32* d_gate;     // when set and *d_gate == 0 the launch does nothing
};

template <class K>
__device__ inline u32 digit_of(K k, int shift) {
    return static_cast<u32>(k >> shift) & 0xffu;
}

__device__ inline u32 seg_count(const Batch& b, u32 seg) {
    if (b.d_gate && *b.d_gate == 0) return 0;
    const u32 n = b.d_count ? b.d_count[seg] : b.n_fixed;
    return n < b.cap ? n : b.cap;
}
__device__ inline u32 tiles_needed(u32 n) { return (n + kTile - 1) / kTile; }

template <class K, Source S>
__device__ inline void load_item(const Batch& b, const K* keys, const u32* vals, u32 seg, u32 i,
                                 K& k, u32& v) {
    if constexpr (S == Source::SapPlanes) {
        k = static_cast<K>(b.planes[static_cast<std::size_t>((i & 1u) * 3u + seg) * b.nbox + (i >> 1)]);
        v = i;
    } else {
        k = keys[seg * b.stride + i];
        v = vals[seg * b.stride + i];
    }
}

template <class K, Source S>
__device__ inline K load_key(const Batch& b, const K* keys, u32 seg, u32 i) {
    if constexpr (S == Source::SapPlanes)
        return static_cast<K>(b.planes[static_cast<std::size_t>((i & 1u) * 3u + seg) * b.nbox + (i >> 1)]);
    else
        return keys[seg * b.stride + i];
}

// Lanes of this wave whose `d` equals mine (only among `valid` lanes).
// A wave whose valid lanes all hold the first one's digit (the common case
// in the upper passes of nearly sorted input) is answered by one ballot.
__device__ inline u64 match_digit(u32 d, bool valid) {
    const u64 vmask = __ballot(valid);
    if (vmask == 0) return 0;
    const u32 dl = static_cast<u32>(__shfl(static_cast<int>(d), static_cast<int>(__builtin_ctzll(vmask))));
    if (__ballot(valid && d == dl) == vmask) return vmask;
    u64 peers = vmask;
#pragma unroll
    for (int bit = 0; bit < 8; ++bit) {
        const bool set = (d >> bit) & 1u;
        const u64 m = __ballot(set);
        peers &= set ? m : ~m;
    }
    return peers;
}

// counts[(seg*tiles + tile)*256 + digit] = #items of the tile with digit
// (tile-major: each tile writes and later reads one contiguous 1 KB row).
template <class K, Source S>
__global__ __launch_bounds__(kBlock) void k_upsweep(Batch b, const K* keys, const u32* vals,
                                                    int shift, u32 tiles, u32* counts) {
    __shared__ u32 hist[kBins];
    const u32 seg = blockIdx.y, tid = threadIdx.x;
    const u32 n = seg_count(b, seg);
    const u32 need = tiles_needed(n);
    for (u32 tile = blockIdx.x; tile < need; tile += gridDim.x) {
        hist[tid] = 0;
        __syncthreads();
        const u32 base = tile * kTile;
        // all of the tile's loads in flight before any is used
        K k[kItems];
#pragma unroll
        for (int j = 0; j < kItems; ++j) {
            const u32 i = base + j * kBlock + tid;
            k[j] = i < n ? load_key<K, S>(b, keys, seg, i) : K(0);
        }
        // counts only (no ranks): a wave with one digit adds once, others
        // add per lane in LDS
#pragma unroll
        for (int j = 0; j < kItems; ++j) {
            const bool valid = base + j * kBlock + tid < n;
            const u32 d = digit_of(k[j], shift);
            const u64 vmask = __ballot(valid);
            if (vmask == 0) continue;
            const u32 leader = static_cast<u32>(__builtin_ctzll(vmask));
            const u32 dl = static_cast<u32>(__shfl(static_cast<int>(d), static_cast<int>(leader)));
            if (__ballot(valid && d == dl) == vmask) {
                if (lane_id() == leader) atomicAdd(&hist[d], static_cast<u32>(__popcll(vmask)));
            } else if (valid) {
                atomicAdd(&hist[d], 1u);
            }
        }
        __syncthreads();
        counts[(static_cast<std::size_t>(seg) * tiles + tile) * kBins + tid] = hist[tid];
        __syncthreads();
    }
}

// Exclusive scan over tiles for each (seg, digit); totals[seg*256+digit].
__global__ __launch_bounds__(kBlock) void k_tile_scan(Batch b, u32 tiles, u32* counts, u32* totals) {
    __shared__ u32 scratch[kWaves + 1];
    const u32 seg = blockIdx.y, digit = blockIdx.x;
    if (b.d_gate && *b.d_gate == 0) return;
    const u32 need = tiles_needed(seg_count(b, seg));
    u32* col = counts + static_cast<std::size_t>(seg) * tiles * kBins + digit;  // stride kBins
    u32 carry = 0;
    for (u32 t0 = 0; t0 < need; t0 += kBlock) {
        const u32 t = t0 + threadIdx.x;
        const u32 v = t < need ? col[static_cast<std::size_t>(t) * kBins] : 0u;
        u32 total;
        const u32 ex = block_exclusive_scan<kBlock>(v, scratch, &total);
        if (t < need) col[static_cast<std::size_t>(t) * kBins] = carry + ex;
        carry += total;
    }
    if (threadIdx.x == 0) totals[seg * kBins + digit] = carry;
}

template <class K, Source S>
__global__ __launch_bounds__(kBlock) void k_scatter(Batch b, const K* keys, const u32* vals,
                                                    K* keys_out, u32* vals_out, int shift,
                                                    u32 tiles, const u32* counts,
                                                    const u32* totals, u32* rank_out) {
    __shared__ u32 s_cnt[kWaves][kBins];
    __shared__ u32 s_lbase[kBins];
    __shared__ u32 s_gbase[kBins];
    __shared__ u32 scratch[kWaves + 1];
    __shared__ K s_keys[kTile];
    __shared__ u32 s_vals[kTile];

    const u32 seg = blockIdx.y, tid = threadIdx.x;
    const u32 n = seg_count(b, seg);
    const u32 need = tiles_needed(n);
    const u32 wave = tid / kWave, lane = lane_id();
    for (u32 tile = blockIdx.x; tile < need; tile += gridDim.x) {
    const u32 base = tile * kTile;
    for (u32 i = tid; i < kWaves * kBins; i += kBlock) (&s_cnt[0][0])[i] = 0;
    __syncthreads();

    K k[kItems];
    u32 v[kItems];
    u32 r[kItems];
    const u32 strip = base + wave * (kItems * kWave);
    // all of the tile's loads in flight before any is used
#pragma unroll
    for (int j = 0; j < kItems; ++j) {
        const u32 i = strip + j * kWave + lane;
        k[j] = 0;
        v[j] = 0;
        if (i < n) load_item<K, S>(b, keys, vals, seg, i, k[j], v[j]);
    }
#pragma unroll
    for (int j = 0; j < kItems; ++j) {
        const bool valid = strip + j * kWave + lane < n;
        const u32 d = digit_of(k[j], shift);
        const u64 peers = match_digit(d, valid);
        r[j] = 0xffffffffu;
        if (valid) {
            const u32 rank = lane_rank(peers);
            const u32 cur = s_cnt[wave][d];
            r[j] = cur + rank;
            if (rank == 0) s_cnt[wave][d] = cur + static_cast<u32>(__popcll(peers));
        }
    }
    __syncthreads();
    // Per digit (one per thread): offsets of each wave's run inside the tile.
    {
        const u32 d = tid;
        u32 run = 0;
        for (int w = 0; w < kWaves; ++w) {
            const u32 c = s_cnt[w][d];
            s_cnt[w][d] = run;
            run += c;
        }
        u32 total;
        const u32 lb = block_exclusive_scan<kBlock>(run, scratch, &total);
        s_lbase[d] = lb;
        const u32 gtot = totals[seg * kBins + d];
        const u32 gb = block_exclusive_scan<kBlock>(gtot, scratch, &total);
        s_gbase[d] = gb + counts[(static_cast<std::size_t>(seg) * tiles + tile) * kBins + d];
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < kItems; ++j) {
        if (r[j] != 0xffffffffu) {
            const u32 d = digit_of(k[j], shift);
            const u32 pos = s_lbase[d] + s_cnt[wave][d] + r[j];
            s_keys[pos] = k[j];
            s_vals[pos] = v[j];
        }
    }
    __syncthreads();
    const u32 cnt = (n - base) < static_cast<u32>(kTile) ? (n - base) : static_cast<u32>(kTile);
    K* ko = keys_out + seg * b.stride;
    u32* vo = vals_out + seg * b.stride;
    u32* ro = rank_out ? rank_out + seg * b.stride : nullptr;
    for (u32 i = tid; i < cnt; i += kBlock) {
        const K key = s_keys[i];
        const u32 d = digit_of(key, shift);
        const u32 g = s_gbase[d] + (i - s_lbase[d]);
        const u32 val = s_vals[i];
        ko[g] = key;
        vo[g] = val;
        if (ro) ro[val] = g;
    }
    __syncthreads();
    }
}

// Scratch words needed for counts (per batch).
inline std::size_t count_words(const Batch& b) {
    const std::size_t tiles = (b.cap + kTile - 1) / kTile;
    return static_cast<std::size_t>(b.segs) * kBins * (tiles ? tiles : 1);
}

// Sort `bits` low bits (multiple of 8) of the keys. Pass 0 reads from
// `src` (or the SAP planes); passes ping-pong between (a) and (b). Returns
// true when the result is in (a), false when in (b).
template <class K>
bool sort(hipStream_t st, const Batch& b, Source first, const K* src_keys, const u32* src_vals,
          K* a_keys, u32* a_vals, K* b_keys, u32* b_vals, int bits, u32* counts, u32* totals,
          Profiler* prof = nullptr, const char* scatter_name = nullptr, u32* rank_out = nullptr) {
    const u32 tiles = (b.cap + kTile - 1) / kTile;
    if (tiles == 0 || b.segs == 0) return true;
    // Device-counted batches loop over the tiles they need inside a
    // bounded grid, so a short list costs a short launch.
    const u32 gx = b.d_count ? (tiles < 1024u ? tiles : 1024u) : tiles;
    const dim3 grid(gx, b.segs), scan_grid(kBins, b.segs);
    const int passes = bits / 8;
    const K* in_k = src_keys;
    const u32* in_v = src_vals;
    K* out_k = b_keys;
    u32* out_v = b_vals;
    bool in_a = false;
    for (int p = 0; p < passes; ++p) {
        const int shift = 8 * p;
        if (p == 0 && first == Source::SapPlanes) {
            hipLaunchKernelGGL((k_upsweep<K, Source::SapPlanes>), grid, dim3(kBlock), 0, st, b, in_k, in_v, shift, tiles, counts);
            hipLaunchKernelGGL(k_tile_scan, scan_grid, dim3(kBlock), 0, st, b, tiles, counts, totals);
            ProfScope ps(scatter_name ? prof : nullptr, scatter_name ? scatter_name : "", st);
            hipLaunchKernelGGL((k_scatter<K, Source::SapPlanes>), grid, dim3(kBlock), 0, st, b, in_k, in_v, out_k, out_v, shift, tiles, counts, totals, p + 1 == passes ? rank_out : nullptr);
        } else {
            hipLaunchKernelGGL((k_upsweep<K, Source::Arrays>), grid, dim3(kBlock), 0, st, b, in_k, in_v, shift, tiles, counts);
            hipLaunchKernelGGL(k_tile_scan, scan_grid, dim3(kBlock), 0, st, b, tiles, counts, totals);
            ProfScope ps(scatter_name ? prof : nullptr, scatter_name ? scatter_name : "", st);
            hipLaunchKernelGGL((k_scatter<K, Source::Arrays>), grid, dim3(kBlock), 0, st, b, in_k, in_v, out_k, out_v, shift, tiles, counts, totals, p + 1 == passes ? rank_out : nullptr);
        }
        FYZZ_HIP_CHECK(hipGetLastError());
        in_a = (out_k == a_keys);
        in_k = out_k;
        in_v = out_v;
        out_k = in_a ? b_keys : a_keys;
        out_v = in_a ? b_vals : a_vals;
    }
    return in_a;
}

}  // namespace fyzz_hip::radix
