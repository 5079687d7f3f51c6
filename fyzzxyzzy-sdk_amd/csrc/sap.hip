// SAP broadphase on gfx950: box bounds + 32-bit endpoint keys, stable
// endpoint radix sort per axis, x-axis sweep for full rebuilds and an
// endpoint-crossing scan for incremental found/lost pair reports.
//
// Semantics follow the reference NativeCpu SAP (paths relative to the
// reference root):
//   bounds and key encoding        src/native/scene.cpp:25-35, 1233-1316
//   endpoint order                 src/native/scene.cpp:1341-1346
//   pair filter and sweep          src/native/scene.cpp:2597-2612, 2642-2703
//   found/lost publication         src/native/scene.cpp:3245-3296
//
// Device layout (HBM, per broadphase object of n boxes):
//   planes[6][n]       current encoded keys (min x,y,z then max x,y,z)
//   rec_cur[n]         32-B record: keys[6], body id, dynamic<<31 | stamp
//   rec_prev[n]        the record before this frame's update (moved boxes)
//   coarse[n]          16-B truncated bounds covering old and new keys
//   sorted[2][3][2n]   (key, payload) per axis, current and previous frame;
//                      payload = slot << 1 | is_max
//   split[2][3][2][n]  gathered walks only: each sorted array split by
//                      endpoint kind (key, payload, coarse bounds of the
//                      other two axes), with each position's index in its
//                      kind's array
//   pairs[2][cap]      sorted u64 (lo slot << 32 | hi slot)
// Boxes are identified by slot (0..n-1, increasing shape id), so slot
// order is shape-id order and every tie-break matches the reference.
#include <algorithm>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/fyzz_hip.h"
#include "context.hpp"
#include "stages.hpp"
#include "box_keys.hpp"
#include "radix.hpp"
#include "select.hpp"

namespace fyzz_hip::sap {

constexpr u64 kFoundBit = 1ull << 63;
constexpr u64 kPairMask = ~kFoundBit;

struct alignas(32) Rec {
    u32 k[8];
};

using namespace fyzz_hip::keys;

// Counters of one broadphase object (device words).
enum Counter : u32 {
    kMoved = 0,
    kDelta,
    kUnique,
    kFound,
    kLost,
    kPairs0,
    kPairs1,
    kKept,
    kError,     // non-finite bounds seen
    kOverflow,  // a capacity was exceeded
    kSweep,
    kLarge,     // this frame's candidate list took the multi-workgroup sort
    kLongWalks, // crossing walks handed to k_cross_long
    kCounters
};

__global__ void k_fill(u32* p, u32 n, u32 v) {
    const u32 i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

__global__ __launch_bounds__(256) void k_bounds_planes(u32 n, const float* pos, const float* quat,
                                                       const float* half, const float* off,
                                                       u32* planes, u32* err) {
    const u32 i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    u32 k[6];
    if (!box_keys(pos, quat, half, off, i, k)) atomicOr(err, 1u);
#pragma unroll
    for (int a = 0; a < 6; ++a) planes[static_cast<std::size_t>(a) * n + i] = k[a];
}

// Coarse bounds: per axis (key_lo >> 16) << 16 | (key_hi >> 16). The
// encoded keys are monotone in the coordinate, so truncating both ends
// keeps every overlap: two boxes whose key intervals overlap on an axis
// have overlapping coarse intervals there. 16 B per box (the fourth word
// is unused), small enough to stay in L2 while the crossing walks test
// candidates against it before loading full records.
__device__ inline u32 coarse_axis(u32 lo, u32 hi) { return (lo & 0xffff0000u) | (hi >> 16); }
__device__ inline bool coarse_overlap(const uint4& a, const uint4& b) {
    const u32 av[3] = {a.x, a.y, a.z}, bv[3] = {b.x, b.y, b.z};
    bool ov = true;
#pragma unroll
    for (int k = 0; k < 3; ++k) ov &= !((av[k] >> 16) > (bv[k] & 0xffffu) || (bv[k] >> 16) > (av[k] & 0xffffu));
    return ov;
}

// The coarse bounds of a sorted entry's box on the two axes other than its
// array's axis (axis a keeps axes (a+1)%3, (a+2)%3), gathered into sorted
// order so that a crossing walk reads them in sequence with the keys and
// payloads. The walked axis itself is left out: two endpoints that cross
// on it are close there anyway.
__device__ inline bool coarse_axis_overlap(u32 a, u32 b) { return ((a >> 16) <= (b & 0xffffu)) & ((b >> 16) <= (a & 0xffffu)); }
// (bitwise, not short-circuit: no branches in the walk loop)
__device__ inline bool coarse_overlap2(const uint2& a, const uint2& b) {
    return coarse_axis_overlap(a.x, b.x) & coarse_axis_overlap(a.y, b.y);
}
// The two components for axis `axis` of bounds already loaded as a value
// (scalar selects; a conditional copy of the vector type compiled (ROCm
// 7.2) into a selection of field addresses that gave the z axis the wrong
// second component).
__device__ inline uint2 coarse_other(const uint4 c, u32 axis) {
    const u32 f = axis == 0u ? c.y : (axis == 1u ? c.z : c.x);
    const u32 s = axis == 0u ? c.z : (axis == 1u ? c.x : c.y);
    return make_uint2(f, s);
}
// The gathered form of a settling pile's walks: each sorted array (new and
// old, 3 axes) split by endpoint kind into two sorted arrays of n entries
// (a subsequence keeps the sorted order), each entry with its key, payload
// and the coarse bounds of the other two axes (both arrays with this
// frame's coarse bounds). A walk only ever acts on endpoints of the other
// kind, so it scans the other kind's array alone. split_pos[j]: the index
// of combined position j within its own kind's array. blockIdx.y = array *
// 3 + axis; tiles of kSplitTile positions.
constexpr u32 kSplitTile = 1024;  // 256 threads x 4 positions
struct SplitArrays {
    u32* key;      // [2 arrays][3 axes][2 kinds][n]
    u32* pay;
    uint2* cs;
    u32* pos;      // [2 arrays][3 axes][2n]
    u32* tile_max; // [6][tiles] max endpoints per tile
};
__device__ inline std::size_t split_off(u32 seg, u32 kind, u32 n) {
    return (static_cast<std::size_t>(seg) * 2u + kind) * n;
}
__global__ __launch_bounds__(256) void k_split_count(u32 n2, const u32* new_pay, const u32* old_pay, SplitArrays S) {
    const u32 seg = blockIdx.y, axis = seg % 3u;
    const u32* P = (seg >= 3u ? old_pay : new_pay) + static_cast<std::size_t>(axis) * n2;
    const u32 tiles = (n2 + kSplitTile - 1) / kSplitTile;
    __shared__ u32 scratch[256 / kWave + 1];
    for (u32 tile = blockIdx.x; tile < tiles; tile += gridDim.x) {
        u32 c = 0;
        for (u32 k = 0; k < 4; ++k) {
            const u32 j = tile * kSplitTile + k * 256u + threadIdx.x;
            if (j < n2) c += P[j] & 1u;
        }
        u32 total;
        (void)block_exclusive_scan<256>(c, scratch, &total);
        if (threadIdx.x == 0) S.tile_max[seg * tiles + tile] = total;
        __syncthreads();
    }
}
__global__ __launch_bounds__(256) void k_split(u32 n2, const u32* new_key, const u32* new_pay, const u32* old_key,
                                               const u32* old_pay, const uint4* coarse, SplitArrays S) {
    const u32 seg = blockIdx.y, axis = seg % 3u, n = n2 >> 1;
    const std::size_t ax = static_cast<std::size_t>(axis) * n2;
    const u32* K = (seg >= 3u ? old_key : new_key) + ax;
    const u32* P = (seg >= 3u ? old_pay : new_pay) + ax;
    const u32 tiles = (n2 + kSplitTile - 1) / kSplitTile;
    __shared__ u32 scratch[256 / kWave + 1];
    for (u32 tile = blockIdx.x; tile < tiles; tile += gridDim.x) {
        // max endpoints before the tile
        u32 before = 0;
        for (u32 t = threadIdx.x; t < tile; t += 256u) before += S.tile_max[seg * tiles + t];
        u32 sum;
        (void)block_exclusive_scan<256>(before, scratch, &sum);
        before = sum;
        // the tile's positions in order: thread t takes 4 consecutive ones
        const u32 j0 = tile * kSplitTile + 4u * threadIdx.x;
        u32 pe[4], kk[4], c = 0;
#pragma unroll
        for (u32 k = 0; k < 4; ++k) {
            const u32 j = j0 + k;
            pe[k] = j < n2 ? P[j] : 0u;
            kk[k] = j < n2 ? K[j] : 0u;
            c += j < n2 ? (pe[k] & 1u) : 0u;
        }
        u32 total;
        u32 mx = before + block_exclusive_scan<256>(c, scratch, &total);
#pragma unroll
        for (u32 k = 0; k < 4; ++k) {
            const u32 j = j0 + k;
            if (j >= n2) break;
            const u32 kind = pe[k] & 1u;
            const u32 idx = kind ? mx : j - mx;
            S.pos[static_cast<std::size_t>(seg) * n2 + j] = idx;
            const std::size_t o = split_off(seg, kind, n) + idx;
            S.key[o] = kk[k];
            S.pay[o] = pe[k];
            S.cs[o] = coarse_other(coarse[pe[k] >> 1], axis);
            mx += kind;
        }
        __syncthreads();
    }
}

// rank_new / rank_old[axis * n2 + endpoint] = the endpoint's position in
// the current / previous sorted array of that axis, written for the
// endpoints of moved boxes (the only walkers); blockIdx.y = array * 3 + axis.
__global__ __launch_bounds__(256) void k_moved_ranks(u32 n2, const u32* new_pay, const u32* old_pay,
                                                     const u32* moved_bits, u32* rank_new, u32* rank_old) {
    const u32 axis = blockIdx.y % 3u;
    const bool old = blockIdx.y >= 3u;
    const std::size_t ax = static_cast<std::size_t>(axis) * n2;
    const u32* P = (old ? old_pay : new_pay) + ax;
    u32* R = (old ? rank_old : rank_new) + ax;
    for (u32 j = blockIdx.x * blockDim.x + threadIdx.x; j < n2; j += gridDim.x * blockDim.x) {
        const u32 pe = P[j];
        const u32 b = pe >> 1;
        if ((moved_bits[b >> 5] >> (b & 31u)) & 1u) R[pe] = j;
    }
}

// Full (re)build: every box gets its keys; no moved list.
__global__ __launch_bounds__(256) void k_bounds_full(u32 n, const float* pos, const float* quat,
                                                     const float* half, const float* off,
                                                     u32* planes, Rec* cur, uint4* coarse, u32 stamp, u32* cnt,
                                                     const uint8_t* sphere) {
    const u32 i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    u32 k[6];
    if (!box_keys(pos, quat, half, off, i, k, sphere)) atomicOr(&cnt[kError], 1u);
    Rec r = cur[i];
#pragma unroll
    for (int a = 0; a < 6; ++a) {
        r.k[a] = k[a];
        planes[static_cast<std::size_t>(a) * n + i] = k[a];
    }
    r.k[7] = (r.k[7] & 0x80000000u) | stamp;
    cur[i] = r;
    coarse[i] = make_uint4(coarse_axis(k[0], k[3]), coarse_axis(k[1], k[4]), coarse_axis(k[2], k[5]), 0u);
}

// Incremental: recompute the listed boxes (or all when dirty == null);
// boxes whose keys changed save their previous record, take this frame's
// stamp and join the moved list. A slot listed twice is processed once.
// A moved box's coarse bounds cover its previous and new keys; a box that
// then rests keeps them, which still covers its (unchanged) keys.
__global__ __launch_bounds__(256) void k_bounds_update(u32 count, const u32* dirty, u32 n,
                                                       const float* pos, const float* quat,
                                                       const float* half, const float* off,
                                                       u32* planes, Rec* cur, Rec* prev, uint4* coarse,
                                                       u32* seen, u32 stamp, u32* moved, u32* moved_bits,
                                                       u32* cnt, const uint8_t* sphere) {
    const u32 t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= count) return;
    const u32 i = dirty ? dirty[t] : t;
    if (i >= n) {
        atomicOr(&cnt[kError], 2u);
        return;
    }
    if (dirty && atomicExch(&seen[i], stamp) == stamp) return;
    u32 k[6];
    if (!box_keys(pos, quat, half, off, i, k, sphere)) atomicOr(&cnt[kError], 1u);
    Rec r = cur[i];
    bool changed = false;
#pragma unroll
    for (int a = 0; a < 6; ++a) changed |= r.k[a] != k[a];
    if (!changed) return;
    prev[i] = r;
    coarse[i] = make_uint4(coarse_axis(min(r.k[0], k[0]), max(r.k[3], k[3])),
                           coarse_axis(min(r.k[1], k[1]), max(r.k[4], k[4])),
                           coarse_axis(min(r.k[2], k[2]), max(r.k[5], k[5])), 0u);
#pragma unroll
    for (int a = 0; a < 6; ++a) {
        r.k[a] = k[a];
        planes[static_cast<std::size_t>(a) * n + i] = k[a];
    }
    r.k[7] = (r.k[7] & 0x80000000u) | stamp;
    cur[i] = r;
    atomicOr(&moved_bits[i >> 5], 1u << (i & 31u));
    const u32 slot = atomicAdd(&cnt[kMoved], 1u);
    if (slot < n) moved[slot] = i;
}

// Inclusive 3-axis overlap of encoded keys (scene.cpp:2606-2612).
__device__ inline bool overlap3(const Rec& a, const Rec& b) {
    return !(a.k[0] > b.k[3] || b.k[0] > a.k[3] || a.k[1] > b.k[4] || b.k[1] > a.k[4] ||
             a.k[2] > b.k[5] || b.k[2] > a.k[5]);
}
__device__ inline bool pair_allowed(const Rec& a, const Rec& b) {
    return a.k[6] != b.k[6] && ((a.k[7] | b.k[7]) & 0x80000000u);
}
__device__ inline u64 pair_key(u32 a, u32 b) {
    const u32 lo = a < b ? a : b, hi = a < b ? b : a;
    return (static_cast<u64>(lo) << 32) | hi;
}
__device__ inline void emit(u64 key, u64* out, u32* cnt, u32 counter, u32 cap) {
    const u32 s = atomicAdd(&cnt[counter], 1u);
    if (s < cap)
        out[s] = key;
    else
        atomicOr(&cnt[kOverflow], 1u);
}

// x sweep over sorted endpoints (scene.cpp:2642-2680). One thread per
// sorted position; a min endpoint of box A scans forward to A's own max
// endpoint and tests the y/z keys of every box that starts inside.
// Adjacent threads scan overlapping windows, so the payload reads of a
// wave are coalesced.
__global__ __launch_bounds__(256) void k_sweep(u32 n2, const u32* xpay, const Rec* cur,
                                               u64* out, u32* cnt, u32 cap) {
    const u32 i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n2) return;
    const u32 pa = xpay[i];
    if (pa & 1u) return;
    const u32 a = pa >> 1;
    const Rec ra = cur[a];
    const u32 stop = pa | 1u;
    for (u32 j = i + 1; j < n2; ++j) {
        const u32 pb = xpay[j];
        if (pb == stop) break;
        if (pb & 1u) continue;
        const u32 b = pb >> 1;
        const Rec rb = cur[b];
        if (!pair_allowed(ra, rb)) continue;
        if (ra.k[1] > rb.k[4] || rb.k[1] > ra.k[4] || ra.k[2] > rb.k[5] || rb.k[2] > ra.k[5]) continue;
        emit(pair_key(a, b), out, cnt, kSweep, cap);
    }
}

// Incremental crossings. A pair can change overlap status only if, on some
// axis, a min endpoint of one box and a max endpoint of the other changed
// their relative order. For every endpoint e that moved from v to v', the
// endpoints that crossed it have old or new positions inside [v, v'] (in
// the total order): scan that interval in the new array and in the old
// array. One wave per (moved box, axis); lanes stride over each interval.
// Every candidate pair is re-tested on all three axes with old and new
// records; a status change emits the pair with kFoundBit for "found".
// One crossing walk: the endpoint (a, kind) of a moved box on one axis,
// in the new array (arr 0) or the old one (arr 1).
struct Walk {
    u32 a, axis, kind, arr;
};

struct CrossArgs {
    u32 n2;
    const u32* cnt_moved;
    const u32* moved;
    const u32* new_key;
    const u32* new_pay;
    const u32* old_key;
    const u32* old_pay;
    const u32* new_rank;
    const u32* old_rank;
    const u32* moved_bits;
    const Rec* cur;
    const Rec* prev;
    const uint4* coarse;
    u64* out;
    u32* cnt;
    u32 cap;
    u32* long_walks;   // walks still inside after kShortWalk entries
    u32 long_cap;
    SplitArrays split;    // the gathered walks' kind-split arrays
};

// Entries a thread walks before handing the rest to a wave. Per-slot walks
// start at random positions, so a thread's sequential steps are separate
// line fetches, while a wave reads 64 consecutive entries per request
// group: handing every walk that crosses more than one entry to a wave
// measured fastest on 1M boxes at 10% moves (sap.pairs 1.40 ms at 16
// entries, 1.03 at 4, 0.995 at 1).
constexpr u32 kShortWalk = 1;

// Walk geometry: the endpoint's rank, the bound (its other value) and the
// direction. e sits at its new value in the new array and at its old value
// in the old one; the crossed entries lie between e and the composite
// (other value, pe), walking up or down from e.
struct WalkGeom {
    const u32* K;
    const u32* P;
    u32 own, bound_v, pe;
    bool up;
};
__device__ inline WalkGeom walk_geom(const CrossArgs& A, const Walk& w, u32 v_old, u32 v_new) {
    WalkGeom g;
    const std::size_t ax = static_cast<std::size_t>(w.axis) * A.n2;
    g.K = (w.arr ? A.old_key : A.new_key) + ax;
    g.P = (w.arr ? A.old_pay : A.new_pay) + ax;
    g.pe = (w.a << 1) | w.kind;
    g.own = (w.arr ? A.old_rank : A.new_rank)[ax + g.pe];
    g.bound_v = w.arr ? v_new : v_old;
    g.up = (w.arr != 0) == (v_new > v_old);  // old array: toward v_new
    return g;
}
// Is the step-th entry from e inside the crossed range? (j receives its
// index and pj its payload, both loaded together)
__device__ inline bool walk_inside(const CrossArgs& A, const WalkGeom& g, u32 step, u32& j, u32& pj) {
    if (g.up) {
        j = g.own + step;
        if (j >= A.n2) return false;
        const u32 kj = g.K[j];
        pj = g.P[j];
        return kj < g.bound_v || (kj == g.bound_v && pj < g.pe);
    }
    if (step > g.own) return false;
    j = g.own - step;
    const u32 kj = g.K[j];
    pj = g.P[j];
    return kj > g.bound_v || (kj == g.bound_v && pj > g.pe);
}
// A crossed entry (payload pf): a candidate pair when it is the other kind
// of endpoint of another box; re-tested on all three axes with old and new
// records. ca: the walking box's coarse bounds (its old and new keys). A
// candidate whose coarse bounds miss them on some axis overlaps it neither
// before nor after, so its status cannot change. The moved bit and the
// coarse bounds are loaded together.
__device__ inline void walk_visit(const CrossArgs& A, const Walk& w, u32 pf, const Rec& an, const Rec& ao,
                                  const uint4& ca) {
    const u32 b = pf >> 1;
    if ((pf & 1u) == w.kind || b == w.a) return;
    const u32 bits = A.moved_bits[b >> 5];
    const uint4 cb = A.coarse[b];
    const bool b_moved = (bits >> (b & 31u)) & 1u;
    // an unmoved endpoint sits at the same value in both arrays: the
    // new-array walk already saw it
    if ((b_moved || !w.arr) && coarse_overlap(ca, cb)) {
        const Rec bn = A.cur[b];
        if (pair_allowed(an, bn)) {
            const Rec bo = b_moved ? A.prev[b] : bn;
            const bool ov_new = overlap3(an, bn);
            const bool ov_old = overlap3(ao, bo);
            if (ov_new != ov_old) emit(pair_key(w.a, b) | (ov_new ? kFoundBit : 0ull), A.out, A.cnt, kDelta, A.cap);
        }
    }
}

// One thread per walk (moved box x axis x endpoint kind x array) for its
// first kShortWalk entries; longer walks continue in k_cross_long.
__global__ __launch_bounds__(256) void k_cross(CrossArgs A) {
    const u32 n_work = min(*A.cnt_moved, A.n2 >> 1) * 12u;
    for (u32 t = blockIdx.x * blockDim.x + threadIdx.x; t < n_work; t += gridDim.x * blockDim.x) {
        const u32 wa = t >> 2;
        const Walk w{A.moved[wa / 3u], wa % 3u, (t >> 1) & 1u, t & 1u};
        const Rec an = A.cur[w.a];
        const Rec ao = A.prev[w.a];
        const u32 v_old = ao.k[w.axis + 3 * w.kind];
        const u32 v_new = an.k[w.axis + 3 * w.kind];
        if (v_old == v_new) continue;
        const WalkGeom g = walk_geom(A, w, v_old, v_new);
        const uint4 ca = A.coarse[w.a];
        for (u32 step = 1;; ++step) {
            u32 j, pj = 0;
            if (!walk_inside(A, g, step, j, pj)) break;
            walk_visit(A, w, pj, an, ao, ca);
            if (step == kShortWalk) {
                const u32 slot = atomicAdd(&A.cnt[kLongWalks], 1u);
                if (slot < A.long_cap)
                    A.long_walks[slot] = t;
                else
                    atomicOr(&A.cnt[kOverflow], 4u);
                break;
            }
        }
    }
}

// The rest of the long walks, one wave each, 64 entries per iteration
// (sorted entries: the in-range lanes form a prefix).
__global__ __launch_bounds__(256) void k_cross_long(CrossArgs A) {
    const u32 lane = lane_id();
    const u32 n_long = min(A.cnt[kLongWalks], A.long_cap);
    const u32 n_waves = gridDim.x * (blockDim.x / kWave);
    for (u32 i = (blockIdx.x * blockDim.x + threadIdx.x) / kWave; i < n_long; i += n_waves) {
        const u32 t = A.long_walks[i];
        const u32 wa = t >> 2;
        const Walk w{A.moved[wa / 3u], wa % 3u, (t >> 1) & 1u, t & 1u};
        const Rec an = A.cur[w.a];
        const Rec ao = A.prev[w.a];
        const WalkGeom g = walk_geom(A, w, ao.k[w.axis + 3 * w.kind], an.k[w.axis + 3 * w.kind]);
        const uint4 ca = A.coarse[w.a];
        for (u32 base = kShortWalk;; base += kWave) {
            u32 j = 0, pj = 0;
            const bool inside = walk_inside(A, g, base + lane + 1, j, pj);
            if (inside) walk_visit(A, w, pj, an, ao, ca);
            if (__ballot(inside) != ~0ull) break;
        }
    }
}

// The same walks when the arrays were split by kind this frame (a settling
// pile: most boxes move and the walks are long). A walk scans the other
// kind's array from the walker's place in the sorted order; each step reads
// the entry's key, payload and the gathered bounds of the two other axes in
// sequence; only candidates that pass them touch per-slot data.
struct SplitWalk {
    const u32* K;
    const u32* P;
    const uint2* C;
    u32 start;  // walking up: the first other-kind index after the walker; down: one past the last before it
    u32 n;
};
__device__ inline bool walk_inside_s(const WalkGeom& g, const SplitWalk& sw, u32 step, u32& pj, uint2& cj) {
    if (g.up) {
        const u32 j = sw.start + step - 1u;
        if (j >= sw.n) return false;
        const u32 kj = sw.K[j];
        pj = sw.P[j];
        cj = sw.C[j];
        return kj < g.bound_v || (kj == g.bound_v && pj < g.pe);
    }
    if (step > sw.start) return false;
    const u32 j = sw.start - step;
    const u32 kj = sw.K[j];
    pj = sw.P[j];
    cj = sw.C[j];
    return kj > g.bound_v || (kj == g.bound_v && pj > g.pe);
}
__device__ inline SplitWalk split_walk(const CrossArgs& A, const Walk& w, u32 own) {
    const u32 seg = w.arr * 3u + w.axis, n = A.n2 >> 1;
    const u32 other = w.kind ^ 1u;
    const std::size_t o = split_off(seg, other, n);
    SplitWalk sw;
    sw.K = A.split.key + o;
    sw.P = A.split.pay + o;
    sw.C = A.split.cs + o;
    sw.n = n;
    // other-kind entries before the walker = own - (same-kind entries before it)
    sw.start = own - A.split.pos[static_cast<std::size_t>(seg) * A.n2 + own];
    return sw;
}
// ca: the walker's own coarse bounds on the same two axes (its old and new
// keys); a candidate that misses them on one of them cannot change status.
__device__ inline void walk_visit_g(const CrossArgs& A, const Walk& w, u32 pf, const uint2& cb, const uint2& ca,
                                    const Rec& an, const Rec& ao) {
    const u32 b = pf >> 1;
    if ((pf & 1u) == w.kind || b == w.a || !coarse_overlap2(ca, cb)) return;
    const bool b_moved = (A.moved_bits[b >> 5] >> (b & 31u)) & 1u;
    // an unmoved endpoint sits at the same value in both arrays: the
    // new-array walk already saw it
    if (b_moved || !w.arr) {
        const Rec bn = A.cur[b];
        if (pair_allowed(an, bn)) {
            const Rec bo = b_moved ? A.prev[b] : bn;
            const bool ov_new = overlap3(an, bn);
            const bool ov_old = overlap3(ao, bo);
            if (ov_new != ov_old) emit(pair_key(w.a, b) | (ov_new ? kFoundBit : 0ull), A.out, A.cnt, kDelta, A.cap);
        }
    }
}
// Entries a gathered walk's thread takes before handing the rest to a wave.
constexpr u32 kShortWalkG = 256;
// Gathered walks are issued in sorted order: work item t is position j of
// array (arr, axis), t = (arr * 3 + axis) * n2 + j, and walks when its
// endpoint belongs to a moved box. Neighbouring lanes then walk
// neighbouring, overlapping windows and share their cache lines, and the
// walker's rank is its position.
__device__ inline bool walker_at(const CrossArgs& A, u32 t, Walk& w, u32& own) {
    const u32 seg = t / A.n2;
    own = t - seg * A.n2;
    w.arr = seg / 3u;
    w.axis = seg - 3u * w.arr;
    const u32 pe = (w.arr ? A.old_pay : A.new_pay)[static_cast<std::size_t>(w.axis) * A.n2 + own];
    w.a = pe >> 1;
    w.kind = pe & 1u;
    return (A.moved_bits[w.a >> 5] >> (w.a & 31u)) & 1u;
}
__device__ inline WalkGeom walk_geom_at(const CrossArgs& A, const Walk& w, u32 own, u32 v_old, u32 v_new) {
    WalkGeom g;
    const std::size_t ax = static_cast<std::size_t>(w.axis) * A.n2;
    g.K = (w.arr ? A.old_key : A.new_key) + ax;
    g.P = (w.arr ? A.old_pay : A.new_pay) + ax;
    g.pe = (w.a << 1) | w.kind;
    g.own = own;
    g.bound_v = w.arr ? v_new : v_old;
    g.up = (w.arr != 0) == (v_new > v_old);
    return g;
}
__global__ __launch_bounds__(256) void k_cross_g(CrossArgs A) {
    const u32 n_work = 6u * A.n2;
    for (u32 t = blockIdx.x * blockDim.x + threadIdx.x; t < n_work; t += gridDim.x * blockDim.x) {
        Walk w;
        u32 own;
        if (!walker_at(A, t, w, own)) continue;
        const Rec an = A.cur[w.a];
        const Rec ao = A.prev[w.a];
        const u32 v_old = ao.k[w.axis + 3 * w.kind];
        const u32 v_new = an.k[w.axis + 3 * w.kind];
        if (v_old == v_new) continue;
        const WalkGeom g = walk_geom_at(A, w, own, v_old, v_new);
        const SplitWalk sw = split_walk(A, w, own);
        const uint2 ca = coarse_other(A.coarse[w.a], w.axis);
        for (u32 step = 1;; ++step) {
            u32 pj = 0;
            uint2 cj;
            if (!walk_inside_s(g, sw, step, pj, cj)) break;
            walk_visit_g(A, w, pj, cj, ca, an, ao);
            if (step == kShortWalkG) {
                const u32 slot = atomicAdd(&A.cnt[kLongWalks], 1u);
                if (slot < A.long_cap)
                    A.long_walks[slot] = t;
                else
                    atomicOr(&A.cnt[kOverflow], 4u);
                break;
            }
        }
    }
}
__global__ __launch_bounds__(256) void k_cross_long_g(CrossArgs A) {
    const u32 lane = lane_id();
    const u32 n_long = min(A.cnt[kLongWalks], A.long_cap);
    const u32 n_waves = gridDim.x * (blockDim.x / kWave);
    for (u32 i = (blockIdx.x * blockDim.x + threadIdx.x) / kWave; i < n_long; i += n_waves) {
        const u32 t = A.long_walks[i];
        Walk w;
        u32 own;
        (void)walker_at(A, t, w, own);
        const Rec an = A.cur[w.a];
        const Rec ao = A.prev[w.a];
        const WalkGeom g = walk_geom_at(A, w, own, ao.k[w.axis + 3 * w.kind], an.k[w.axis + 3 * w.kind]);
        const SplitWalk sw = split_walk(A, w, own);
        const uint2 ca = coarse_other(A.coarse[w.a], w.axis);
        for (u32 base = kShortWalkG;; base += kWave) {
            u32 pj = 0;
            uint2 cj = make_uint2(0, 0);
            const bool inside = walk_inside_s(g, sw, base + lane + 1, pj, cj);
            if (inside) walk_visit_g(A, w, pj, cj, ca, an, ao);
            if (__ballot(inside) != ~0ull) break;
        }
    }
}

// ---- compaction predicates / writers ----
struct Unique {
    const u64* a;
    __device__ bool operator()(u32 i) const { return i == 0 || a[i] != a[i - 1]; }
};
__device__ inline bool contains(const u64* a, u32 n, u64 x) {
    u32 lo = 0, hi = n;
    while (lo < hi) {
        const u32 mid = (lo + hi) >> 1;
        if (a[mid] < x)
            lo = mid + 1;
        else
            hi = mid;
    }
    return lo < n && a[lo] == x;
}
__device__ inline u32 lower_bound64(const u64* a, u32 n, u64 x) {
    u32 lo = 0, hi = n;
    while (lo < hi) {
        const u32 mid = (lo + hi) >> 1;
        if (a[mid] < x)
            lo = mid + 1;
        else
            hi = mid;
    }
    return lo;
}
struct NotIn {
    const u64* a;
    const u64* other;
    const u32* n_other;
    __device__ bool operator()(u32 i) const { return !contains(other, *n_other, a[i]); }
};
struct Copy {
    const u64* src;
    u64* dst;
    __device__ void operator()(u32 i, u32 o) const { dst[o] = src[i]; }
};

// This frame's candidate changes (pair keys with kFoundBit for "found",
// duplicates possible) -> sorted unique lost list, then found list, in one
// launch of kListBlocks workgroups (plain launch; every workgroup of the
// small grid is resident, and the grid barrier's spin is bounded).
//   n <= kSmall: workgroup 0 sorts the list in LDS (bitonic) and splits it;
//                the others return.
//   larger:      the keys are distributed into kBuckets key ranges (found
//                bit, then the lower slot's range: ranges in key order),
//                each bucket is sorted in LDS by one workgroup, then a
//                grid-wide scan of unique entries splits the sorted list.
//                If a bucket exceeds kSmall entries, the list is instead
//                sorted as kSmall-entry LDS chunks followed by rounds of
//                pairwise merges (merge path, 16 outputs per thread).
constexpr u32 kSmall = 8192;
constexpr u32 kListThreads = 1024;
constexpr u32 kListBlocks = 64;
constexpr u32 kBuckets = 128;  // 2 (found bit) x 64 lower-slot ranges
constexpr u32 kMergeRun = 16;  // outputs per thread in a merge round
// count, generation, per-workgroup totals, per-workgroup bucket histograms
constexpr u32 kListBarWords = 2 + 2 * kListBlocks + kListBlocks * kBuckets;
constexpr u32 kListSpinLimit = 1u << 22;

// bar[0] arrivals, bar[1] generation. Relaxed agent-scope atomics with a
// release fence before arriving and an acquire fence after leaving, so
// global writes before the barrier are visible after it.
__device__ inline void list_grid_sync(u32* bar, u32& gen, u32* cnt) {
    __syncthreads();
    if (threadIdx.x == 0) {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        const u32 g0 = gen;
        if (__hip_atomic_fetch_add(&bar[0], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == gridDim.x - 1) {
            __hip_atomic_store(&bar[0], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "agent");
            __hip_atomic_store(&bar[1], g0 + 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        } else {
            u32 spins = 0;
            while (__hip_atomic_load(&bar[1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == g0) {
                if (++spins > kListSpinLimit) {
                    atomicOr(&cnt[kError], 4u);  // workgroups not co-resident
                    break;
                }
                __builtin_amdgcn_s_sleep(1);
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        gen = g0 + 1u;
    }
    __syncthreads();
}

// Bitonic sort of s[0, p2) in LDS (p2 a power of two >= 64) by the workgroup.
__device__ inline void lds_bitonic(u64* s, u32 p2) {
    const u32 tid = threadIdx.x;
    for (u32 k = 2; k <= p2; k <<= 1) {
        for (u32 j = k >> 1; j > 0; j >>= 1) {
            for (u32 i = tid; i < p2; i += kListThreads) {
                const u32 ixj = i ^ j;
                if (ixj > i) {
                    const u64 a = s[i], b = s[ixj];
                    const bool asc = (i & k) == 0;
                    if ((a > b) == asc) {
                        s[i] = b;
                        s[ixj] = a;
                    }
                }
            }
            __syncthreads();
        }
    }
}

// Merge-path split of output diagonal d of runs A (la) and B (lb): i entries
// come from A (A first on equal keys).
__device__ inline u32 merge_split(const u64* A, u32 la, const u64* B, u32 lb, u32 d) {
    u32 lo = d > lb ? d - lb : 0u, hi = d < la ? d : la;
    while (lo < hi) {
        const u32 mid = (lo + hi) >> 1;
        if (A[mid] <= B[d - mid - 1])
            lo = mid + 1;
        else
            hi = mid;
    }
    return lo;
}

// Bucket of a pair key: found bit, then the lower slot >> shift (slots below
// 2^(shift + 6)), so buckets are key ranges in increasing order.
__device__ inline u32 key_bucket(u64 x, u32 shift) {
    return static_cast<u32>(x >> 63) * (kBuckets / 2) + min(static_cast<u32>((x >> 32) & 0x7fffffffu) >> shift,
                                                             kBuckets / 2 - 1);
}

__global__ __launch_bounds__(kListThreads) void k_lists(const u64* d_in, u64* tmp, const u32* d_n, u32 cap,
                                                        u64* found, u64* lost, u32* cnt, u32* bar,
                                                        u32 slot_shift) {
    __shared__ u64 s[kSmall];
    __shared__ u32 scratch[kListThreads / kWave + 1];
    __shared__ u32 lhist[kBuckets], loff[kBuckets], ltot[kBuckets];
    __shared__ u32 s_overflow;
    const u32 n = min(*d_n, cap);
    const u32 tid = threadIdx.x;
    if (n <= kSmall) {
        if (blockIdx.x != 0) return;
        u32 p2 = 64;
        while (p2 < n) p2 <<= 1;
        for (u32 i = tid; i < p2; i += kListThreads) s[i] = i < n ? d_in[i] : ~0ull;
        __syncthreads();
        lds_bitonic(s, p2);
        constexpr u32 per = kSmall / kListThreads;
        const u32 first = tid * per;
        u32 cu = 0, cl = 0;
        for (u32 q = 0; q < per; ++q) {
            const u32 i = first + q;
            if (i < n && (i == 0 || s[i] != s[i - 1])) {
                ++cu;
                cl += (s[i] & kFoundBit) ? 0u : 1u;
            }
        }
        u32 total_u, total_l;
        u32 u = block_exclusive_scan<kListThreads>(cu, scratch, &total_u);
        (void)block_exclusive_scan<kListThreads>(cl, scratch, &total_l);
        for (u32 q = 0; q < per; ++q) {
            const u32 i = first + q;
            if (i < n && (i == 0 || s[i] != s[i - 1])) {
                const u64 x = s[i];
                if (u < total_l)
                    lost[u] = x;
                else
                    found[u - total_l] = x & kPairMask;
                ++u;
            }
        }
        if (tid == 0) {
            cnt[kUnique] = total_u;
            cnt[kLost] = total_l;
            cnt[kFound] = total_u - total_l;
        }
        return;
    }
    if (tid == 0) cnt[kLarge] = 1u;
    u32 gen = tid == 0 ? __hip_atomic_load(&bar[1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : 0u;
    u32* ghist = bar + 2 + 2 * kListBlocks;  // [workgroup][bucket]
    const u32 per_block = (n + gridDim.x - 1) / gridDim.x;
    const u32 b0 = min(n, blockIdx.x * per_block), b1 = min(n, b0 + per_block);
    // bucket histogram of this workgroup's contiguous range
    for (u32 k = tid; k < kBuckets; k += kListThreads) lhist[k] = 0;
    __syncthreads();
    for (u32 i = b0 + tid; i < b1; i += kListThreads) atomicAdd(&lhist[key_bucket(d_in[i], slot_shift)], 1u);
    __syncthreads();
    for (u32 k = tid; k < kBuckets; k += kListThreads) ghist[blockIdx.x * kBuckets + k] = lhist[k];
    list_grid_sync(bar, gen, cnt);
    // bucket starts and this workgroup's offsets; any bucket above kSmall
    // sends every workgroup to the merge path (all see the same totals)
    if (tid == 0) s_overflow = 0;
    __syncthreads();
    for (u32 k = tid; k < kBuckets; k += kListThreads) {
        u32 total = 0, mine = 0;
        for (u32 b = 0; b < gridDim.x; ++b) {
            const u32 h = ghist[b * kBuckets + k];
            if (b < blockIdx.x) mine += h;
            total += h;
        }
        ltot[k] = total;
        loff[k] = mine;
        if (total > kSmall) s_overflow = 1;
    }
    __syncthreads();
    const bool overflow = s_overflow != 0;
    if (!overflow) {
        if (tid == 0) {  // bucket starts (exclusive scan of the totals)
            u32 run = 0;
            for (u32 k = 0; k < kBuckets; ++k) {
                lhist[k] = run;
                run += ltot[k];
            }
        }
        __syncthreads();
        for (u32 k = tid; k < kBuckets; k += kListThreads) loff[k] += lhist[k];
        __syncthreads();
        // scatter into bucket order (order inside a bucket is arbitrary: the
        // bucket is sorted next)
        for (u32 i = b0 + tid; i < b1; i += kListThreads) {
            const u64 x = d_in[i];
            tmp[atomicAdd(&loff[key_bucket(x, slot_shift)], 1u)] = x;
        }
        list_grid_sync(bar, gen, cnt);
        for (u32 k = blockIdx.x; k < kBuckets; k += gridDim.x) {
            const u32 start = lhist[k], len = ltot[k];
            if (len == 0) continue;
            u32 p2 = 64;
            while (p2 < len) p2 <<= 1;
            for (u32 i = tid; i < p2; i += kListThreads) s[i] = i < len ? tmp[start + i] : ~0ull;
            __syncthreads();
            lds_bitonic(s, p2);
            for (u32 i = tid; i < len; i += kListThreads) tmp[start + i] = s[i];
            __syncthreads();
        }
        list_grid_sync(bar, gen, cnt);
    }
    const u64* src = tmp;
    if (overflow) {
    // chunks sorted in LDS: d_in -> tmp
    const u32 chunks = (n + kSmall - 1) / kSmall;
    for (u32 c = blockIdx.x; c < chunks; c += gridDim.x) {
        const u32 base = c * kSmall, len = min(kSmall, n - base);
        for (u32 i = tid; i < kSmall; i += kListThreads) s[i] = i < len ? d_in[base + i] : ~0ull;
        __syncthreads();
        lds_bitonic(s, kSmall);
        for (u32 i = tid; i < len; i += kListThreads) tmp[base + i] = s[i];
        __syncthreads();
    }
    list_grid_sync(bar, gen, cnt);
    // merge rounds: tmp -> d_in scratch -> ... (d_in is no longer needed)
    u64* msrc = tmp;
    u64* dst = const_cast<u64*>(d_in);
    const u32 nthreads = gridDim.x * kListThreads;
    const u32 gt = blockIdx.x * kListThreads + tid;
    const u32 tiles = (n + kMergeRun - 1) / kMergeRun;
    for (u32 L = kSmall; L < n; L <<= 1) {
        for (u32 t = gt; t < tiles; t += nthreads) {
            const u32 o0 = t * kMergeRun;
            const u32 base = o0 / (2 * L) * (2 * L);
            const u32 la = min(L, n - base);
            const u32 lb = min(L, n - base - la);
            const u64* A = msrc + base;
            const u64* B = A + la;
            const u32 d = o0 - base;
            const u32 end = min(o0 + kMergeRun, n) - base;
            u32 i = merge_split(A, la, B, lb, d);
            u32 j = d - i;
            for (u32 o = d; o < end; ++o) {
                const bool take_a = j >= lb || (i < la && A[i] <= B[j]);
                dst[base + o] = take_a ? A[i++] : B[j++];
            }
        }
        list_grid_sync(bar, gen, cnt);
        u64* t2 = msrc;
        msrc = dst;
        dst = t2;
    }
    src = msrc;
    }
    // unique entries -> lost (no kFoundBit, sorted first) then found
    const u32 per_thread = (per_block + kListThreads - 1) / kListThreads;
    const u32 t0 = min(b1, b0 + tid * per_thread), t1 = min(b1, t0 + per_thread);
    u32 cu = 0, cl = 0;
    for (u32 i = t0; i < t1; ++i) {
        const u64 x = src[i];
        if (i == 0 || x != src[i - 1]) {
            ++cu;
            cl += (x & kFoundBit) ? 0u : 1u;
        }
    }
    u32 block_u, block_l;
    u32 u = block_exclusive_scan<kListThreads>(cu, scratch, &block_u);
    (void)block_exclusive_scan<kListThreads>(cl, scratch, &block_l);
    if (tid == 0) {
        bar[2 + 2 * blockIdx.x] = block_u;
        bar[3 + 2 * blockIdx.x] = block_l;
    }
    list_grid_sync(bar, gen, cnt);
    u32 before_u = 0, total_u = 0, total_l = 0;
    for (u32 b = 0; b < gridDim.x; ++b) {
        const u32 bu = bar[2 + 2 * b], bl = bar[3 + 2 * b];
        if (b < blockIdx.x) before_u += bu;
        total_u += bu;
        total_l += bl;
    }
    u += before_u;
    for (u32 i = t0; i < t1; ++i) {
        const u64 x = src[i];
        if (i == 0 || x != src[i - 1]) {
            if (u < total_l)
                lost[u] = x;
            else
                found[u - total_l] = x & kPairMask;
            ++u;
        }
    }
    if (blockIdx.x == 0 && tid == 0) {
        cnt[kUnique] = total_u;
        cnt[kLost] = total_l;
        cnt[kFound] = total_u - total_l;
    }
}

// new list = (old - lost) + found in one pass: lost is a subset of old and
// found is disjoint from it, so every output position is a rank sum.
// src[pos]: the new pair's index in the old list (kNone: found this frame),
// which the caller would otherwise find by a binary search per pair.
__global__ void k_apply_delta(u32 cap, const u64* old, const u32* d_no, const u64* found, const u32* d_nf,
                              const u64* lost, const u32* d_nl, u64* out, u32* d_nout, u32* cnt, u32* src) {
    const u32 no = *d_no, nf = *d_nf, nl = *d_nl;
    const u32 total = no - nl + nf;
    const u32 t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t == 0) {
        *d_nout = min(total, cap);
        if (total > cap) atomicOr(&cnt[kOverflow], 2u);
    }
    for (u32 i = t; i < no + nf; i += gridDim.x * blockDim.x) {
        u32 pos, from;
        u64 x;
        if (i < no) {
            x = old[i];
            const u32 l = lower_bound64(lost, nl, x);
            if (l < nl && lost[l] == x) continue;
            pos = i - l + lower_bound64(found, nf, x);
            from = i;
        } else {
            const u32 j = i - no;
            x = found[j];
            pos = lower_bound64(old, no, x) - lower_bound64(lost, nl, x) + j;
            from = 0xffffffffu;
        }
        if (pos < cap) {
            out[pos] = x;
            src[pos] = from;
        }
    }
}

struct Sap {
    fyzz_hip_context* ctx = nullptr;
    u32 n = 0;
    u32 pair_cap = 0;
    u32 delta_cap = 0;
    u32 frame = 0;
    int cur = 0;  // index of the current sorted arrays / pair list
    bool built = false;
    DevBuf<u32> long_walks;
    DevBuf<u32> planes, seen, moved, moved_bits, counters, sorted_key[2], sorted_val[2], rank[2], tmp_key, tmp_val,
        radix_counts, radix_totals, sel_tiles, dummy_a, dummy_b;
    DevBuf<Rec> rec_cur, rec_prev;
    DevBuf<uint4> coarse;
    DevBuf<u32> split_key, split_pay, split_pos, split_tiles;  // the gathered walks' kind-split arrays
    DevBuf<uint2> split_cs;
    DevBuf<u32> list_bar;          // grid barrier and totals of k_lists
    DevBuf<u64> pairs[2], work_a, work_b, found, lost, kept;
    DevBuf<u32> src;         // previous-list index of each current pair (incremental frames)
    bool src_valid = false;  // the last frame was incremental, so src is current
    const uint8_t* sphere = nullptr;  // per slot: a sphere (set by the world), or null
    std::vector<u32> host_counters;

    u32 list_cap() const { return std::max(pair_cap, delta_cap); }
};

// Pair keys (found << 63 | lo << 32 | hi, slots below 2^B) packed into
// their 2B + 1 significant bits in the same order, so the list sort runs
// ceil((2B + 1) / 8) passes instead of 8; unpacked on the way back.
__global__ void k_pack_pairs(const u32* d_n, u32 cap, const u32* gate, u64* keys, u32 bits) {
    if (gate && *gate == 0) return;
    const u32 n = min(*d_n, cap);
    const u64 mask = (1ull << bits) - 1ull;
    for (u32 i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const u64 k = keys[i];
        keys[i] = ((k >> 63) << (2 * bits)) | (((k >> 32) & mask) << bits) | (k & mask);
    }
}
__global__ void k_unpack_pairs(const u32* d_n, u32 cap, const u32* gate, const u64* src, u64* keys, u32 bits) {
    if (gate && *gate == 0) return;
    const u32 n = min(*d_n, cap);
    const u64 mask = (1ull << bits) - 1ull;
    for (u32 i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const u64 k = src[i];
        keys[i] = ((k >> (2 * bits)) << 63) | (((k >> bits) & mask) << 32) | (k & mask);
    }
}

void sort_u64(Sap& s, hipStream_t st, const u32* d_count, u64* keys, u64* tmp, const u32* gate = nullptr) {
    radix::Batch b{s.list_cap(), 1, s.list_cap(), 0, d_count, nullptr, 0, gate};
    u32 bits = 1;
    while (bits < 31 && (1ull << bits) < s.n) ++bits;  // slots < 2^bits
    const int key_bits = static_cast<int>((2 * bits + 1 + 7) / 8 * 8);
    const dim3 grid(std::min<u32>(grid_for(s.list_cap(), 256), 4096u));
    hipLaunchKernelGGL(k_pack_pairs, grid, dim3(256), 0, st, d_count, s.list_cap(), gate, keys, bits);
    const bool in_a = radix::sort<u64>(st, b, radix::Source::Arrays, keys, s.dummy_a.p, keys,
                                       s.dummy_a.p, tmp, s.dummy_b.p, key_bits, s.radix_counts.p,
                                       s.radix_totals.p);
    hipLaunchKernelGGL(k_unpack_pairs, grid, dim3(256), 0, st, d_count, s.list_cap(), gate, in_a ? keys : tmp, keys,
                       bits);
    FYZZ_HIP_CHECK(hipGetLastError());
}

void advance(Sap& s, const float* pos, const float* quat, const float* half, const float* off,
             u32 n_dirty, const u32* dirty, bool force_full) {
    hipStream_t st = s.ctx->stream;
    const u32 n = s.n, n2 = 2 * n;
    s.frame = (s.frame + 1) & 0x7fffffffu;
    if (s.frame == 0) s.frame = 1;
    const u32 stamp = s.frame;
    const bool full = force_full || !s.built;
    const int nxt = s.cur ^ 1;
    u32* cnt = s.counters.p;
    // reset the per-frame counters, keeping the current pair count
    hipLaunchKernelGGL(k_fill, dim3(1), dim3(64), 0, st, cnt, static_cast<u32>(kPairs0), 0u);
    hipLaunchKernelGGL(k_fill, dim3(1), dim3(64), 0, st, cnt + kKept, static_cast<u32>(kCounters - kKept), 0u);
    const u32 pairs_cur = s.cur == 0 ? kPairs0 : kPairs1;
    const u32 pairs_nxt = s.cur == 0 ? kPairs1 : kPairs0;
    if (!s.built) hipLaunchKernelGGL(k_fill, dim3(1), dim3(64), 0, st, cnt + pairs_cur, 1u, 0u);

    Profiler* prof = &s.ctx->prof;
    {
    ProfScope ps(prof, "sap.bounds", st);
    if (full) {
        if (n) hipLaunchKernelGGL(k_bounds_full, dim3(grid_for(n, 256)), dim3(256), 0, st, n, pos, quat, half, off,
                                  s.planes.p, s.rec_cur.p, s.coarse.p, stamp, cnt, s.sphere);
    } else {
        const u32 words = (n + 31) / 32;
        if (words) hipLaunchKernelGGL(k_fill, dim3(grid_for(words, 256)), dim3(256), 0, st, s.moved_bits.p, words, 0u);
        const u32 count = dirty ? n_dirty : n;
        if (count)
            hipLaunchKernelGGL(k_bounds_update, dim3(grid_for(count, 256)), dim3(256), 0, st, count, dirty, n, pos,
                               quat, half, off, s.planes.p, s.rec_cur.p, s.rec_prev.p, s.coarse.p, s.seen.p, stamp,
                               s.moved.p, s.moved_bits.p, cnt, s.sphere);
    }
    FYZZ_HIP_CHECK(hipGetLastError());
    }
    // endpoint sort, 3 axes in one batch, from the key planes
    if (n) {
        ProfScope ps(prof, "sap.sort", st);
        radix::Batch b{n2, 3, n2, n2, nullptr, s.planes.p, n};
        const bool in_a = radix::sort<u32>(st, b, radix::Source::SapPlanes, nullptr, nullptr,
                                           s.sorted_key[nxt].p, s.sorted_val[nxt].p, s.tmp_key.p,
                                           s.tmp_val.p, 32, s.radix_counts.p, s.radix_totals.p,
                                           prof, "sap.sort.scatter");
        if (!in_a) {
            FYZZ_HIP_CHECK(hipMemcpyAsync(s.sorted_key[nxt].p, s.tmp_key.p, sizeof(u32) * 3 * n2, hipMemcpyDeviceToDevice, st));
            FYZZ_HIP_CHECK(hipMemcpyAsync(s.sorted_val[nxt].p, s.tmp_val.p, sizeof(u32) * 3 * n2, hipMemcpyDeviceToDevice, st));
        }
    }
    if (full) {
        long pi = prof->begin("sap.pairs", st);
        if (n) hipLaunchKernelGGL(k_sweep, dim3(grid_for(n2, 256)), dim3(256), 0, st, n2, s.sorted_val[nxt].p,
                                  s.rec_cur.p, s.work_a.p, cnt, s.pair_cap);
        prof->end("sap.pairs", pi, st);
        ProfScope ps(prof, "sap.lists", st);
        sort_u64(s, st, cnt + kSweep, s.work_a.p, s.work_b.p);
        // unique -> pairs[nxt]
        sel::select(st, s.pair_cap, cnt + kSweep, Unique{s.work_a.p}, Copy{s.work_a.p, s.pairs[nxt].p},
                    s.sel_tiles.p, cnt + pairs_nxt);
        sel::select(st, s.pair_cap, cnt + pairs_nxt, NotIn{s.pairs[nxt].p, s.pairs[s.cur].p, cnt + pairs_cur},
                    Copy{s.pairs[nxt].p, s.found.p}, s.sel_tiles.p, cnt + kFound);
        sel::select(st, s.pair_cap, cnt + pairs_cur, NotIn{s.pairs[s.cur].p, s.pairs[nxt].p, cnt + pairs_nxt},
                    Copy{s.pairs[s.cur].p, s.lost.p}, s.sel_tiles.p, cnt + kLost);
    } else {
        long pi = prof->begin("sap.pairs", st);
        // The gather reads every entry once, so it pays only when the walks
        // visit many entries: when every box is recomputed or at least half
        // of them are listed (a settling pile). Otherwise the walks load
        // each candidate's bounds by slot.
        const bool gather = n && (!dirty || 2ull * n_dirty >= n);
        const SplitArrays split{s.split_key.p, s.split_pay.p, s.split_cs.p, s.split_pos.p, s.split_tiles.p};
        // per-slot walks start from their endpoint's position in each array
        // (the SapBox1D min/max indices of the reference GPU path), for the
        // moved boxes only; the sorted-order walks start at their position
        if (n && !gather)
            hipLaunchKernelGGL(k_moved_ranks, dim3(std::min<u32>(grid_for(n2, 256), 2048u), 6), dim3(256), 0, st,
                               n2, s.sorted_val[nxt].p, s.sorted_val[s.cur].p, s.moved_bits.p, s.rank[0].p,
                               s.rank[1].p);
        if (gather)
        {
            const u32 tiles = (n2 + kSplitTile - 1) / kSplitTile;
            const dim3 grid(std::min<u32>(tiles, 1024u), 6);
            hipLaunchKernelGGL(k_split_count, grid, dim3(256), 0, st, n2, s.sorted_val[nxt].p, s.sorted_val[s.cur].p,
                               split);
            hipLaunchKernelGGL(k_split, grid, dim3(256), 0, st, n2, s.sorted_key[nxt].p, s.sorted_val[nxt].p,
                               s.sorted_key[s.cur].p, s.sorted_val[s.cur].p, s.coarse.p, split);
        }
        CrossArgs A{n2, cnt + kMoved, s.moved.p, s.sorted_key[nxt].p, s.sorted_val[nxt].p, s.sorted_key[s.cur].p,
                    s.sorted_val[s.cur].p, s.rank[0].p, s.rank[1].p, s.moved_bits.p, s.rec_cur.p,
                    s.rec_prev.p, s.coarse.p, s.work_a.p, cnt, s.delta_cap, s.long_walks.p, 12u * n,
                    split};
        const dim3 short_grid(std::min<u32>(grid_for(static_cast<std::size_t>(n) * 12, 256), 8192u));
        const dim3 sorted_grid(std::min<u32>(grid_for(6ull * n2, 256), 8192u));
        const dim3 long_grid(std::min<u32>(grid_for(static_cast<std::size_t>(n) * 12 * kWave, 256), 4096u));
        if (gather) {
            hipLaunchKernelGGL(k_cross_g, sorted_grid, dim3(256), 0, st, A);
            hipLaunchKernelGGL(k_cross_long_g, long_grid, dim3(256), 0, st, A);
        } else {
            hipLaunchKernelGGL(k_cross, short_grid, dim3(256), 0, st, A);
            hipLaunchKernelGGL(k_cross_long, long_grid, dim3(256), 0, st, A);
        }
        FYZZ_HIP_CHECK(hipGetLastError());
        prof->end("sap.pairs", pi, st);
        ProfScope ps(prof, "sap.lists", st);
        u32 slot_bits = 6;
        while (slot_bits < 31 && (1ull << slot_bits) < n) ++slot_bits;  // slots < 2^slot_bits
        hipLaunchKernelGGL(k_lists, dim3(kListBlocks), dim3(kListThreads), 0, st, s.work_a.p, s.work_b.p,
                           cnt + kDelta, s.delta_cap, s.found.p, s.lost.p, cnt, s.list_bar.p, slot_bits - 6);
        hipLaunchKernelGGL(k_apply_delta, dim3(std::min<u32>(grid_for(static_cast<std::size_t>(s.pair_cap) + s.delta_cap, 256), 2048u)), dim3(256), 0, st,
                           s.pair_cap, s.pairs[s.cur].p, cnt + pairs_cur, s.found.p, cnt + kFound, s.lost.p,
                           cnt + kLost, s.pairs[nxt].p, cnt + pairs_nxt, cnt, s.src.p);
    }
    FYZZ_HIP_CHECK(hipGetLastError());
    s.cur = nxt;
    s.built = true;
    s.src_valid = !full;
}

}  // namespace fyzz_hip::sap

struct fyzz_hip_sap {
    fyzz_hip::sap::Sap s;
};

namespace fyzz_hip::sap {
std::size_t sort_keys_scratch_words(u32 cap) {
    const radix::Batch b{cap, 1, cap, 0, nullptr, nullptr, 0, nullptr};
    return radix::count_words(b) + radix::kBins;
}
void sort_keys(hipStream_t st, u64* keys, u32* vals, u64* tmp_keys, u32* tmp_vals, u32 cap, const u32* d_count,
               int bits, u32* scratch) {
    const radix::Batch b{cap, 1, cap, 0, d_count, nullptr, 0, nullptr};
    u32* counts = scratch;
    u32* totals = scratch + radix::count_words(b);
    (void)radix::sort<u64>(st, b, radix::Source::Arrays, keys, vals, keys, vals, tmp_keys, tmp_vals, bits, counts,
                           totals);
    FYZZ_HIP_CHECK(hipGetLastError());
}
bool built(fyzz_hip_sap* h) { return h->s.built; }
void set_spheres(fyzz_hip_sap* h, const uint8_t* sphere) { h->s.sphere = sphere; }
void replay_advance(fyzz_hip_sap* h) {
    Sap& s = h->s;
    s.frame = (s.frame + 1) & 0x7fffffffu;
    if (s.frame == 0) s.frame = 1;
    s.cur ^= 1;
    s.src_valid = true;  // a recorded frame is incremental
}
Lists lists(fyzz_hip_sap* h) {
    Sap& s = h->s;
    const u32 cur = s.cur == 0 ? kPairs0 : kPairs1;
    const u32 prev = s.cur == 0 ? kPairs1 : kPairs0;
    return {s.pairs[s.cur].p, s.counters.p + cur, s.pairs[s.cur ^ 1].p, s.counters.p + prev,
            s.pair_cap, s.counters.p + kError, s.counters.p + kOverflow, s.lost.p, s.counters.p + kLost,
            s.counters.p + kFound, s.src_valid ? s.src.p : nullptr};
}
}  // namespace fyzz_hip::sap

using namespace fyzz_hip;

extern "C" {

int fyzz_hip_sap_box_bounds(fyzz_hip_context* ctx, uint32_t n, const float* pos, const float* quat,
                            const float* half, const float* contact_offset, uint32_t* bounds,
                            uint32_t* error_flag) {
    return guarded(ctx, [&] {
        FYZZ_REQUIRE(n == 0 || (pos && quat && half && contact_offset && bounds && error_flag),
                     kInvalidArgument, "sap_box_bounds: null device pointer");
        if (!n) return;
        hipLaunchKernelGGL(sap::k_bounds_planes, dim3(grid_for(n, 256)), dim3(256), 0, ctx->stream, n, pos, quat,
                           half, contact_offset, bounds, error_flag);
        FYZZ_HIP_CHECK(hipGetLastError());
    });
}

int fyzz_hip_sap_sort_endpoints(fyzz_hip_context* ctx, uint32_t n, const uint32_t* bounds,
                                uint32_t* values, uint32_t* payloads, uint32_t* scratch,
                                uint64_t scratch_words) {
    return guarded(ctx, [&] {
        FYZZ_REQUIRE(n == 0 || (bounds && values && payloads && scratch), kInvalidArgument,
                     "sap_sort_endpoints: null device pointer");
        FYZZ_REQUIRE(n <= 0x7fffffffu, kInvalidArgument, "sap_sort_endpoints: n too large");
        if (!n) return;
        const u32 n2 = 2 * n;
        radix::Batch b{n2, 3, n2, n2, nullptr, bounds, n};
        const std::size_t need = 2ull * 3 * n2 + radix::count_words(b) + 3 * radix::kBins;
        FYZZ_REQUIRE(scratch_words >= need, kInvalidArgument,
                     "sap_sort_endpoints: scratch too small (need " + std::to_string(need) + " words)");
        u32* tk = scratch;
        u32* tv = tk + 3ull * n2;
        u32* counts = tv + 3ull * n2;
        u32* totals = counts + radix::count_words(b);
        const bool in_a = radix::sort<u32>(ctx->stream, b, radix::Source::SapPlanes, nullptr, nullptr, values,
                                           payloads, tk, tv, 32, counts, totals);
        if (!in_a) {
            FYZZ_HIP_CHECK(hipMemcpyAsync(values, tk, sizeof(u32) * 3 * n2, hipMemcpyDeviceToDevice, ctx->stream));
            FYZZ_HIP_CHECK(hipMemcpyAsync(payloads, tv, sizeof(u32) * 3 * n2, hipMemcpyDeviceToDevice, ctx->stream));
        }
    });
}

uint64_t fyzz_hip_sap_sort_scratch_words(uint32_t n) {
    const u32 n2 = 2 * n;
    radix::Batch b{n2, 3, n2, n2, nullptr, nullptr, n};
    return 2ull * 3 * n2 + radix::count_words(b) + 3 * radix::kBins;
}

int fyzz_hip_sap_create(fyzz_hip_context* ctx, uint32_t n, const uint32_t* body_ids,
                        const uint8_t* dynamic, uint32_t pair_capacity, uint32_t delta_capacity,
                        fyzz_hip_sap** out) {
    if (!out) return kInvalidArgument;
    *out = nullptr;
    return guarded(ctx, [&] {
        FYZZ_REQUIRE(n <= 0x3fffffffu, kInvalidArgument, "sap_create: too many boxes");
        FYZZ_REQUIRE(n == 0 || (body_ids && dynamic), kInvalidArgument, "sap_create: null host array");
        auto h = std::make_unique<fyzz_hip_sap>();
        sap::Sap& s = h->s;
        s.ctx = ctx;
        s.n = n;
        s.pair_cap = std::max<u32>(pair_capacity, 1024);
        s.delta_cap = std::max<u32>(delta_capacity, 1024);
        const u32 n2 = 2 * n;
        s.planes.reserve(6ull * n);
        s.seen.reserve(n);
        s.moved.reserve(n);
        s.long_walks.reserve(12ull * std::max<u32>(n, 1));
        s.moved_bits.reserve((n + 31) / 32 + 1);
        s.counters.reserve(sap::kCounters);
        for (int i = 0; i < 2; ++i) {
            s.sorted_key[i].reserve(3ull * n2);
            s.sorted_val[i].reserve(3ull * n2);
            s.rank[i].reserve(3ull * n2);
            s.pairs[i].reserve(s.pair_cap);
        }
        s.tmp_key.reserve(3ull * n2);
        s.tmp_val.reserve(3ull * n2);
        const u32 lcap = s.list_cap();
        radix::Batch be{n2, 3, n2, n2, nullptr, nullptr, n};
        radix::Batch bl{lcap, 1, lcap, 0, nullptr, nullptr, 0};
        s.radix_counts.reserve(std::max(radix::count_words(be), radix::count_words(bl)));
        s.radix_totals.reserve(3 * radix::kBins);
        s.sel_tiles.reserve(sel::tiles_for(lcap) + 1);
        s.dummy_a.reserve(lcap);
        s.dummy_b.reserve(lcap);
        s.rec_cur.reserve(n);
        s.rec_prev.reserve(n);
        s.coarse.reserve(n);
        // kind-split arrays of both sorted arrays: 2 x 3 x 2n entries each
        s.split_key.reserve(6ull * n2);
        s.split_pay.reserve(6ull * n2);
        s.split_cs.reserve(6ull * n2);
        s.split_pos.reserve(6ull * n2);
        s.split_tiles.reserve(6ull * ((n2 + sap::kSplitTile - 1) / sap::kSplitTile) + 1);
        s.list_bar.reserve(sap::kListBarWords);
        FYZZ_HIP_CHECK(hipMemset(s.list_bar.p, 0, sizeof(u32) * sap::kListBarWords));
        s.work_a.reserve(lcap);
        s.work_b.reserve(lcap);
        s.found.reserve(lcap);
        s.lost.reserve(lcap);
        s.src.reserve(std::max<u32>(s.pair_cap, 1));
        s.kept.reserve(s.pair_cap);
        std::vector<sap::Rec> init(n);
        for (u32 i = 0; i < n; ++i) {
            for (int k = 0; k < 6; ++k) init[i].k[k] = 0;
            init[i].k[6] = body_ids[i];
            init[i].k[7] = dynamic[i] ? 0x80000000u : 0u;
        }
        if (n) {
            FYZZ_HIP_CHECK(hipMemcpy(s.rec_cur.p, init.data(), sizeof(sap::Rec) * n, hipMemcpyHostToDevice));
            FYZZ_HIP_CHECK(hipMemcpy(s.rec_prev.p, init.data(), sizeof(sap::Rec) * n, hipMemcpyHostToDevice));
            FYZZ_HIP_CHECK(hipMemset(s.seen.p, 0, sizeof(u32) * n));
        }
        FYZZ_HIP_CHECK(hipMemset(s.counters.p, 0, sizeof(u32) * sap::kCounters));
        *out = h.release();
    });
}

int fyzz_hip_sap_destroy(fyzz_hip_sap* sap) {
    if (!sap) return kInvalidArgument;
    // Device buffers are released on the device they were allocated on;
    // the context must still be alive (destroy children first).
    try {
        (void)hipSetDevice(sap->s.ctx->device);
        delete sap;
    } catch (...) {
        return kDeviceError;
    }
    return kOk;
}

int fyzz_hip_sap_advance(fyzz_hip_sap* sap, const float* pos, const float* quat, const float* half,
                         const float* contact_offset, uint32_t n_dirty, const uint32_t* dirty_slots,
                         int force_full) {
    if (!sap) return kInvalidArgument;
    return guarded(sap->s.ctx, [&] {
        FYZZ_REQUIRE(sap->s.n == 0 || (pos && quat && half && contact_offset), kInvalidArgument,
                     "sap_advance: null device pointer");
        sap::advance(sap->s, pos, quat, half, contact_offset, n_dirty, dirty_slots, force_full != 0);
    });
}

int fyzz_hip_sap_counts(fyzz_hip_sap* sap, uint64_t* n_pairs, uint64_t* n_found, uint64_t* n_lost,
                        uint32_t* status_flags) {
    if (!sap) return kInvalidArgument;
    return guarded(sap->s.ctx, [&] {
        sap::Sap& s = sap->s;
        s.host_counters.resize(sap::kCounters);
        FYZZ_HIP_CHECK(hipMemcpyAsync(s.host_counters.data(), s.counters.p, sizeof(u32) * sap::kCounters,
                                      hipMemcpyDeviceToHost, s.ctx->stream));
        FYZZ_HIP_CHECK(hipStreamSynchronize(s.ctx->stream));
        const auto& c = s.host_counters;
        if (n_pairs) *n_pairs = c[s.cur == 0 ? sap::kPairs0 : sap::kPairs1];
        if (n_found) *n_found = c[sap::kFound];
        if (n_lost) *n_lost = c[sap::kLost];
        if (status_flags) *status_flags = (c[sap::kError] ? 1u : 0u) | (c[sap::kOverflow] ? 2u : 0u);
    });
}

// which: 0 current pair list, 1 found, 2 lost. Slot pairs (lo, hi).
int fyzz_hip_sap_fetch(fyzz_hip_sap* sap, int which, uint32_t* out_pairs, uint64_t capacity,
                       uint64_t* written) {
    if (!sap) return kInvalidArgument;
    return guarded(sap->s.ctx, [&] {
        FYZZ_REQUIRE(which >= 0 && which <= 2, kInvalidArgument, "sap_fetch: which must be 0, 1 or 2");
        uint64_t np = 0, nf = 0, nl = 0;
        uint32_t flags = 0;
        int rc = fyzz_hip_sap_counts(sap, &np, &nf, &nl, &flags);
        if (rc) throw Error(rc, sap->s.ctx->last_error);
        FYZZ_REQUIRE(!(flags & 2u), kOverflow, "sap: pair capacity exceeded");
        FYZZ_REQUIRE(!(flags & 1u), kOverflow, "sap: non-finite box bounds");
        sap::Sap& s = sap->s;
        const u64* src = which == 0 ? s.pairs[s.cur].p : which == 1 ? s.found.p : s.lost.p;
        const uint64_t n = which == 0 ? np : which == 1 ? nf : nl;
        const uint64_t m = std::min<uint64_t>(n, capacity);
        std::vector<u64> tmp(m);
        if (m) FYZZ_HIP_CHECK(hipMemcpyAsync(tmp.data(), src, sizeof(u64) * m, hipMemcpyDeviceToHost, s.ctx->stream));
        FYZZ_HIP_CHECK(hipStreamSynchronize(s.ctx->stream));
        for (uint64_t i = 0; i < m; ++i) {
            out_pairs[2 * i] = static_cast<u32>(tmp[i] >> 32);
            out_pairs[2 * i + 1] = static_cast<u32>(tmp[i] & 0xffffffffu);
        }
        if (written) *written = n;
    });
}

int fyzz_hip_sap_endpoints(fyzz_hip_sap* sap, uint32_t* values, uint32_t* payloads) {
    if (!sap) return kInvalidArgument;
    return guarded(sap->s.ctx, [&] {
        sap::Sap& s = sap->s;
        FYZZ_REQUIRE(s.built, kLogicError, "sap_endpoints: no frame advanced yet");
        FYZZ_REQUIRE(values && payloads, kInvalidArgument, "sap_endpoints: null device pointer");
        const std::size_t bytes = sizeof(u32) * 6ull * s.n;
        if (!bytes) return;
        FYZZ_HIP_CHECK(hipMemcpyAsync(values, s.sorted_key[s.cur].p, bytes, hipMemcpyDeviceToDevice, s.ctx->stream));
        FYZZ_HIP_CHECK(hipMemcpyAsync(payloads, s.sorted_val[s.cur].p, bytes, hipMemcpyDeviceToDevice, s.ctx->stream));
    });
}

}  // extern "C"

// ---- endpoint sort with a library-held workspace (no allocation) -------
//
// fyzz_hip_sap_sort_endpoints_fixed runs the same radix sort as
// fyzz_hip_sap_sort_endpoints with its workspace in `scratch`, a
// file-scope device array: every device the library runs on holds one
// copy until the process ends. Host code never passes `scratch` itself
// (on the host it names a stand-in, not device memory): the device
// address of each device's copy is looked up once per device, with the
// device current, when a context is created on it, and kept in
// g_fixed_scratch.
namespace fyzz_hip::sap {
constexpr u32 kFixedSortMaxBoxes = 131072;
constexpr std::size_t kFixedSortKeys = 2ull * kFixedSortMaxBoxes;  // endpoints per axis
constexpr std::size_t kFixedSortWords =
    2ull * 3 * kFixedSortKeys + 3ull * radix::kBins * ((kFixedSortKeys + radix::kTile - 1) / radix::kTile) +
    3ull * radix::kBins;
constexpr int kMaxDevices = 64;
}  // namespace fyzz_hip::sap

static __device__ __align__(256) fyzz_hip::u32 scratch[fyzz_hip::sap::kFixedSortWords];

namespace fyzz_hip::sap {
namespace {
std::mutex g_fixed_mu;
u32* g_fixed_scratch[kMaxDevices] = {};
}  // namespace

// Called with `device` current (context creation). Idempotent.
void bind_fixed_workspace(int device) {
    FYZZ_REQUIRE(device >= 0 && device < kMaxDevices, kInvalidArgument, "device index out of range");
    std::lock_guard<std::mutex> lock(g_fixed_mu);
    if (g_fixed_scratch[device]) return;
    void* p = nullptr;
    FYZZ_HIP_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    FYZZ_REQUIRE(p != nullptr, kDeviceError, "workspace symbol has no device address");
    g_fixed_scratch[device] = static_cast<u32*>(p);
}

u32* fixed_workspace(int device) {
    FYZZ_REQUIRE(device >= 0 && device < kMaxDevices, kInvalidArgument, "device index out of range");
    std::lock_guard<std::mutex> lock(g_fixed_mu);
    u32* p = g_fixed_scratch[device];
    FYZZ_REQUIRE(p != nullptr, kLogicError, "workspace not bound on this device (create a context first)");
    return p;
}
}  // namespace fyzz_hip::sap

extern "C" {

int fyzz_hip_sap_sort_endpoints_fixed(fyzz_hip_context* ctx, uint32_t n, const uint32_t* bounds, uint32_t* values,
                                      uint32_t* payloads) {
    return guarded(ctx, [&] {
        FYZZ_REQUIRE(n <= sap::kFixedSortMaxBoxes, kInvalidArgument,
                     "sap_sort_endpoints_fixed: n exceeds the fixed workspace (" +
                         std::to_string(sap::kFixedSortMaxBoxes) + " boxes)");
        FYZZ_REQUIRE(n == 0 || (bounds && values && payloads), kInvalidArgument,
                     "sap_sort_endpoints_fixed: null device pointer");
        if (!n) return;
        u32* ws = sap::fixed_workspace(ctx->device);
        const u32 n2 = 2 * n;
        radix::Batch b{n2, 3, n2, n2, nullptr, bounds, n};
        const std::size_t need = 2ull * 3 * n2 + radix::count_words(b) + 3 * radix::kBins;
        FYZZ_REQUIRE(need <= sap::kFixedSortWords, kLogicError, "sap_sort_endpoints_fixed: workspace sizing");
        u32* tk = ws;
        u32* tv = tk + 3ull * n2;
        u32* counts = tv + 3ull * n2;
        u32* totals = counts + radix::count_words(b);
        const bool in_a = radix::sort<u32>(ctx->stream, b, radix::Source::SapPlanes, nullptr, nullptr, values,
                                           payloads, tk, tv, 32, counts, totals);
        if (!in_a) {
            FYZZ_HIP_CHECK(hipMemcpyAsync(values, tk, sizeof(u32) * 3 * n2, hipMemcpyDeviceToDevice, ctx->stream));
            FYZZ_HIP_CHECK(hipMemcpyAsync(payloads, tv, sizeof(u32) * 3 * n2, hipMemcpyDeviceToDevice, ctx->stream));
        }
    });
}

uint32_t fyzz_hip_sap_sort_fixed_max_boxes(void) { return fyzz_hip::sap::kFixedSortMaxBoxes; }

}  // extern "C"
