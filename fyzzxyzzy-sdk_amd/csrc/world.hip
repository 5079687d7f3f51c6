// Device-resident step pipeline: body records, one broadphase over the
// dynamic boxes, the plane x box pairs, persistent box/box and plane/box
// manifolds, the touching-manager list with its friction history, and the
// Dy step, enqueued on the context's stream with no host readback between
// stages.
//
// Stands in for fyzz::World's step() over the native GPU backend
// (reference include/fyzzxyzzy/world.h:400-470, src/native/scene.cpp
// Scene::step 10940-11020) for the admitted class: one static floor plane
// and dynamic boxes. Per step:
//   poses -> SoA            (update_shape_bounds input, scene.cpp:1233)
//   broadphase advance      (update_bounds/update_pairs, 1336-2704)
//   plane x every box       (update_gpu_pairs, 2396-2418)
//   box/box + plane/box PCM (update_contacts, 4012-4758)
//   touching managers       (the descriptor stream; see below)
//   Dy                      (6940-10707)
// Manager (solver) order: touching box/box managers in broadphase pair
// order (lo slot, hi slot), then touching plane managers in slot order.
// The reference orders descriptors by its island contact lists
// (scene.cpp:7174-7260), which this pipeline does not rebuild; every stage
// reproduces the reference for a given order.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <iterator>
#include <memory>
#include <string>
#include <vector>

#include "../../include/fyzz_hip.h"
#include "context.hpp"
#include "box_keys.hpp"
#include "select.hpp"
#include "stages.hpp"

namespace fyzz_hip::world {

constexpr float kContactOffset = 0.02f;    // per shape
constexpr float kContactDistance = 0.04f;  // sum of the pair's offsets
constexpr u32 kNone = 0xffffffffu;
constexpr u32 kCacheWords = 56;            // 224-byte manifold cache record
constexpr u32 kTotals = 9;                 // running totals (fyzz_hip_world_totals)

struct alignas(16) CacheRec {
    u32 w[kCacheWords];
};

__global__ void k_gather(u32 n, const fyzz_dy_body* bodies, float* pos, float* quat) {
    const u32 i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const fyzz_dy_body& b = bodies[i];
    pos[3 * i] = b.position[0];
    pos[3 * i + 1] = b.position[1];
    pos[3 * i + 2] = b.position[2];
    quat[4 * i] = b.orientation[0];
    quat[4 * i + 1] = b.orientation[1];
    quat[4 * i + 2] = b.orientation[2];
    quat[4 * i + 3] = b.orientation[3];
}

__global__ void k_fill(u32* p, u32 n, u32 v) {
    for (u32 i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = v;
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

// Gu order of a box/box pair (first = higher shape id, normals in the
// second's frame), the pair's previous-list index (its cache and friction
// history) and the broadphase pair count per box.
// known: the broadphase's own previous-list index of each pair (an
// incremental frame records it), else found by a binary search.
__global__ void k_pairs(u32 cap, const u64* cur, const u32* d_cur, const u64* prev, const u32* d_prev, u32* first,
                        u32* second, u32* src, u32* counted, const u32* known, u32 n_dyn) {
    const u32 n = min(*d_cur, cap), np = min(*d_prev, cap);
    for (u32 t = blockIdx.x * blockDim.x + threadIdx.x; t < n; t += gridDim.x * blockDim.x) {
        const u64 x = cur[t];
        const u32 lo = static_cast<u32>(x >> 32), hi = static_cast<u32>(x & 0xffffffffu);
        // Gu order: two dynamic bodies the higher id first; a static body
        // (slots from n_dyn on) second (scene.cpp:3278-3287)
        const bool stat = hi >= n_dyn;
        first[t] = stat ? lo : hi;
        second[t] = stat ? hi : lo;
        if (known) {
            src[t] = known[t];
        } else {
            const u32 j = lower_bound64(prev, np, x);
            src[t] = (j < np && prev[j] == x) ? j : kNone;
        }
        atomicAdd(&counted[lo], 1u);
        if (!stat) atomicAdd(&counted[hi], 1u);
    }
}

__global__ void k_counted(u32 n, fyzz_dy_body* bodies, const u32* counted, u32 planes) {
    const u32 i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) bodies[i].counted_interactions = counted[i] + planes;
}

__global__ void k_step_begin(u32* d_step) { *d_step += 1u; }

__global__ void k_total(const u32* d_nbb, u32 cap, u32 extra, u32* d_total) {
    *d_total = min(*d_nbb, cap) + extra;
}

// A touching manager enters the solver list unless its island sleeps
// (body0 is dynamic in both kinds; a box/box manager's bodies share an
// island, so they sleep together).
struct MgrPred {
    const u32* d_nbb;
    u32 cap;
    const u32* counts_bb;
    const u32* counts_pl;
    const u32* first;
    const fyzz_dy_body* bodies;
    __device__ bool operator()(u32 i) const {
        const u32 nbb = min(*d_nbb, cap);
        const bool touching = i < nbb ? counts_bb[i] != 0u : counts_pl[i - nbb] != 0u;
        if (!touching) return false;
        const u32 b0 = i < nbb ? first[i] : i - nbb;
        return !(bodies[b0].flags & 1u);
    }
};

struct MgrWrite {
    const u32* d_nbb;
    u32 cap;
    const u32* first;
    const u32* second;
    const u32* src;
    const u32* counts_bb;
    const float* contacts_bb;
    const u32* counts_pl;
    const float* contacts_pl;
    const u32* stamp_prev;
    const u32* d_step;  // this step's number (1, 2, ...); history must carry the previous one
    fyzz_dy_manager* managers;
    u32* mgr_slot;
    u32 n_dyn;  // slots from here on are static bodies (solver static 1 + slot - n_dyn)
    __device__ void operator()(u32 i, u32 o) const {
        const u32 nbb = min(*d_nbb, cap);
        const u32 step = *d_step;
        const u32 want_stamp = step > 1u ? step - 1u : kNone;
        fyzz_dy_manager M;
        const float* c;
        u32 slot, hist_slot;
        if (i < nbb) {
            M.body0 = static_cast<int>(first[i]);
            M.body1 = second[i] < n_dyn ? static_cast<int>(second[i]) : -2 - static_cast<int>(second[i] - n_dyn);
            M.count = counts_bb[i];
            c = contacts_bb + 28ull * i;
            slot = i;
            hist_slot = src[i];
        } else {
            const u32 b = i - nbb;
            M.body0 = static_cast<int>(b);
            M.body1 = -1;
            M.count = counts_pl[b];
            c = contacts_pl + 28ull * b;
            slot = cap + b;
            hist_slot = slot;
        }
        M.history = (hist_slot != kNone && want_stamp != kNone && stamp_prev[hist_slot] == want_stamp) ? hist_slot
                                                                                                       : kNone;
        for (u32 k = 0; k < 4; ++k) {
            fyzz_dy_contact& d = M.contacts[k];
            const bool live = k < M.count;
            d.position[0] = live ? c[7 * k] : 0.f;
            d.position[1] = live ? c[7 * k + 1] : 0.f;
            d.position[2] = live ? c[7 * k + 2] : 0.f;
            d.normal[0] = live ? c[7 * k + 3] : 0.f;
            d.normal[1] = live ? c[7 * k + 4] : 0.f;
            d.normal[2] = live ? c[7 * k + 5] : 0.f;
            d.separation = live ? c[7 * k + 6] : 0.f;
        }
        managers[o] = M;
        mgr_slot[o] = slot;
    }
};

// the step's Dy inputs with each manager's history moved to its own index
// (the differentiable form records the contact sets this way)
__global__ void k_record_dy(const u32* d_nm, u32 cap_m, const fyzz_dy_manager* managers,
                            const fyzz_dy_patch* hist, fyzz_dy_manager* out_m, fyzz_dy_patch* out_h) {
    const u32 n = min(*d_nm, cap_m);
    for (u32 m = blockIdx.x * blockDim.x + threadIdx.x; m < n; m += gridDim.x * blockDim.x) {
        fyzz_dy_manager M = managers[m];
        if (M.history != kNone) {
            out_h[m] = hist[M.history];
            M.history = m;
        } else {
            out_h[m] = fyzz_dy_patch{};
        }
        out_m[m] = M;
    }
}

// The same records appended at *cursor of caller arenas, without a host
// round trip: nothing is copied when they do not fit; k_append_advance then
// poisons the cursor (all ones) so the caller learns it once, afterwards.
__global__ void k_append_dy(const u32* d_nm, u32 cap_m, const fyzz_dy_manager* managers, const fyzz_dy_patch* hist,
                            fyzz_dy_manager* out_m, fyzz_dy_patch* out_h, u64 capacity, const u64* cursor) {
    const u32 n = min(*d_nm, cap_m);
    const u64 c = *cursor;
    if (c == ~0ull || c + n > capacity) return;
    for (u32 m = blockIdx.x * blockDim.x + threadIdx.x; m < n; m += gridDim.x * blockDim.x) {
        fyzz_dy_manager M = managers[m];
        if (M.history != kNone) {
            out_h[c + m] = hist[M.history];
            M.history = m;  // index within this step's records
        } else {
            out_h[c + m] = fyzz_dy_patch{};
        }
        out_m[c + m] = M;
    }
}
__global__ void k_append_advance(const u32* d_nm, u32 cap_m, u64 capacity, u64* cursor, u32* count) {
    const u32 n = min(*d_nm, cap_m);
    const u64 c = *cursor;
    *count = n;
    *cursor = (c == ~0ull || c + n > capacity) ? ~0ull : c + n;
}

// friction patches solved this step, keyed for the next step's lookup
__global__ void k_scatter_patches(const u32* d_nm, u32 cap_m, const fyzz_dy_patch* patches, const u32* mgr_slot,
                                  fyzz_dy_patch* hist_next, u32* stamp_next, const u32* d_step) {
    const u32 n = min(*d_nm, cap_m);
    const u32 stamp = *d_step;
    for (u32 m = blockIdx.x * blockDim.x + threadIdx.x; m < n; m += gridDim.x * blockDim.x) {
        const u32 s = mgr_slot[m];
        hist_next[s] = patches[m];
        stamp_next[s] = stamp;
    }
}

// ---- per-scene broadphase (independent scenes in one world)
//
// Scene s owns box slots [first[s], first[s+1]). One workgroup per scene
// computes the boxes' keys (the broadphase's encoding) into LDS and tests
// every pair of the scene once with the broadphase's pair filter (different
// bodies, at least one dynamic, inclusive 3-axis overlap of the keys,
// scene.cpp:2606-2612), so the pair set equals the SAP's for that scene.
// The test writes each box's row of partners as a bitmask (bit j of row i:
// box j > i overlaps box i) and counts the row; after the scan of the
// scenes' totals, a second pass scans the rows and writes each row's
// partners in increasing order: rows in (lo, hi) order, scenes concatenated
// in slot order, so the whole list is sorted like the SAP's.
constexpr u32 kSceneMax = 1024;  // boxes per scene in this path
constexpr int kSceneBlock = 256;
__device__ inline u32 scene_words(u32 n) { return (n + 31u) / 32u; }

__global__ __launch_bounds__(kSceneBlock) void k_scene_masks(u32 n_scenes, const u32* first, const u32* mask_off,
                                                             const float* pos, const float* quat, const float* half,
                                                             const float* offset, u32* mask, u32* row_cnt,
                                                             u32* scene_cnt, u32* status, const uint8_t* sphere) {
    // a box's keys as (min x, min y, min z, max x) and (max y, max z): two
    // LDS reads per tested pair
    __shared__ uint4 klo[kSceneMax];
    __shared__ uint2 khi[kSceneMax];
    __shared__ u32 scratch[kSceneBlock / kWave + 1];
    for (u32 sc = blockIdx.x; sc < n_scenes; sc += gridDim.x) {
        const u32 b0 = first[sc], n = first[sc + 1] - b0;
        if (n > kSceneMax) {
            if (threadIdx.x == 0) {
                atomicOr(&status[0], 4u);
                scene_cnt[sc] = 0;
            }
            continue;
        }
        for (u32 i = threadIdx.x; i < n; i += kSceneBlock) {
            u32 k[6];
            if (!keys::box_keys(pos, quat, half, offset, b0 + i, k, sphere)) atomicOr(&status[0], 1u);
            klo[i] = make_uint4(k[0], k[1], k[2], k[3]);
            khi[i] = make_uint2(k[4], k[5]);
        }
        __syncthreads();
        const u32 words = scene_words(n);
        u32* M = mask + mask_off[sc];
        u32 total = 0;
        // a thread takes row r and row n-1-r, so every thread tests about n
        // pairs (row i tests the n-1-i boxes above it)
        for (u32 r = threadIdx.x; r < (n + 1) / 2; r += kSceneBlock) {
            for (u32 h = 0; h < 2; ++h) {
                const u32 i = h ? n - 1 - r : r;
                if (h && i == r) break;
                const uint4 al = klo[i];
                const uint2 ah = khi[i];
                u32* row = M + static_cast<std::size_t>(i) * words;
                u32 c = 0;
                for (u32 wd = 0; wd < words; ++wd) {
                    const u32 j0 = wd * 32u;
                    u32 bits = 0;
                    for (u32 j = max(j0, i + 1); j < min(j0 + 32u, n); ++j) {
                        const uint4 bl = klo[j];
                        const uint2 bh = khi[j];
                        // inclusive overlap on all three axes (bitwise: no branches)
                        const bool ov = (al.x <= bl.w) & (bl.x <= al.w) & (al.y <= bh.x) & (bl.y <= ah.x) &
                                        (al.z <= bh.y) & (bl.z <= ah.y);
                        bits |= static_cast<u32>(ov) << (j - j0);
                    }
                    row[wd] = bits;
                    c += __popc(bits);
                }
                row_cnt[b0 + i] = c;
                total += c;
            }
        }
        u32 sum;
        (void)block_exclusive_scan<kSceneBlock>(total, scratch, &sum);
        if (threadIdx.x == 0) scene_cnt[sc] = sum;
        __syncthreads();
    }
}

__global__ __launch_bounds__(kSceneBlock) void k_scene_write(u32 n_scenes, const u32* first, const u32* mask_off,
                                                             const u32* mask, const u32* row_cnt,
                                                             const u32* scene_base, u64* out, u32 cap, u32* status) {
    __shared__ u32 row[kSceneMax];
    __shared__ u32 scratch[kSceneBlock / kWave + 1];
    for (u32 sc = blockIdx.x; sc < n_scenes; sc += gridDim.x) {
        const u32 b0 = first[sc], n = first[sc + 1] - b0;
        if (n > kSceneMax) continue;
        u32 carry = 0;
        for (u32 t0 = 0; t0 < n; t0 += kSceneBlock) {
            const u32 i = t0 + threadIdx.x;
            const u32 v = i < n ? row_cnt[b0 + i] : 0u;
            u32 total;
            const u32 ex = block_exclusive_scan<kSceneBlock>(v, scratch, &total);
            if (i < n) row[i] = carry + ex;
            carry += total;
        }
        __syncthreads();
        const u32 base = scene_base[sc];
        if (threadIdx.x == 0 && base + carry > cap) atomicOr(&status[0], 2u);
        const u32 words = scene_words(n);
        const u32* M = mask + mask_off[sc];
        for (u32 i = threadIdx.x; i < n; i += kSceneBlock) {
            u32 o = base + row[i];
            const u32* rw = M + static_cast<std::size_t>(i) * words;
            for (u32 wd = (i + 1) / 32u; wd < words; ++wd) {
                u32 bits = rw[wd];
                while (bits) {
                    const u32 j = wd * 32u + static_cast<u32>(__ffs(bits)) - 1u;
                    bits &= bits - 1u;
                    if (o < cap) out[o] = (static_cast<u64>(b0 + i) << 32) | (b0 + j);
                    ++o;
                }
            }
        }
        __syncthreads();
    }
}

struct SceneCount {
    const u32* cnt;
    __device__ u32 operator()(u32 i) const { return cnt[i]; }
};

__global__ void k_clamp(u32* d_n, u32 cap) { *d_n = min(*d_n, cap); }

// ---- contact events (Scene::publish_contact_events, scene.cpp:9149-9259;
// touch transitions of update_contacts, 4640-4731)

__device__ inline u64 pair_key(u32 a, u32 b) {
    return a < b ? (static_cast<u64>(a) << 32) | b : (static_cast<u64>(b) << 32) | a;
}

// manager index of each touching pair slot
__global__ void k_slot_mgr(const u32* d_nm, u32 cap_m, const u32* mgr_slot, u32* slot_mgr) {
    const u32 n = min(*d_nm, cap_m);
    for (u32 m = blockIdx.x * blockDim.x + threadIdx.x; m < n; m += gridDim.x * blockDim.x) slot_mgr[mgr_slot[m]] = m;
}

struct EvArgs {
    u32 cap;                  // pair capacity
    const u32* d_nbb;         // current pair count
    const u64* prev;          // previous pair list
    const u32* d_prev;
    const u64* lost;          // pairs that left the list (sorted)
    const u32* d_lost;
    const u32* src;           // previous index of each current pair
    const u32* first;         // Gu first slot (higher) of each current pair
    const u32* second;
    const u32* counts_bb;
    const u32* counts_pl;
    const u32* touch_prev;    // touching flags aligned with the previous list
    const u32* touch_pl;      // plane touching flags of the previous step
    const u32* shape;         // shape (= body) id of each box slot
    u32 floor_shape;
    const u64* k1;            // compacted stream keys and sources
    const u32* s1;
    const u64* k2;
    const u32* s2;
    const u64* k3;
    const u32* s3;
    const u32* d_counts;      // n1, n2, n3, total
    const fyzz_dy_body* bodies;  // sleeping flags after the island pass
};

__device__ inline bool prev_touch_bb(const EvArgs& A, u32 i) {
    return A.src[i] != kNone && A.touch_prev[A.src[i]] != 0u;
}
__device__ inline u32 lost_prev_index(const EvArgs& A, u32 j) {
    const u32 np = min(*A.d_prev, A.cap);
    return lower_bound64(A.prev, np, A.lost[j]);
}

// A touch that persists inside a sleeping island is not reported
// (processPersistentContactEvents needs an active body, scene.cpp:
// 9199-9215); found and lost touches always are.
struct EvPred1 {
    EvArgs A;
    __device__ bool operator()(u32 i) const {
        const bool before = prev_touch_bb(A, i);
        if (A.counts_bb[i] == 0u) return before;
        return !(before && (A.bodies[A.first[i]].flags & 1u));
    }
};
struct EvWrite1 {
    EvArgs A;
    u64* k;
    u32* s;
    __device__ void operator()(u32 i, u32 o) const {
        k[o] = pair_key(A.shape[A.second[i]], A.shape[A.first[i]]);
        s[o] = i;
    }
};
struct EvPred2 {
    EvArgs A;
    __device__ bool operator()(u32 b) const {
        const bool before = A.touch_pl[b] != 0u;
        if (A.counts_pl[b] == 0u) return before;
        return !(before && (A.bodies[b].flags & 1u));
    }
};
struct EvWrite2 {
    EvArgs A;
    u64* k;
    u32* s;
    __device__ void operator()(u32 b, u32 o) const {
        k[o] = pair_key(A.floor_shape, A.shape[b]);
        s[o] = b;
    }
};
struct EvPred3 {
    EvArgs A;
    __device__ bool operator()(u32 j) const {
        const u32 np = min(*A.d_prev, A.cap);
        const u32 o = lost_prev_index(A, j);
        return o < np && A.prev[o] == A.lost[j] && A.touch_prev[o] != 0u;
    }
};
struct EvWrite3 {
    EvArgs A;
    u64* k;
    u32* s;
    __device__ void operator()(u32 j, u32 o) const {
        const u64 x = A.lost[j];
        k[o] = pair_key(A.shape[static_cast<u32>(x >> 32)], A.shape[static_cast<u32>(x & 0xffffffffu)]);
        s[o] = j;
    }
};

__global__ void k_ev_total(u32* d_counts) { d_counts[3] = d_counts[0] + d_counts[1] + d_counts[2]; }

// Gu-order position of every stream item (keys are distinct across the
// three streams), its event header and its point count.
__global__ void k_ev_merge(EvArgs A, fyzz_contact_event* ev, u32* ev_src, u32* ev_pts) {
    const u32 n1 = A.d_counts[0], n2 = A.d_counts[1], n3 = A.d_counts[2];
    const u32 total = n1 + n2 + n3;
    for (u32 t = blockIdx.x * blockDim.x + threadIdx.x; t < total; t += gridDim.x * blockDim.x) {
        u32 stream, i, pos, count;
        u64 key;
        bool touching, before;
        if (t < n1) {
            stream = 0;
            i = t;
            key = A.k1[i];
            pos = i + lower_bound64(A.k2, n2, key) + lower_bound64(A.k3, n3, key);
            const u32 p = A.s1[i];
            count = A.counts_bb[p];
            touching = count != 0u;
            before = prev_touch_bb(A, p);
        } else if (t < n1 + n2) {
            stream = 1;
            i = t - n1;
            key = A.k2[i];
            pos = i + lower_bound64(A.k1, n1, key) + lower_bound64(A.k3, n3, key);
            const u32 b = A.s2[i];
            count = A.counts_pl[b];
            touching = count != 0u;
            before = A.touch_pl[b] != 0u;
        } else {
            stream = 2;
            i = t - n1 - n2;
            key = A.k3[i];
            pos = i + lower_bound64(A.k1, n1, key) + lower_bound64(A.k2, n2, key);
            count = 0;
            touching = false;
            before = true;
        }
        const u32 a = static_cast<u32>(key >> 32), b = static_cast<u32>(key & 0xffffffffu);
        fyzz_contact_event e{};
        e.body0 = a;  // one shape per body: body ids are the shape ids, a < b
        e.body1 = b;
        e.kind = touching ? (before ? 1 : 0) : 2;
        e.contact_count = static_cast<uint8_t>(touching ? count : 0u);
        e.shape0 = a;
        e.shape1 = b;
        ev[pos] = e;
        ev_src[pos] = (stream << 30) | i;
        ev_pts[pos] = touching ? count : 0u;
    }
}

struct EvCount {
    const u32* ev_pts;
    __device__ u32 operator()(u32 i) const { return ev_pts[i]; }
};

// points of each event: published contacts, the outward normal flip
// (gu_order_reversed != body(a) > body(b), scene.cpp:9238-9240), the
// solver's impulses as double and their sum
__global__ void k_ev_points(EvArgs A, const u32* ev_src, const u32* ev_off, fyzz_contact_event* ev,
                            fyzz_contact_event_point* pts, const u32* slot_mgr, const fyzz_dy_manager* managers,
                            const float* impulses, u32 pt_cap) {
    const u32 total = A.d_counts[3];
    for (u32 e = blockIdx.x * blockDim.x + threadIdx.x; e < total; e += gridDim.x * blockDim.x) {
        const u32 off = ev_off[e];
        ev[e].point_offset = off;
        const u32 count = ev[e].contact_count;
        if (!count) {
            ev[e].normal_impulse = 0.0;
            continue;
        }
        const u32 stream = ev_src[e] >> 30, i = ev_src[e] & 0x3fffffffu;
        u32 slot;
        bool flip;
        if (stream == 0) {
            slot = A.s1[i];
            // Gu first the higher body (two dynamic ones, or a static one
            // created before its dynamic partner): reversed
            flip = A.shape[A.first[slot]] > A.shape[A.second[slot]];
        } else {
            const u32 b = A.s2[i];
            slot = A.cap + b;
            flip = A.floor_shape < A.shape[b];  // static first: reversed
        }
        const u32 m = slot_mgr[slot];
        const fyzz_dy_manager& M = managers[m];
        double sum = 0.0;
        for (u32 k = 0; k < count; ++k) {
            fyzz_contact_event_point q{};
            const fyzz_dy_contact& c = M.contacts[k];
            for (int d = 0; d < 3; ++d) {
                q.position[d] = c.position[d];
                q.normal[d] = flip ? -c.normal[d] : c.normal[d];
            }
            q.separation = c.separation;
            q.normal_impulse = static_cast<double>(impulses[4ull * m + k]);
            sum += q.normal_impulse;
            if (off + k < pt_cap) pts[off + k] = q;
        }
        ev[e].normal_impulse = sum;
    }
}

// this step's touch state for the next step's transitions
// Flag words of the body records as bytes (bit 0 asleep, bit 1 ready).
__global__ void k_flag_bytes(u32 n, const fyzz_dy_body* bodies, uint8_t* out) {
    for (u32 i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        out[i] = static_cast<uint8_t>(bodies[i].flags & 0xffu);
}

// The solver list in a given order: order[o] is the candidate index
// (box/box pair slot, or pair count + box for a plane manager) of the o-th
// manager.
__global__ void k_mgr_gather(const u32* order, u32 count, MgrWrite wr) {
    for (u32 o = blockIdx.x * blockDim.x + threadIdx.x; o < count; o += gridDim.x * blockDim.x) wr(order[o], o);
}

__global__ void k_touch(u32 cap, const u32* d_nbb, const u32* counts_bb, u32* touch_cur, u32 n, const u32* counts_pl,
                        u32* touch_pl) {
    const u32 nbb = min(*d_nbb, cap);
    for (u32 i = blockIdx.x * blockDim.x + threadIdx.x; i < nbb || i < n; i += gridDim.x * blockDim.x) {
        if (i < nbb) touch_cur[i] = counts_bb[i] != 0u ? 1u : 0u;
        if (counts_pl && i < n) touch_pl[i] = counts_pl[i] != 0u ? 1u : 0u;
    }
}

// running totals over steps (algorithmic byte accounting of the solve):
// [0] steps, [1] broadphase pairs, [2] managers, [3] contact rows,
// [4] friction rows, [5] partitions, [6] position+velocity iterations,
// [7] partition dependency depth, [8] broadphase found + lost pairs
__global__ void k_accumulate(const u32* d_pairs, u32 cap, const u32* d_nm, const u32* dy_ctr, const u32* d_found,
                             const u32* d_lost, u64* totals) {
    totals[0] += 1;
    totals[1] += min(*d_pairs, cap);
    totals[2] += *d_nm;
    totals[3] += dy_ctr[dy::kCtrContacts];
    totals[4] += dy_ctr[dy::kCtrFriction];
    totals[5] += dy_ctr[dy::kCtrParts];
    totals[6] += dy_ctr[dy::kCtrPos] + max(dy_ctr[dy::kCtrVel], 1u);
    totals[7] += dy_ctr[dy::kCtrRounds];
    if (d_found) totals[8] += *d_found;
    if (d_lost) totals[8] += *d_lost;
}

// ---- islands: sleep and wake (Scene::activate_published_islands,
// scene.cpp:6148-6262; Scene::deactivate_ready_accurate_islands,
// 6465-6536)
//
// Body flags: bit0 sleeping, bit1 ready for sleep (the last sleep check),
// bit3 activating (a wake queued for this step's island pass: an authored
// wake of a sleeping body, or a peer woken by an authored pose).
// Island membership is computed each step, not stored: connected
// components of the dynamic bodies over an edge set, by union-find (a
// larger root is hooked under a smaller one, so a component's root is its
// smallest slot whatever the order of the races).
//   speculative islands (wake): this step's broadphase pairs and the pairs
//     it lost (those are split only by this step's third pass, after the
//     wake: scene.cpp:10747-10793)
//   accurate islands (sleep): this step's touching box/box pairs. Only
//     ready bodies can sleep, so the union takes the touches between two
//     ready bodies; a ready body touching one that is not ready poisons
//     its component (that body is in the same island).
// The union kernels hook roots with agent-scope compare-and-swap and walk
// parents with plain loads (see unite); the kernels after them flatten the
// parents into labels (kernel boundaries make the union's results
// visible). Every kernel reads its gate from device
// counters, so the pass costs a few empty launches on steps where nothing
// can sleep or wake, and it can be captured in a graph.
constexpr u32 kSleeping = 1u, kReady = 2u, kActivating = 8u;
enum IslandCounter : u32 {
    kIslSleeping = 0,
    kIslActivating,
    kIslReadyAwake,
    kIslSleepingActivating,  // sleeping bodies carrying an activation (authored wakes)
    kIslMixed,               // speculative edges between a sleeping and an awake body
    kIslGiant,               // the largest speculative component after the link rounds (its root)
    kIslCounters
};

// Union-find on hardware whose L2 caches are not coherent across XCDs:
// plain loads may return an older parent, which is still an ancestor (a
// parent only ever moves up, to a smaller slot), so a walk from it still
// ends at a node that was a root once. Only hooking is atomic (agent-scope
// compare-and-swap, performed coherently): a hook on a node that is no
// longer a root fails and returns its current parent, where the walk goes
// on. Every failure moves strictly up, so the union ends; plain loads keep
// the shared upper nodes in each XCD's L2 instead of sending every read of
// a giant component's root to memory.
__device__ inline u32 find_plain(const u32* parent, u32 x) {
    u32 p = parent[x];
    while (p != x) {
        x = p;
        p = parent[x];
    }
    return x;
}
__device__ inline void unite(u32* parent, u32 a, u32 b) {
    // most edges join bodies already hanging from the same node
    if (parent[a] == parent[b]) return;
    a = find_plain(parent, a);
    b = find_plain(parent, b);
    while (a != b) {
        const u32 hi = a > b ? a : b, lo = a > b ? b : a;
        u32 expect = hi;
        if (__hip_atomic_compare_exchange_strong(parent + hi, &expect, lo, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                                 __HIP_MEMORY_SCOPE_AGENT))
            return;
        const u32 up = find_plain(parent, expect);  // hi had been hooked: go on from its parent
        if (hi == a)
            a = up;
        else
            b = up;
    }
}
// After the union kernel: the root by plain loads (parents only point down
// to smaller slots, so the walk ends).
__device__ inline u32 root_of(const u32* parent, u32 x) {
    u32 p = parent[x];
    while (p != x) {
        x = p;
        p = parent[x];
    }
    return x;
}

struct IslandArgs {
    u32 n, cap;
    fyzz_dy_body* bodies;
    const u64* cur;        // this step's broadphase pairs (lo, hi)
    const u32* d_cur;
    const u64* gone;       // pairs the speculative graph still holds: this step's lost pairs, or the
    const u32* d_gone;     // whole previous list where the broadphase lists no losses
    const u32* src;        // previous index of each current pair (kNone: found this step)
    const u32* first;      // Gu order of each current pair
    const u32* second;
    const u32* counts_bb;  // this step's contact counts (touching: non-zero)
    const u32* touch_prev; // the last step's touching flags, by previous index
    const u32* counts_pl;  // plane contact counts per box (null: no floor)
    const u32* touch_pl;   // the last step's plane touching flags
    u32* spec;             // union-find parents, then labels: speculative islands
    u32* acc;              // union-find parents, then labels: accurate islands (ready bodies)
    u32* root_wake;        // per speculative root: a member is activating
    u32* root_poison;      // per accurate root: a member touches a body that is not ready
    u32* root_awake;       // per accurate root: a member is awake
    u32* ctr;              // kIslCounters
    uint8_t* asleep;       // per body: sleeping, as the pass found it (a compact copy for edge scans)
};

__device__ inline void set_bits(fyzz_dy_body& b, u32 on, u32 off) {
    if (on) atomicOr(&b.flags, on);
    if (off) atomicAnd(&b.flags, ~off);
}

// Wake marks of this step's broadphase and touch events
// (activate_published_islands' first two loops): a pair found by the
// broadphase activates both bodies unless both sleep; a touch found by
// the narrowphase activates both and clears their readiness. Also resets
// the union-find parents, the root flags and the counters for the pass.
__global__ void k_isl_mark(IslandArgs A) {
    const u32 nbb = min(*A.d_cur, A.cap);
    const u32 stride = gridDim.x * blockDim.x;
    const u32 t0 = blockIdx.x * blockDim.x + threadIdx.x;
    if (t0 < kIslCounters) A.ctr[t0] = 0u;
    for (u32 i = t0; i < A.n; i += stride) {
        A.spec[i] = i;
        A.acc[i] = i;
        A.root_wake[i] = 0u;
        A.root_poison[i] = 0u;
        A.root_awake[i] = 0u;
    }
    for (u32 i = t0; i < nbb; i += stride) {
        // `first` is dynamic; `second` may be a static body (slot >= n),
        // which is no island node: a found pair with it activates nothing,
        // a found touch only the dynamic side
        const u32 a = A.first[i], b = A.second[i];
        const bool dyn_b = b < A.n;
        const u32 s = A.src[i];
        if (dyn_b && s == kNone && !((A.bodies[a].flags & A.bodies[b].flags) & kSleeping)) {
            set_bits(A.bodies[a], kActivating, 0u);
            set_bits(A.bodies[b], kActivating, 0u);
        }
        if (A.counts_bb[i] != 0u && !(s != kNone && A.touch_prev[s] != 0u)) {
            set_bits(A.bodies[a], kActivating, kReady);
            if (dyn_b) set_bits(A.bodies[b], kActivating, kReady);
        }
    }
    if (A.counts_pl)
        for (u32 i = t0; i < A.n; i += stride)
            if (A.counts_pl[i] != 0u && A.touch_pl[i] == 0u) set_bits(A.bodies[i], kActivating, kReady);
}

// Counters of the pass: sleeping, activating, awake-and-ready bodies.
__global__ void k_isl_count(IslandArgs A) {
    __shared__ u32 s[kIslCounters];
    if (threadIdx.x < kIslCounters) s[threadIdx.x] = 0u;
    __syncthreads();
    constexpr u32 kCounted = 4;  // the counters filled here
    u32 c[kCounted] = {0u, 0u, 0u, 0u};
    for (u32 i = blockIdx.x * blockDim.x + threadIdx.x; i < A.n; i += gridDim.x * blockDim.x) {
        const u32 f = A.bodies[i].flags;
        A.asleep[i] = static_cast<uint8_t>(f & kSleeping);
        c[kIslSleeping] += (f & kSleeping) ? 1u : 0u;
        c[kIslActivating] += (f & kActivating) ? 1u : 0u;
        c[kIslReadyAwake] += ((f & (kReady | kSleeping)) == kReady) ? 1u : 0u;
        c[kIslSleepingActivating] += ((f & (kActivating | kSleeping)) == (kActivating | kSleeping)) ? 1u : 0u;
    }
    for (u32 k = 0; k < kCounted; ++k)
        for (int d = kWave / 2; d > 0; d >>= 1) c[k] += __shfl_xor(c[k], d, kWave);
    if (lane_id() == 0)
        for (u32 k = 0; k < kCounted; ++k)
            if (c[k]) atomicAdd(&s[k], c[k]);
    __syncthreads();
    if (threadIdx.x < kCounted && s[threadIdx.x]) atomicAdd(&A.ctr[threadIdx.x], s[threadIdx.x]);
}

__device__ inline bool wake_pass(const IslandArgs& A) {
    return A.ctr[kIslSleeping] != 0u && A.ctr[kIslActivating] != 0u;
}
// The components matter only when a sleeping body could reach an
// activation: one carries an activation itself, or an edge joins a
// sleeping body to an awake one. Otherwise every sleeping body's island
// is made of sleeping bodies without activations, and nothing wakes.
__device__ inline bool spec_needed(const IslandArgs& A) {
    return wake_pass(A) && (A.ctr[kIslSleepingActivating] != 0u || A.ctr[kIslMixed] != 0u);
}

// Speculative edges: this step's pairs, then the pairs it lost.
__device__ inline u64 spec_edge(const IslandArgs& A, u32 i, u32 nbb) { return i < nbb ? A.cur[i] : A.gone[i - nbb]; }
__device__ inline u32 edge_lo(u64 x) { return static_cast<u32>(x >> 32); }
__device__ inline u32 edge_hi(u64 x) { return static_cast<u32>(x & 0xffffffffu); }
// The link edges of the speculative union's first round: each body's
// first two pairs as the lower slot (the list is sorted by lower slot).
__device__ inline bool link_edge(const IslandArgs& A, u32 i, u32 nbb) {
    if (i >= nbb) return false;
    const u32 lo = edge_lo(A.cur[i]);
    return i == 0 || edge_lo(A.cur[i - 1]) != lo || i == 1 || edge_lo(A.cur[i - 2]) != lo;
}

// Counts the edges between a sleeping and an awake body (sleep bits from
// the compact per-body copy k_isl_count wrote).
__global__ void k_isl_spec_mixed(IslandArgs A) {
    if (!wake_pass(A)) return;
    const u32 nbb = min(*A.d_cur, A.cap), ng = min(*A.d_gone, A.cap);
    u32 mixed = 0;
    for (u32 i = blockIdx.x * blockDim.x + threadIdx.x; i < nbb + ng; i += gridDim.x * blockDim.x) {
        const u64 x = spec_edge(A, i, nbb);
        if (edge_hi(x) >= A.n) continue;  // a static body joins no island
        mixed += A.asleep[edge_lo(x)] != A.asleep[edge_hi(x)] ? 1u : 0u;
    }
    for (int d = kWave / 2; d > 0; d >>= 1) mixed += __shfl_xor(mixed, d, kWave);
    if (lane_id() == 0 && mixed) atomicAdd(&A.ctr[kIslMixed], mixed);
}

// When the components are needed: links each body to its first two pairs
// (Afforest sampling: most bodies join one giant component through a
// small share of the edges).
__global__ void k_isl_spec_link(IslandArgs A) {
    if (!spec_needed(A)) return;
    const u32 nbb = min(*A.d_cur, A.cap);
    for (u32 i = blockIdx.x * blockDim.x + threadIdx.x; i < nbb; i += gridDim.x * blockDim.x)
        if (link_edge(A, i, nbb) && edge_hi(A.cur[i]) < A.n) unite(A.spec, edge_lo(A.cur[i]), edge_hi(A.cur[i]));
}

// Flattens the linked parents into labels.
__global__ void k_isl_spec_compress(IslandArgs A) {
    if (!spec_needed(A)) return;
    for (u32 i = blockIdx.x * blockDim.x + threadIdx.x; i < A.n; i += gridDim.x * blockDim.x)
        A.spec[i] = root_of(A.spec, i);
}

// The most frequent label among 256 sampled bodies (one workgroup): the
// giant component, whose internal edges the full union skips.
__global__ void k_isl_spec_giant(IslandArgs A) {
    if (!spec_needed(A)) return;
    constexpr u32 kSamples = 256;
    __shared__ u32 lab[kSamples];
    for (u32 k = threadIdx.x; k < kSamples; k += blockDim.x)
        lab[k] = A.n ? A.spec[static_cast<u32>((static_cast<u64>(k) * A.n) / kSamples)] : 0u;
    __syncthreads();
    // count, for each sampled label, how often it occurs; keep the best
    u32 best = 0u, best_count = 0u;
    for (u32 k = threadIdx.x; k < kSamples; k += blockDim.x) {
        u32 c = 0;
        for (u32 j = 0; j < kSamples; ++j) c += lab[j] == lab[k] ? 1u : 0u;
        if (c > best_count || (c == best_count && lab[k] < best)) {
            best = lab[k];
            best_count = c;
        }
    }
    __shared__ u64 s_best;
    if (threadIdx.x == 0) s_best = ~0ull;
    __syncthreads();
    // largest count first, then smallest label
    atomicMin(&s_best, (static_cast<u64>(kSamples - best_count) << 32) | best);
    __syncthreads();
    if (threadIdx.x == 0) A.ctr[kIslGiant] = static_cast<u32>(s_best & 0xffffffffu);
}

// The full speculative union: every edge not already inside the giant
// component (both labels equal the giant root: their bodies stay joined).
__global__ void k_isl_union_spec(IslandArgs A) {
    if (!spec_needed(A)) return;
    const u32 nbb = min(*A.d_cur, A.cap), ng = min(*A.d_gone, A.cap);
    const u32 giant = A.ctr[kIslGiant];
    for (u32 i = blockIdx.x * blockDim.x + threadIdx.x; i < nbb + ng; i += gridDim.x * blockDim.x) {
        if (link_edge(A, i, nbb)) continue;  // already united
        const u64 x = spec_edge(A, i, nbb);
        const u32 a = edge_lo(x), b = edge_hi(x);
        if (b >= A.n) continue;  // a static body joins no island
        if (A.spec[a] == giant && A.spec[b] == giant) continue;
        unite(A.spec, a, b);
    }
}

// Labels (flattened roots) and the roots of the speculative islands that
// hold an activating body.
__global__ void k_isl_wake_roots(IslandArgs A) {
    if (!spec_needed(A)) return;
    for (u32 i = blockIdx.x * blockDim.x + threadIdx.x; i < A.n; i += gridDim.x * blockDim.x) {
        const u32 r = root_of(A.spec, i);
        if (A.bodies[i].flags & kActivating) A.root_wake[r] = 1u;
        A.spec[i] = r;
    }
}

// Wakes every sleeping member of those islands (wake counter 0.4 s at
// least, readiness and sleep filter cleared) and consumes the activation
// marks.
__global__ void k_isl_wake(IslandArgs A) {
    if (A.ctr[kIslActivating] == 0u) return;
    const bool wake = spec_needed(A);
    for (u32 i = blockIdx.x * blockDim.x + threadIdx.x; i < A.n; i += gridDim.x * blockDim.x) {
        fyzz_dy_body& b = A.bodies[i];
        u32 f = b.flags & ~kActivating;
        if (wake && (f & kSleeping) && A.root_wake[root_of(A.spec, i)]) {
            f &= ~(kSleeping | kReady);
            b.wake_counter = b.wake_counter > 0.4f ? b.wake_counter : 0.4f;
            for (int k = 0; k < 3; ++k) b.sleep_linear_acc[k] = b.sleep_angular_acc[k] = 0.0f;
        }
        b.flags = f;
    }
}

// Accurate union over touches between two ready bodies; a touch between a
// ready body and one that is not poisons the ready one.
__global__ void k_isl_union_acc(IslandArgs A) {
    if (A.ctr[kIslReadyAwake] == 0u) return;
    const u32 nbb = min(*A.d_cur, A.cap);
    for (u32 i = blockIdx.x * blockDim.x + threadIdx.x; i < nbb; i += gridDim.x * blockDim.x) {
        if (A.counts_bb[i] == 0u) continue;
        const u32 a = A.first[i], b = A.second[i];
        if (b >= A.n) continue;  // touching a static body does not hold an island awake
        const bool ra = A.bodies[a].flags & kReady, rb = A.bodies[b].flags & kReady;
        if (ra && rb)
            unite(A.acc, a, b);
        else if (ra || rb)
            A.root_poison[ra ? a : b] = 1u;  // per body for now; folded into its root below
    }
}

// Per accurate component of ready bodies: poisoned, awake.
__global__ void k_isl_sleep_roots(IslandArgs A) {
    if (A.ctr[kIslReadyAwake] == 0u) return;
    for (u32 i = blockIdx.x * blockDim.x + threadIdx.x; i < A.n; i += gridDim.x * blockDim.x) {
        const u32 f = A.bodies[i].flags;
        if (!(f & kReady)) continue;
        const u32 r = root_of(A.acc, i);
        if (r != i && A.root_poison[i]) A.root_poison[r] = 1u;
        if (!(f & kSleeping)) A.root_awake[r] = 1u;
    }
}

// A component of ready bodies that is a whole island (unpoisoned), with
// one awake, goes to sleep: sleeping, wake counter 0, velocities and
// sleep filter cleared.
__global__ void k_isl_sleep(IslandArgs A) {
    if (A.ctr[kIslReadyAwake] == 0u) return;
    for (u32 i = blockIdx.x * blockDim.x + threadIdx.x; i < A.n; i += gridDim.x * blockDim.x) {
        fyzz_dy_body& b = A.bodies[i];
        const u32 f = b.flags;
        if ((f & (kReady | kSleeping)) != kReady) continue;
        const u32 r = root_of(A.acc, i);
        if (A.root_poison[r] || !A.root_awake[r]) continue;
        b.flags = f | kSleeping;
        b.wake_counter = 0.0f;
        for (int k = 0; k < 3; ++k)
            b.linear_velocity[k] = b.angular_velocity[k] = b.sleep_linear_acc[k] = b.sleep_angular_acc[k] = 0.0f;
    }
}

// ---- authored edits after the first step

// Scene::set_pose's wake of the body itself (scene.cpp:637-647) and
// set_velocity / set_angular_velocity's (730-781): the wake counter is
// raised to 0.4 s and the sleep flags cleared; a sleeping body also drops
// its sleep filter.
// A sleeping body's wake is also queued for the next step's island pass
// (queue_node_activation: its island wakes with it); `queue` queues it
// regardless (a peer woken by an authored pose).
__host__ __device__ inline void wake_record(fyzz_dy_body& b, bool queue = false) {
    if (b.flags & 1u) {
        for (int k = 0; k < 3; ++k) b.sleep_linear_acc[k] = b.sleep_angular_acc[k] = 0.0f;
        queue = true;
    }
    b.wake_counter = b.wake_counter > 0.4f ? b.wake_counter : 0.4f;
    b.flags = (b.flags & ~3u) | (queue ? 8u : 0u);
}

// Scene::set_pose's invalidations (scene.cpp:651-700) for the box in
// `slot`: every current broadphase pair holding it and its plane pair
// lose their PCM cache (ShapeInteraction::resetManagerCachedState). When
// the pose changed, their friction history is erased as well and the
// other body of every touching pair (sleeping islands' included) that
// sleeps or is ready to sleep is woken and queued for the next island
// pass. `pairs` is the list the next step starts from; `cache_bb`,
// `stamp` and `counts_bb` (touching: non-zero) are aligned with it.
__global__ void k_author_pose(u32 slot, int changed, u32 cap, const u64* pairs, const u32* d_pairs,
                              CacheRec* cache_bb, CacheRec* cache_pl, int has_plane, u32* stamp,
                              const u32* counts_bb, fyzz_dy_body* bodies, u32 n_dyn) {
    const u32 np = d_pairs ? min(*d_pairs, cap) : 0u;
    const u32 stride = gridDim.x * blockDim.x;
    const u32 t0 = blockIdx.x * blockDim.x + threadIdx.x;
    for (u32 i = t0; i < np; i += stride) {
        const u64 x = pairs[i];
        if (static_cast<u32>(x >> 32) != slot && static_cast<u32>(x & 0xffffffffu) != slot) continue;
        for (u32 k = 0; k < kCacheWords; ++k) cache_bb[i].w[k] = 0u;
        if (changed) stamp[i] = kNone;
    }
    if (t0 == 0 && has_plane) {
        for (u32 k = 0; k < kCacheWords; ++k) cache_pl[slot].w[k] = 0u;
        if (changed) stamp[cap + slot] = kNone;
    }
    if (!changed) return;
    // the other body of every touching pair (sleeping islands' included)
    for (u32 i = t0; i < np; i += stride) {
        if (counts_bb[i] == 0u) continue;
        const u64 x = pairs[i];
        const u32 lo = static_cast<u32>(x >> 32), hi = static_cast<u32>(x & 0xffffffffu);
        if (lo != slot && hi != slot) continue;
        const u32 peer = lo == slot ? hi : lo;
        if (peer >= n_dyn) continue;  // a static peer has nothing to wake
        fyzz_dy_body& b = bodies[peer];
        // every writer of one peer stores the same values
        if (b.flags & 3u) wake_record(b, true);
    }
}

// Scene::add_dynamic_box mass properties (scene.cpp:408-475): mass in
// double, the unit-density inertia tensor scaled by density in float.
void box_mass(const float h[3], float density, float& inv_mass, float inv_inertia[3]) {
    const double mass = 8.0 * double(h[0]) * double(h[1]) * double(h[2]) * double(density);
    inv_mass = static_cast<float>(1.0 / mass);
    const float unit_mass = (h[0] * h[1] * h[2]) * 8.0f;
    const float ex = h[0] * h[0], ey = h[1] * h[1], ez = h[2] * h[2];
    const float unit_scale = (unit_mass * 1.0f) / 3.0f;
    const float d = density;
    const float ix = ((ey + ez) * unit_scale) * d, iy = ((ex + ez) * unit_scale) * d,
                iz = ((ex + ey) * unit_scale) * d;
    inv_inertia[0] = 1.0f / ix;
    inv_inertia[1] = 1.0f / iy;
    inv_inertia[2] = 1.0f / iz;
}

}  // namespace fyzz_hip::world

using namespace fyzz_hip;

struct fyzz_hip_world {
    fyzz_hip_context* ctx = nullptr;
    u32 max_bodies = 0;
    // authoring (host) until the first step
    std::vector<fyzz_dy_body> host_bodies;
    std::vector<float> host_half;
    std::vector<uint8_t> host_sphere;  // per slot: a sphere (radius in its three half extents)
    // static bodies other than the floor (slot_of_id = -2 - index): SAP
    // slots n.. after the dynamic ones, solver statics 1.. after the floor
    struct StaticBody {
        float half[3], pos[3], quat[4];
        uint8_t sphere;
        float static_friction, dynamic_friction, restitution;
    };
    std::vector<StaticBody> host_statics;
    bool has_spheres = false;
    std::vector<int> slot_of_id;  // body id -> box slot, -1 for the floor
    bool has_floor = false;
    u32 floor_id = 0;
    fyzz_dy_static floor{};
    bool uploaded = false;
    int req_partition_blocks = 0, req_pgs_blocks = 0;  // launch shapes set before the Dy workspace exists
    u32 n = 0;
    u32 pair_cap = 0, mgr_cap = 0;
    u32 step_index = 0;
    int ccur = 0, hcur = 0;
    fyzz_hip_sap* sap = nullptr;
    fyzz_hip_dy* dy = nullptr;
    DevBuf<fyzz_dy_body> bodies;
    DevBuf<fyzz_dy_static> statics;
    DevBuf<float> pos, quat, half, offset, contacts_bb, contacts_pl, impulses;
    DevBuf<u32> first, second, src, counted, counts_bb, flags_bb, counts_pl, flags_pl, pcm_scratch, tiles, d_total,
        d_nm, mgr_slot, parts, stamp[2];
    DevBuf<world::CacheRec> cache_bb[2], cache_pl;
    DevBuf<fyzz_dy_manager> managers;
    DevBuf<fyzz_dy_patch> patches, hist[2];
    DevBuf<fyzz_friction_anchors> anchors;  // friction write-back, filled on demand
    // island pass: union-find parents (speculative, accurate), per-root flags, counters
    DevBuf<u32> isl_spec, isl_acc, isl_root_wake, isl_root_poison, isl_root_awake, isl_ctr;
    DevBuf<uint8_t> isl_asleep;
    DevBuf<u64> totals;
    DevBuf<u32> d_step;
    // graph stepping: one graph per parity of the double-buffer toggles
    hipStream_t gstream = nullptr;
    hipGraphExec_t gexec[2] = {nullptr, nullptr};
    float gdt[2] = {0.0f, 0.0f};
    hipEvent_t gev_in = nullptr, gev_out = nullptr;
    // independent scenes (per-scene broadphase instead of one SAP)
    std::vector<u32> scene_counts;
    u32 n_scenes = 0;
    int scur = 0;
    DevBuf<u32> scene_first, scene_cnt, scene_base, scene_n, scene_status;
    DevBuf<u32> scene_mask_off, scene_mask, scene_rows;  // per-scene pair bitmasks (rows of partners) and row counts
    DevBuf<u64> scene_pairs[2];
    // contact events
    bool notify = false;
    DevBuf<u32> shape, touch_bb[2], touch_pl, slot_mgr, s1, s2, s3, ev_counts, ev_src, ev_pts, ev_off, d_npts;
    DevBuf<u64> k1, k2, k3;
    DevBuf<fyzz_contact_event> events;
    DevBuf<fyzz_contact_event_point> points;
    u32 ev_cap = 0, pt_cap = 0;
    // the native scene's solver order (host island authority), chosen
    // before the first step
    bool native_order = false;
    fyzz_hip_islands* order_auth = nullptr;
    std::vector<u32> id_of_slot;
    DevBuf<u64> native_keys, native_keys2, native_keys3, native_found, native_order_keys;  // compacted pairs;
                                                                                         // the order as slot keys
    DevBuf<u32> native_planes, native_planes2, native_counts, native_n;
    std::vector<uint8_t> native_sleeping;
    DevBuf<u32> native_list;
    DevBuf<uint8_t> flag_bytes;
    DevBuf<uint8_t> sphere;  // per slot, when the world holds spheres
    // a static body created before a dynamic one: slot order is not id
    // order, so the event streams built in slot order are sorted by id
    bool ids_out_of_slot_order = false;
    DevBuf<u64> ev_tmp_keys;
    DevBuf<u32> ev_tmp_vals, ev_sort_scratch;

    ~fyzz_hip_world() {
        if (order_auth) fyzz_hip_islands_destroy(order_auth);
        for (auto& g : gexec)
            if (g) (void)hipGraphExecDestroy(g);
        if (gev_in) (void)hipEventDestroy(gev_in);
        if (gev_out) (void)hipEventDestroy(gev_out);
        if (gstream) (void)hipStreamDestroy(gstream);
        if (dy) fyzz_hip_dy_destroy(dy);
        if (sap) fyzz_hip_sap_destroy(sap);
    }
    // broadphase slot of a dynamic or static (non-floor) body id
    u32 slot(u32 id) const {
        const int s = slot_of_id[id];
        return s >= 0 ? static_cast<u32>(s) : n_dynamic_slots() + static_cast<u32>(-2 - s);
    }
    u32 n_dynamic_slots() const { return static_cast<u32>(host_bodies.size()); }
    fyzz_dy_body& body(u32 id) {
        FYZZ_REQUIRE(id < slot_of_id.size() && slot_of_id[id] >= 0, kInvalidArgument, "world: not a dynamic body id");
        return host_bodies[static_cast<u32>(slot_of_id[id])];
    }
    void require_authoring() const {
        FYZZ_REQUIRE(!uploaded, kLogicError, "world: authoring calls must precede the first step");
    }
};

namespace {

fyzz_dy_body default_body() {
    fyzz_dy_body b{};
    b.orientation[3] = 1.0f;
    b.wake_counter = 0.4f;
    b.flags = 4u;
    b.angular_damping = 0.05f;
    b.max_linear_velocity = 1.0e16f;
    b.max_angular_velocity = 100.0f;
    b.static_friction = 0.5f;
    b.dynamic_friction = 0.5f;
    b.restitution = 0.2f;
    b.iterations = 4u | (1u << 8);
    return b;
}

void reserve_events(fyzz_hip_world* w);

// Captured step graphs hold launch shapes and whether events publish.
void drop_graphs(fyzz_hip_world* w) {
    for (auto& g : w->gexec)
        if (g) {
            FYZZ_HIP_CHECK(hipGraphExecDestroy(g));
            g = nullptr;
        }
}

// The solver's statics: the floor, then the static bodies (solver static
// 1 + k is static body k; a manager names it as -2 - k).
void upload_statics(fyzz_hip_world* w) {
    std::vector<fyzz_dy_static> rec(1 + w->host_statics.size());
    rec[0] = w->floor;
    for (std::size_t k = 0; k < w->host_statics.size(); ++k) {
        const auto& S = w->host_statics[k];
        fyzz_dy_static& r = rec[1 + k];
        for (int a = 0; a < 3; ++a) r.position[a] = S.pos[a];
        for (int a = 0; a < 4; ++a) r.orientation[a] = S.quat[a];
        r.static_friction = S.static_friction;
        r.dynamic_friction = S.dynamic_friction;
        r.restitution = S.restitution;
    }
    hipStream_t st = w->ctx->stream;
    FYZZ_HIP_CHECK(hipMemcpyAsync(w->statics.p, rec.data(), sizeof(fyzz_dy_static) * rec.size(), hipMemcpyHostToDevice,
                                  st));
    FYZZ_HIP_CHECK(hipStreamSynchronize(st));
}

void upload(fyzz_hip_world* w) {
    const u32 n = static_cast<u32>(w->host_bodies.size());
    const u32 ns = static_cast<u32>(w->host_statics.size());
    const u32 slots = n + ns;  // broadphase slots: the dynamic bodies, then the statics
    w->n = n;
    w->pair_cap = std::max<u32>(8 * slots, 4096);
    w->mgr_cap = w->pair_cap + n;
    const std::size_t nn = std::max<u32>(n, 1), ss = std::max<u32>(slots, 1);
    hipStream_t st = w->ctx->stream;
    w->bodies.reserve(nn);
    w->statics.reserve(1 + ns);
    w->pos.reserve(3 * ss);
    w->quat.reserve(4 * ss);
    w->half.reserve(3 * ss);
    w->offset.reserve(ss);
    w->counted.reserve(nn);
    w->counts_pl.reserve(nn);
    w->flags_pl.reserve(nn);
    w->contacts_pl.reserve(28 * nn);
    w->cache_pl.reserve(nn);
    for (auto* b : {&w->first, &w->second, &w->src, &w->counts_bb, &w->flags_bb}) b->reserve(w->pair_cap);
    w->pcm_scratch.reserve(w->pair_cap + 1);
    w->contacts_bb.reserve(28ull * w->pair_cap);
    w->cache_bb[0].reserve(w->pair_cap);
    w->cache_bb[1].reserve(w->pair_cap);
    // sized for the largest stream that uses it (manager list, contact
    // events): never grown after the first step, while kernels may hold it
    w->tiles.reserve(sel::tiles_for(std::max(w->mgr_cap, 2 * w->pair_cap + n)) + 1);
    w->d_total.reserve(1);
    w->d_nm.reserve(1);
    w->mgr_slot.reserve(w->mgr_cap);
    w->shape.reserve(ss);
    w->touch_pl.reserve(nn);
    FYZZ_HIP_CHECK(hipMemsetAsync(w->touch_pl.p, 0, sizeof(u32) * nn, st));
    for (int k = 0; k < 2; ++k) {
        w->touch_bb[k].reserve(w->pair_cap);
        FYZZ_HIP_CHECK(hipMemsetAsync(w->touch_bb[k].p, 0, sizeof(u32) * w->pair_cap, st));
    }
    {
        std::vector<u32> shape_of_slot(slots);
        for (std::size_t id = 0; id < w->slot_of_id.size(); ++id)
            if (w->slot_of_id[id] >= 0 || w->slot_of_id[id] <= -2) shape_of_slot[w->slot(static_cast<u32>(id))] = static_cast<u32>(id);
        if (slots)
            FYZZ_HIP_CHECK(hipMemcpyAsync(w->shape.p, shape_of_slot.data(), sizeof(u32) * slots,
                                          hipMemcpyHostToDevice, st));
        FYZZ_HIP_CHECK(hipStreamSynchronize(st));
    }
    for (auto* b : {&w->isl_spec, &w->isl_acc, &w->isl_root_wake, &w->isl_root_poison, &w->isl_root_awake})
        b->reserve(nn);
    w->isl_ctr.reserve(world::kIslCounters);
    w->isl_asleep.reserve(nn);
    w->totals.reserve(world::kTotals);
    FYZZ_HIP_CHECK(hipMemsetAsync(w->totals.p, 0, world::kTotals * sizeof(u64), st));
    w->d_step.reserve(1);
    FYZZ_HIP_CHECK(hipMemsetAsync(w->d_step.p, 0, sizeof(u32), st));
    w->parts.reserve(w->mgr_cap);
    w->impulses.reserve(4ull * w->mgr_cap);
    w->managers.reserve(w->mgr_cap);
    w->patches.reserve(w->mgr_cap);
    for (int k = 0; k < 2; ++k) {
        w->hist[k].reserve(w->mgr_cap);
        w->stamp[k].reserve(w->mgr_cap);
        FYZZ_HIP_CHECK(hipMemsetAsync(w->stamp[k].p, 0xff, sizeof(u32) * w->mgr_cap, st));
    }
    FYZZ_HIP_CHECK(hipMemsetAsync(w->cache_pl.p, 0, sizeof(world::CacheRec) * nn, st));
    if (n) {
        FYZZ_HIP_CHECK(hipMemcpyAsync(w->bodies.p, w->host_bodies.data(), sizeof(fyzz_dy_body) * n,
                                      hipMemcpyHostToDevice, st));
        FYZZ_HIP_CHECK(hipMemcpyAsync(w->half.p, w->host_half.data(), sizeof(float) * 3 * n, hipMemcpyHostToDevice,
                                      st));
    }
    if (slots) {
        // statics: their poses and extents sit after the dynamic slots and stay
        std::vector<float> spos(3 * ns), squat(4 * ns), shalf(3 * ns);
        std::vector<uint8_t> sph(w->host_sphere);
        for (u32 k = 0; k < ns; ++k) {
            const auto& S = w->host_statics[k];
            for (int a = 0; a < 3; ++a) spos[3 * k + a] = S.pos[a], shalf[3 * k + a] = S.half[a];
            for (int a = 0; a < 4; ++a) squat[4 * k + a] = S.quat[a];
            sph.push_back(S.sphere);
        }
        if (ns) {
            FYZZ_HIP_CHECK(hipMemcpyAsync(w->pos.p + 3 * n, spos.data(), sizeof(float) * 3 * ns, hipMemcpyHostToDevice, st));
            FYZZ_HIP_CHECK(hipMemcpyAsync(w->quat.p + 4 * n, squat.data(), sizeof(float) * 4 * ns, hipMemcpyHostToDevice, st));
            FYZZ_HIP_CHECK(hipMemcpyAsync(w->half.p + 3 * n, shalf.data(), sizeof(float) * 3 * ns, hipMemcpyHostToDevice, st));
        }
        std::vector<float> off(slots, world::kContactOffset);
        FYZZ_HIP_CHECK(hipMemcpyAsync(w->offset.p, off.data(), sizeof(float) * slots, hipMemcpyHostToDevice, st));
        if (w->has_spheres) {
            w->sphere.reserve(slots);
            FYZZ_HIP_CHECK(hipMemcpyAsync(w->sphere.p, sph.data(), slots, hipMemcpyHostToDevice, st));
        }
        FYZZ_HIP_CHECK(hipStreamSynchronize(st));
    }
    upload_statics(w);
    int rc = 0;
    if (w->n_scenes) {
        u64 total = 0;
        std::vector<u32> first(w->n_scenes + 1);
        for (u32 k = 0; k < w->n_scenes; ++k) {
            first[k] = static_cast<u32>(total);
            FYZZ_REQUIRE(w->scene_counts[k] <= world::kSceneMax, kInvalidArgument,
                         "world: a scene holds more boxes than the per-scene broadphase takes (1024)");
            total += w->scene_counts[k];
        }
        first[w->n_scenes] = static_cast<u32>(total);
        FYZZ_REQUIRE(total == n, kInvalidArgument, "world: scene box counts do not add up to the boxes authored");
        // each scene's partner bitmasks: one row of ceil(n/32) words per box
        std::vector<u32> mask_off(w->n_scenes + 1);
        u64 words = 0;
        for (u32 k = 0; k < w->n_scenes; ++k) {
            mask_off[k] = static_cast<u32>(words);
            words += static_cast<u64>(w->scene_counts[k]) * ((w->scene_counts[k] + 31u) / 32u);
            FYZZ_REQUIRE(words < (1ull << 32), kInvalidArgument, "world: scenes too large for the per-scene broadphase");
        }
        mask_off[w->n_scenes] = static_cast<u32>(words);
        w->scene_mask_off.reserve(w->n_scenes + 1);
        w->scene_mask.reserve(std::max<u64>(words, 1));
        w->scene_rows.reserve(std::max<u32>(n, 1));
        FYZZ_HIP_CHECK(hipMemcpyAsync(w->scene_mask_off.p, mask_off.data(), sizeof(u32) * mask_off.size(),
                                      hipMemcpyHostToDevice, st));
        w->scene_first.reserve(w->n_scenes + 1);
        w->scene_cnt.reserve(w->n_scenes);
        w->scene_base.reserve(w->n_scenes);
        w->scene_n.reserve(4);
        w->scene_status.reserve(4);
        w->scene_pairs[0].reserve(w->pair_cap);
        w->scene_pairs[1].reserve(w->pair_cap);
        FYZZ_HIP_CHECK(hipMemcpyAsync(w->scene_first.p, first.data(), sizeof(u32) * first.size(),
                                      hipMemcpyHostToDevice, st));
        FYZZ_HIP_CHECK(hipMemsetAsync(w->scene_n.p, 0, 4 * sizeof(u32), st));
        FYZZ_HIP_CHECK(hipMemsetAsync(w->scene_status.p, 0, 4 * sizeof(u32), st));
        FYZZ_HIP_CHECK(hipStreamSynchronize(st));
    } else {
        std::vector<u32> ids(slots);
        std::vector<uint8_t> dyn(slots, 0);
        for (u32 i = 0; i < slots; ++i) ids[i] = i, dyn[i] = i < n ? 1 : 0;
        rc = fyzz_hip_sap_create(w->ctx, slots, ids.data(), dyn.data(), w->pair_cap, w->pair_cap, &w->sap);
        if (rc) throw Error(rc, w->ctx->last_error);
        if (w->has_spheres) sap::set_spheres(w->sap, w->sphere.p);
    }
    rc = fyzz_hip_dy_create(w->ctx, n, w->mgr_cap, &w->dy);
    if (rc) throw Error(rc, w->ctx->last_error);
    if (w->req_partition_blocks) dy::set_coop_blocks(w->dy, w->req_partition_blocks, w->req_pgs_blocks);
    {
        // a static id below a dynamic one breaks id order = slot order
        int max_dyn = -1, min_static = 0x7fffffff;
        for (std::size_t id = 0; id < w->slot_of_id.size(); ++id) {
            if (w->slot_of_id[id] >= 0) max_dyn = static_cast<int>(id);
            else if (w->slot_of_id[id] <= -2) min_static = std::min(min_static, static_cast<int>(id));
        }
        w->ids_out_of_slot_order = min_static < max_dyn;
    }
    if (w->notify) reserve_events(w);
    if (w->native_order) {
        const u32 ids = static_cast<u32>(w->slot_of_id.size());
        std::vector<uint8_t> dynamic(ids, 0);
        w->id_of_slot.assign(slots, 0);
        for (u32 i = 0; i < ids; ++i) {
            if (w->slot_of_id[i] >= 0) dynamic[i] = 1;
            if (w->slot_of_id[i] >= 0 || w->slot_of_id[i] <= -2) w->id_of_slot[w->slot(i)] = i;
        }
        FYZZ_REQUIRE(ids > 0 && fyzz_hip_islands_create(ids, dynamic.data(), &w->order_auth) == 0, kOutOfMemory,
                     "world: cannot create the island authority");
        w->native_list.reserve(w->mgr_cap);
        w->flag_bytes.reserve(nn);
        w->native_keys.reserve(w->pair_cap);
        w->native_keys2.reserve(w->pair_cap);
        w->native_keys3.reserve(w->pair_cap);
        w->native_found.reserve(w->pair_cap);
        w->native_order_keys.reserve(w->mgr_cap);
        w->native_planes.reserve(nn);
        w->native_planes2.reserve(nn);
        w->native_counts.reserve(8);
        w->native_n.reserve(1);
    }
    FYZZ_HIP_CHECK(hipStreamSynchronize(st));
    w->uploaded = true;
}

sap::Lists sap_broadphase(fyzz_hip_world* w) {
    const int rc = fyzz_hip_sap_advance(w->sap, w->pos.p, w->quat.p, w->half.p, w->offset.p, w->n, nullptr, 0);
    if (rc) throw Error(rc, w->ctx->last_error);
    return sap::lists(w->sap);
}

sap::Lists scene_broadphase(fyzz_hip_world* w) {
    using namespace fyzz_hip::world;
    hipStream_t st = w->ctx->stream;
    const int cur = w->scur, nxt = cur ^ 1;
    const u32 S = w->n_scenes, cap = w->pair_cap;
    ProfScope ps(&w->ctx->prof, "world.scene_pairs", st);
    const u32 grid = std::min<u32>(S, 4096u);
    hipLaunchKernelGGL(k_scene_masks, dim3(grid), dim3(kSceneBlock), 0, st, S, w->scene_first.p,
                       w->scene_mask_off.p, w->pos.p, w->quat.p, w->half.p, w->offset.p, w->scene_mask.p,
                       w->scene_rows.p, w->scene_cnt.p, w->scene_status.p, w->has_spheres ? w->sphere.p : nullptr);
    hipLaunchKernelGGL(k_fill, dim3(1), dim3(64), 0, st, w->scene_n.p + 2, 1u, S);
    sel::exclusive_scan(st, S, w->scene_n.p + 2, SceneCount{w->scene_cnt.p}, w->scene_base.p, w->tiles.p,
                        w->scene_n.p + nxt);
    hipLaunchKernelGGL(k_scene_write, dim3(grid), dim3(kSceneBlock), 0, st, S, w->scene_first.p,
                       w->scene_mask_off.p, w->scene_mask.p, w->scene_rows.p, w->scene_base.p,
                       w->scene_pairs[nxt].p, cap, w->scene_status.p);
    hipLaunchKernelGGL(k_clamp, dim3(1), dim3(1), 0, st, w->scene_n.p + nxt, cap);
    FYZZ_HIP_CHECK(hipGetLastError());
    w->scur = nxt;
    return {w->scene_pairs[nxt].p, w->scene_n.p + nxt, w->scene_pairs[cur].p, w->scene_n.p + cur, cap,
            w->scene_status.p, w->scene_status.p + 1, nullptr, nullptr, nullptr};
}

// The contact-event buffers, sized from the world's capacities once, before
// any step that publishes (never inside a captured step).
void reserve_events(fyzz_hip_world* w) {
    if (w->ev_cap) return;
    const u32 cap = w->pair_cap, n = w->n;
    w->ev_cap = 2 * cap + n;
    w->pt_cap = 4 * (cap + n);
    w->slot_mgr.reserve(w->mgr_cap);
    w->s1.reserve(cap);
    w->k1.reserve(cap);
    w->s2.reserve(std::max<u32>(n, 1));
    w->k2.reserve(std::max<u32>(n, 1));
    w->s3.reserve(cap);
    w->k3.reserve(cap);
    if (w->ids_out_of_slot_order) {
        w->ev_tmp_keys.reserve(cap);
        w->ev_tmp_vals.reserve(cap);
        w->ev_sort_scratch.reserve(sap::sort_keys_scratch_words(cap));
    }
    w->ev_counts.reserve(4);
    w->ev_src.reserve(w->ev_cap);
    w->ev_pts.reserve(w->ev_cap);
    w->ev_off.reserve(w->ev_cap);
    w->d_npts.reserve(1);
    w->events.reserve(w->ev_cap);
    w->points.reserve(w->pt_cap);
}

void publish_events(fyzz_hip_world* w, const sap::Lists& L, int c_prev, int c_cur) {
    using namespace fyzz_hip::world;
    hipStream_t st = w->ctx->stream;
    const u32 cap = w->pair_cap, n = w->n;
    reserve_events(w);
    ProfScope ps(&w->ctx->prof, "world.events", st);
    EvArgs A{cap,          L.d_cur,         L.prev,         L.d_prev,        L.lost,       L.d_lost,
             w->src.p,     w->first.p,      w->second.p,    w->counts_bb.p,  w->counts_pl.p,
             w->touch_bb[c_prev].p, w->touch_pl.p, w->shape.p, w->floor_id,
             w->k1.p,      w->s1.p,         w->k2.p,        w->s2.p,         w->k3.p,      w->s3.p,
             w->ev_counts.p, w->bodies.p};
    hipLaunchKernelGGL(k_slot_mgr, dim3(std::min<u32>(grid_for(w->mgr_cap, 256), 2048u)), dim3(256), 0, st,
                       w->d_nm.p, w->mgr_cap, w->mgr_slot.p, w->slot_mgr.p);
    sel::select(st, cap, L.d_cur, EvPred1{A}, EvWrite1{A, w->k1.p, w->s1.p}, w->tiles.p, w->ev_counts.p);
    if (w->has_floor) {
        hipLaunchKernelGGL(k_fill, dim3(1), dim3(64), 0, st, w->d_npts.p, 1u, n);
        sel::select(st, n, w->d_npts.p, EvPred2{A}, EvWrite2{A, w->k2.p, w->s2.p}, w->tiles.p, w->ev_counts.p + 1);
    } else {
        hipLaunchKernelGGL(k_fill, dim3(1), dim3(64), 0, st, w->ev_counts.p + 1, 1u, 0u);
    }
    sel::select(st, cap, L.d_lost, EvPred3{A}, EvWrite3{A, w->k3.p, w->s3.p}, w->tiles.p, w->ev_counts.p + 2);
    if (w->ids_out_of_slot_order) {
        // streams 1 and 3 were compacted in slot order; the merge needs id order
        const int bits = w->slot_of_id.size() <= (1u << 16) ? 48 : 64;
        sap::sort_keys(st, w->k1.p, w->s1.p, w->ev_tmp_keys.p, w->ev_tmp_vals.p, cap, w->ev_counts.p, bits,
                       w->ev_sort_scratch.p);
        sap::sort_keys(st, w->k3.p, w->s3.p, w->ev_tmp_keys.p, w->ev_tmp_vals.p, cap, w->ev_counts.p + 2, bits,
                       w->ev_sort_scratch.p);
    }
    hipLaunchKernelGGL(k_ev_total, dim3(1), dim3(1), 0, st, w->ev_counts.p);
    const u32 grid = std::min<u32>(grid_for(w->ev_cap, 256), 2048u);
    hipLaunchKernelGGL(k_ev_merge, dim3(grid), dim3(256), 0, st, A, w->events.p, w->ev_src.p, w->ev_pts.p);
    sel::exclusive_scan(st, w->ev_cap, w->ev_counts.p + 3, EvCount{w->ev_pts.p}, w->ev_off.p, w->tiles.p,
                        w->d_npts.p);
    hipLaunchKernelGGL(k_ev_points, dim3(grid), dim3(256), 0, st, A, w->ev_src.p, w->ev_off.p, w->events.p,
                       w->points.p, w->slot_mgr.p, w->managers.p, w->impulses.p, w->pt_cap);
    FYZZ_HIP_CHECK(hipGetLastError());
    (void)c_cur;
}

// The island pass of a step, after the narrowphase and before the solver
// list: wakes (activate_published_islands), then sleeps (the accurate
// third pass's deactivate_ready_accurate_islands), as Scene::step orders
// them (scene.cpp:10747-10793). c: the parity of the previous step's
// pair-aligned buffers (its touching flags).
void islands(fyzz_hip_world* w, const sap::Lists& L, int c) {
    using namespace fyzz_hip::world;
    hipStream_t st = w->ctx->stream;
    const u32 n = w->n, cap = w->pair_cap;
    ProfScope ps(&w->ctx->prof, "world.islands", st);
    // the speculative graph keeps this step's lost pairs until its third
    // pass; worlds whose broadphase lists no losses (per-scene) keep the
    // whole previous list, a superset with the same components
    const u64* gone = L.lost ? L.lost : L.prev;
    const u32* d_gone = L.lost ? L.d_lost : L.d_prev;
    IslandArgs A{n,           cap,         w->bodies.p,    L.cur,         L.d_cur,
                 gone,        d_gone,      w->src.p,       w->first.p,    w->second.p,
                 w->counts_bb.p, w->touch_bb[c].p, w->has_floor ? w->counts_pl.p : nullptr, w->touch_pl.p,
                 w->isl_spec.p, w->isl_acc.p, w->isl_root_wake.p, w->isl_root_poison.p, w->isl_root_awake.p,
                 w->isl_ctr.p, w->isl_asleep.p};
    const dim3 gp(std::min<u32>(grid_for(std::max(cap, n), 256), 2048u)), gb(std::min<u32>(grid_for(n, 256), 1024u));
    hipLaunchKernelGGL(k_isl_mark, gp, dim3(256), 0, st, A);
    hipLaunchKernelGGL(k_isl_count, gb, dim3(256), 0, st, A);
    hipLaunchKernelGGL(k_isl_spec_mixed, gp, dim3(256), 0, st, A);
    hipLaunchKernelGGL(k_isl_spec_link, gp, dim3(256), 0, st, A);
    hipLaunchKernelGGL(k_isl_spec_compress, gb, dim3(256), 0, st, A);
    hipLaunchKernelGGL(k_isl_spec_giant, dim3(1), dim3(256), 0, st, A);
    hipLaunchKernelGGL(k_isl_union_spec, gp, dim3(256), 0, st, A);
    hipLaunchKernelGGL(k_isl_wake_roots, gb, dim3(256), 0, st, A);
    hipLaunchKernelGGL(k_isl_wake, gb, dim3(256), 0, st, A);
    hipLaunchKernelGGL(k_isl_union_acc, gp, dim3(256), 0, st, A);
    hipLaunchKernelGGL(k_isl_sleep_roots, gb, dim3(256), 0, st, A);
    hipLaunchKernelGGL(k_isl_sleep, gb, dim3(256), 0, st, A);
    FYZZ_HIP_CHECK(hipGetLastError());
}

// Compactions for the native order: this step's found pairs (no previous
// index), touch changes of box/box pairs and of plane managers.
struct NatFoundPred {
    const u32* src;
    __device__ bool operator()(u32 i) const { return src[i] == ::fyzz_hip::world::kNone; }
};
struct NatKeyWrite {
    const u64* cur;
    u64* out;
    __device__ void operator()(u32 i, u32 o) const { out[o] = cur[i]; }
};
// touch changes: a current pair that touches now and did not before
// (`want` 1), or that touched before and no longer does (`want` 0)
struct NatTouchChangePred {
    const u32* counts_bb;
    const u32* src;
    const u32* touch_prev;
    u32 want;
    __device__ bool operator()(u32 i) const {
        const bool now = counts_bb[i] != 0u;
        const bool before = src[i] != ::fyzz_hip::world::kNone && touch_prev[src[i]] != 0u;
        return want ? (now && !before) : (!now && before);
    }
};
// a pair the broadphase lost that touched on the last step
struct NatLostTouchPred {
    const u64* lost;
    const u64* prev;
    const u32* d_prev;
    u32 cap;
    const u32* touch_prev;
    __device__ bool operator()(u32 j) const {
        const u32 np = min(*d_prev, cap);
        const u32 o = ::fyzz_hip::world::lower_bound64(prev, np, lost[j]);
        return o < np && prev[o] == lost[j] && touch_prev[o] != 0u;
    }
};
struct NatPlanePred {
    const u32* counts_pl;
    const u32* touch_pl;
    u32 want;  // 1: began touching, 0: stopped
    __device__ bool operator()(u32 b) const {
        const bool now = counts_pl[b] != 0u, before = touch_pl[b] != 0u;
        return want ? (now && !before) : (!now && before);
    }
};
struct NatSlotWrite {
    u32* out;
    __device__ void operator()(u32 b, u32 o) const { out[o] = b; }
};
// The solver list's candidate index of each ordered manager: a plane
// manager (key hi bit set, box slot below) at pair count + slot, a box/box
// pair found by binary search in the current pair list. A key that names
// no touching manager raises bit 0 of *error.
__global__ void k_native_index(const u64* order, u32 count, const u64* cur, const u32* d_cur, u32 cap,
                               const u32* counts_bb, const u32* counts_pl, u32 n, u32* out, u32* error) {
    const u32 nbb = min(*d_cur, cap);
    for (u32 o = blockIdx.x * blockDim.x + threadIdx.x; o < count; o += gridDim.x * blockDim.x) {
        const u64 k = order[o];
        u32 idx;
        bool ok;
        if (k >> 63) {
            const u32 b = static_cast<u32>(k & 0xffffffffu);
            ok = b < n && counts_pl && counts_pl[b] != 0u;
            idx = nbb + (b < n ? b : 0u);
        } else {
            const u32 j = ::fyzz_hip::world::lower_bound64(cur, nbb, k);
            ok = j < nbb && cur[j] == k && counts_bb[j] != 0u;
            idx = j < nbb ? j : 0u;
        }
        out[o] = idx;
        if (!ok) atomicOr(error, 1u);
    }
}

// The native scene's solver order for this step (Scene::step's island
// bookkeeping, scene.cpp:10747-10793 and 7174-7260): the device compacts
// the step's found pairs and its touch changes (c_prev: the parity of the
// last step's touching flags); the host turns them, the broadphase's lost
// pairs and the last solve's ready flags into the authority's change
// lists, runs it, and hands the order back as pair keys, which the device
// maps to solver-list candidates in w->native_list. Synchronizes. Returns
// the order's length.
u32 native_order_list(fyzz_hip_world* w, const sap::Lists& L, int c_prev) {
    using namespace fyzz_hip::world;
    hipStream_t st = w->ctx->stream;
    const u32 n = w->n, cap = w->pair_cap;
    const u32 ids = static_cast<u32>(w->slot_of_id.size());
    // [0] touches found, [1] pairs found, [2] plane touches found, [3] errors,
    // [4] touches lost (pairs kept), [5] touches lost with their pair, [6] plane touches lost
    u32* cnt = w->native_counts.p;
    const u32* tprev = w->touch_bb[c_prev].p;
    hipLaunchKernelGGL(k_fill, dim3(1), dim3(64), 0, st, cnt, 8u, 0u);
    hipLaunchKernelGGL(k_flag_bytes, dim3(std::min<u32>(grid_for(n, 256), 1024u)), dim3(256), 0, st, n, w->bodies.p,
                       w->flag_bytes.p);
    sel::select(st, cap, L.d_cur, NatTouchChangePred{w->counts_bb.p, w->src.p, tprev, 1u},
                NatKeyWrite{L.cur, w->native_keys.p}, w->tiles.p, cnt + 0);
    sel::select(st, cap, L.d_cur, NatFoundPred{w->src.p}, NatKeyWrite{L.cur, w->native_found.p}, w->tiles.p,
                cnt + 1);
    sel::select(st, cap, L.d_cur, NatTouchChangePred{w->counts_bb.p, w->src.p, tprev, 0u},
                NatKeyWrite{L.cur, w->native_keys2.p}, w->tiles.p, cnt + 4);
    sel::select(st, cap, L.d_lost, NatLostTouchPred{L.lost, L.prev, L.d_prev, cap, tprev},
                NatKeyWrite{L.lost, w->native_keys3.p}, w->tiles.p, cnt + 5);
    if (w->has_floor) {
        hipLaunchKernelGGL(k_fill, dim3(1), dim3(64), 0, st, w->native_n.p, 1u, n);
        sel::select(st, n, w->native_n.p, NatPlanePred{w->counts_pl.p, w->touch_pl.p, 1u},
                    NatSlotWrite{w->native_planes.p}, w->tiles.p, cnt + 2);
        sel::select(st, n, w->native_n.p, NatPlanePred{w->counts_pl.p, w->touch_pl.p, 0u},
                    NatSlotWrite{w->native_planes2.p}, w->tiles.p, cnt + 6);
    }
    u32 c[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    u32 nlost = 0, nbb = 0;
    FYZZ_HIP_CHECK(hipMemcpyAsync(c, cnt, sizeof c, hipMemcpyDeviceToHost, st));
    FYZZ_HIP_CHECK(hipMemcpyAsync(&nlost, L.d_lost, sizeof nlost, hipMemcpyDeviceToHost, st));
    FYZZ_HIP_CHECK(hipMemcpyAsync(&nbb, L.d_cur, sizeof nbb, hipMemcpyDeviceToHost, st));
    FYZZ_HIP_CHECK(hipStreamSynchronize(st));
    for (int k : {0, 1, 4, 5}) c[k] = std::min(c[k], cap);
    c[2] = std::min(c[2], n);
    c[6] = std::min(c[6], n);
    nlost = std::min(nlost, cap);
    nbb = std::min(nbb, cap);
    std::vector<u64> tkeys(c[0]), fkeys(c[1]), lkeys(nlost), tl_keep(c[4]), tl_gone(c[5]);
    std::vector<u32> planes(c[2]), planes_lost(c[6]);
    std::vector<uint8_t> flags(n);
    auto fetch = [&](void* dst, const void* src, std::size_t bytes) {
        if (bytes) FYZZ_HIP_CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, st));
    };
    fetch(tkeys.data(), w->native_keys.p, 8ull * c[0]);
    fetch(fkeys.data(), w->native_found.p, 8ull * c[1]);
    fetch(lkeys.data(), L.lost, 8ull * nlost);
    fetch(tl_keep.data(), w->native_keys2.p, 8ull * c[4]);
    fetch(tl_gone.data(), w->native_keys3.p, 8ull * c[5]);
    fetch(planes.data(), w->native_planes.p, 4ull * c[2]);
    fetch(planes_lost.data(), w->native_planes2.p, 4ull * c[6]);
    fetch(flags.data(), w->flag_bytes.p, n);
    FYZZ_HIP_CHECK(hipStreamSynchronize(st));

    const auto& id = w->id_of_slot;
    auto id_key = [](u32 a, u32 b) { return a < b ? (u64(a) << 32) | b : (u64(b) << 32) | a; };
    auto slot_pair = [&](u64 k) { return id_key(id[static_cast<u32>(k >> 32)], id[static_cast<u32>(k)]); };
    // broadphase changes (the plane pairs every box from the first step on)
    std::vector<u64> found, lost;
    found.reserve(fkeys.size() + (w->step_index == 1 && w->has_floor ? n : 0));
    for (u64 k : fkeys) found.push_back(slot_pair(k));
    if (w->step_index == 1 && w->has_floor)
        for (u32 k = 0; k < n; ++k) found.push_back(id_key(w->floor_id, id[k]));
    for (u64 k : lkeys) lost.push_back(slot_pair(k));
    std::sort(found.begin(), found.end());
    std::sort(lost.begin(), lost.end());
    // touch changes in id order
    std::vector<u64> tfound, tlost;
    for (u32 b : planes) tfound.push_back(id_key(w->floor_id, id[b]));
    for (u64 k : tkeys) tfound.push_back(slot_pair(k));
    for (u32 b : planes_lost) tlost.push_back(id_key(w->floor_id, id[b]));
    for (u64 k : tl_keep) tlost.push_back(slot_pair(k));
    for (u64 k : tl_gone) tlost.push_back(slot_pair(k));
    std::sort(tfound.begin(), tfound.end());
    std::sort(tlost.begin(), tlost.end());
    auto unpack = [](const std::vector<u64>& v) {
        std::vector<u32> out(2 * v.size());
        for (std::size_t i = 0; i < v.size(); ++i) {
            out[2 * i] = static_cast<u32>(v[i] >> 32);
            out[2 * i + 1] = static_cast<u32>(v[i]);
        }
        return out;
    };
    const auto pf = unpack(found), pl = unpack(lost), tf = unpack(tfound), tl = unpack(tlost);
    std::vector<uint8_t> ready(ids, 0);
    for (u32 k = 0; k < n; ++k) ready[id[k]] = (flags[k] & 2u) ? 1 : 0;
    fyzz_island_events ev{pf.data(), static_cast<u32>(found.size()), pl.data(), static_cast<u32>(lost.size()),
                          tf.data(), static_cast<u32>(tfound.size()), tl.data(), static_cast<u32>(tlost.size()),
                          ready.data()};
    const u32 most = nbb + (w->has_floor ? n : 0u);  // every touching manager at most
    std::vector<u32> order(2 * std::max<u32>(most, 1u)), tasks(ids + 1);
    w->native_sleeping.assign(ids, 0);
    fyzz_island_order out{order.data(), most, 0, tasks.data(), ids + 1, 0, w->native_sleeping.data()};
    const int rc = fyzz_hip_islands_step(w->order_auth, &ev, &out);
    FYZZ_REQUIRE(rc == 0, kLogicError,
                 std::string("world: native solver order: ") + fyzz_hip_islands_last_error(w->order_auth));
    // the order as slot keys (a plane manager: high bit | box slot) for the device's lookup
    std::vector<u64> okeys(out.manager_count);
    for (u32 m = 0; m < out.manager_count; ++m) {
        const u32 p = order[2 * m], q = order[2 * m + 1];
        if (w->has_floor && (p == w->floor_id || q == w->floor_id)) {
            okeys[m] = (1ull << 63) | w->slot(p == w->floor_id ? q : p);
        } else {
            u32 sp = w->slot(p), sq = w->slot(q);
            if (sp > sq) std::swap(sp, sq);
            okeys[m] = (u64(sp) << 32) | sq;
        }
    }
    if (!okeys.empty()) {
        FYZZ_HIP_CHECK(hipMemcpy(w->native_order_keys.p, okeys.data(), 8ull * okeys.size(), hipMemcpyHostToDevice));
        hipLaunchKernelGGL(k_native_index, dim3(std::min<u32>(grid_for(out.manager_count, 256), 2048u)), dim3(256), 0,
                           st, w->native_order_keys.p, out.manager_count, L.cur, L.d_cur, cap, w->counts_bb.p,
                           w->has_floor ? w->counts_pl.p : nullptr, n, w->native_list.p, cnt + 3);
        FYZZ_HIP_CHECK(hipGetLastError());
    }
    return out.manager_count;
}

// The device island pass and the host authority decide sleep by the same
// rules; the solver list would be wrong if they disagreed.
void native_check_sleep(fyzz_hip_world* w) {
    using namespace fyzz_hip::world;
    hipStream_t st = w->ctx->stream;
    const u32 n = w->n;
    hipLaunchKernelGGL(k_flag_bytes, dim3(std::min<u32>(grid_for(n, 256), 1024u)), dim3(256), 0, st, n, w->bodies.p,
                       w->flag_bytes.p);
    std::vector<uint8_t> flags(n);
    u32 error = 0;
    FYZZ_HIP_CHECK(hipMemcpyAsync(flags.data(), w->flag_bytes.p, n, hipMemcpyDeviceToHost, st));
    FYZZ_HIP_CHECK(hipMemcpyAsync(&error, w->native_counts.p + 3, sizeof error, hipMemcpyDeviceToHost, st));
    FYZZ_HIP_CHECK(hipStreamSynchronize(st));
    FYZZ_REQUIRE(!error, kLogicError, "world: native order names a manager without contacts");
    for (u32 k = 0; k < n; ++k)
        FYZZ_REQUIRE(((flags[k] & 1u) != 0) == (w->native_sleeping[w->id_of_slot[k]] != 0), kLogicError,
                     "world: native order: the island pass and the island authority disagree on sleep");
}

void step(fyzz_hip_world* w, float dt) {
    using namespace fyzz_hip::world;
    if (!w->uploaded) upload(w);
    const u32 n = w->n;
    hipStream_t st = w->ctx->stream;
    ++w->step_index;
    if (!n) return;
    const u32 cap = w->pair_cap;
    hipLaunchKernelGGL(k_step_begin, dim3(1), dim3(1), 0, st, w->d_step.p);
    hipLaunchKernelGGL(k_gather, dim3(grid_for(n, 256)), dim3(256), 0, st, n, w->bodies.p, w->pos.p, w->quat.p);
    const sap::Lists L = w->n_scenes ? scene_broadphase(w) : sap_broadphase(w);
    hipLaunchKernelGGL(k_fill, dim3(std::min<u32>(grid_for(n, 256), 1024u)), dim3(256), 0, st, w->counted.p, n, 0u);
    hipLaunchKernelGGL(k_pairs, dim3(std::min<u32>(grid_for(cap, 256), 4096u)), dim3(256), 0, st, cap, L.cur, L.d_cur,
                       L.prev, L.d_prev, w->first.p, w->second.p, w->src.p, w->counted.p, L.src, n);
    const u32 planes = w->has_floor ? 1u : 0u;
    hipLaunchKernelGGL(k_counted, dim3(grid_for(n, 256)), dim3(256), 0, st, n, w->bodies.p, w->counted.p, planes);
    const int c = w->ccur, c2 = c ^ 1;  // previous / current pair-aligned buffers
    pcm::ListPairs lp{cap, L.d_cur, w->first.p, w->second.p, w->cache_bb[c].p, w->src.p, w->cache_bb[c2].p};
    const uint8_t* sphere = w->has_spheres ? w->sphere.p : nullptr;
    pcm::box_box_list(w->ctx, lp, w->pos.p, w->quat.p, w->half.p, kContactDistance, w->contacts_bb.p,
                      w->counts_bb.p, w->flags_bb.p, w->pcm_scratch.p, sphere);
    w->ccur = c2;
    if (w->has_floor) {
        pcm::plane_all(w->ctx, n, w->pos.p, w->quat.p, w->half.p, w->floor.position, kContactDistance, w->cache_pl.p,
                       w->contacts_pl.p, w->counts_pl.p, w->flags_pl.p, sphere);
    }
    const u32 native_count = w->native_order ? native_order_list(w, L, c) : 0u;
    islands(w, L, c);
    if (w->native_order) native_check_sleep(w);
    // touching managers, in order
    const u32 extra = w->has_floor ? n : 0u;
    hipLaunchKernelGGL(k_total, dim3(1), dim3(1), 0, st, L.d_cur, cap, extra, w->d_total.p);
    const int h = w->hcur;
    {
        ProfScope ps(&w->ctx->prof, "world.managers", st);
        MgrWrite wr{L.d_cur,       cap,          w->first.p,         w->second.p,   w->src.p,
                    w->counts_bb.p, w->contacts_bb.p, w->counts_pl.p, w->contacts_pl.p, w->stamp[h].p,
                    w->d_step.p,   w->managers.p, w->mgr_slot.p, n};
        if (w->native_order) {
            hipLaunchKernelGGL(k_fill, dim3(1), dim3(64), 0, st, w->d_nm.p, 1u, native_count);
            if (native_count)
                hipLaunchKernelGGL(k_mgr_gather, dim3(std::min<u32>(grid_for(native_count, 256), 2048u)), dim3(256),
                                   0, st, w->native_list.p, native_count, wr);
        } else {
            MgrPred pred{L.d_cur, cap, w->counts_bb.p, w->counts_pl.p, w->first.p, w->bodies.p};
            sel::select(st, w->mgr_cap, w->d_total.p, pred, wr, w->tiles.p, w->d_nm.p);
        }
    }
    dy::step_device(w->dy, dt, n, w->bodies.p, 1 + static_cast<u32>(w->host_statics.size()), w->statics.p, w->mgr_cap, w->d_nm.p, w->managers.p, w->hist[h].p,
                    w->mgr_cap, w->patches.p, w->impulses.p, w->parts.p);
    hipLaunchKernelGGL(k_scatter_patches, dim3(std::min<u32>(grid_for(w->mgr_cap, 256), 2048u)), dim3(256), 0, st,
                       w->d_nm.p, w->mgr_cap, w->patches.p, w->mgr_slot.p, w->hist[h ^ 1].p, w->stamp[h ^ 1].p,
                       w->d_step.p);
    hipLaunchKernelGGL(k_accumulate, dim3(1), dim3(1), 0, st, L.d_cur, cap, w->d_nm.p, dy::counters(w->dy),
                       L.d_found, L.d_lost, w->totals.p);
    if (w->notify) publish_events(w, L, c, c2);
    hipLaunchKernelGGL(k_touch, dim3(std::min<u32>(grid_for(std::max(cap, n), 256), 2048u)), dim3(256), 0, st, cap,
                       L.d_cur, w->counts_bb.p, w->touch_bb[c2].p, n, w->has_floor ? w->counts_pl.p : nullptr,
                       w->touch_pl.p);
    FYZZ_HIP_CHECK(hipGetLastError());
    w->hcur = h ^ 1;
}

// One step through a captured graph. The step's host-side toggles (pair
// list, caches, friction history) flip every step, so a graph per parity
// is captured on first use (or when dt changes) and replayed afterwards;
// the history stamps come from the device step counter. The graphs run on
// the world's own stream, ordered after and before the context stream's
// work by events.
void step_graph(fyzz_hip_world* w, float dt) {
    if (!w->uploaded) upload(w);
    if (!w->n) {
        ++w->step_index;
        return;
    }
    // The incremental SAP builds its endpoint arrays on its first frame;
    // the graphs record the incremental frame (whose stamps only label
    // records, as the world lists no dirty boxes).
    if (!w->n_scenes && !sap::built(w->sap)) {
        step(w, dt);
        return;
    }
    if (!w->gstream) {
        FYZZ_HIP_CHECK(hipStreamCreateWithFlags(&w->gstream, hipStreamNonBlocking));
        FYZZ_HIP_CHECK(hipEventCreateWithFlags(&w->gev_in, hipEventDisableTiming));
        FYZZ_HIP_CHECK(hipEventCreateWithFlags(&w->gev_out, hipEventDisableTiming));
    }
    const int parity = static_cast<int>(w->step_index & 1u);
    hipStream_t outer = w->ctx->stream;
    FYZZ_HIP_CHECK(hipEventRecord(w->gev_in, outer));
    FYZZ_HIP_CHECK(hipStreamWaitEvent(w->gstream, w->gev_in, 0));
    if (!w->gexec[parity] || w->gdt[parity] != dt) {
        if (w->gexec[parity]) {
            FYZZ_HIP_CHECK(hipGraphExecDestroy(w->gexec[parity]));
            w->gexec[parity] = nullptr;
        }
        const bool prof = w->ctx->prof.enabled;
        w->ctx->prof.enabled = false;
        w->ctx->stream = w->gstream;
        dy::set_plain_launch(w->dy, true);
        hipGraph_t graph = nullptr;
        auto restore = [&] {
            w->ctx->stream = outer;
            w->ctx->prof.enabled = prof;
            dy::set_plain_launch(w->dy, false);
        };
        FYZZ_HIP_CHECK(hipStreamBeginCapture(w->gstream, hipStreamCaptureModeThreadLocal));
        try {
            step(w, dt);
        } catch (...) {
            (void)hipStreamEndCapture(w->gstream, &graph);
            if (graph) (void)hipGraphDestroy(graph);
            restore();
            throw;
        }
        const hipError_t e = hipStreamEndCapture(w->gstream, &graph);
        restore();
        FYZZ_HIP_CHECK(e);
        const hipError_t ie = hipGraphInstantiate(&w->gexec[parity], graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        FYZZ_HIP_CHECK(ie);
        w->gdt[parity] = dt;
    } else {
        // the toggles a recorded step flips
        ++w->step_index;
        w->ccur ^= 1;
        w->hcur ^= 1;
        w->scur ^= 1;
        if (!w->n_scenes) sap::replay_advance(w->sap);
    }
    FYZZ_HIP_CHECK(hipGraphLaunch(w->gexec[parity], w->gstream));
    FYZZ_HIP_CHECK(hipEventRecord(w->gev_out, w->gstream));
    FYZZ_HIP_CHECK(hipStreamWaitEvent(outer, w->gev_out, 0));
}

// The pair list the next step starts from (the last step's current list).
sap::Lists current_lists(fyzz_hip_world* w) {
    if (!w->n_scenes) return sap::lists(w->sap);
    sap::Lists L{};
    L.cur = w->scene_pairs[w->scur].p;
    L.d_cur = w->scene_n.p + w->scur;
    L.cap = w->pair_cap;
    return L;
}

// One body record of a stepped world, read and written synchronously on
// the world's stream (authoring calls are host API calls, not step work).
fyzz_dy_body read_body(fyzz_hip_world* w, u32 slot) {
    fyzz_dy_body b{};
    hipStream_t st = w->ctx->stream;
    FYZZ_HIP_CHECK(hipMemcpyAsync(&b, w->bodies.p + slot, sizeof b, hipMemcpyDeviceToHost, st));
    FYZZ_HIP_CHECK(hipStreamSynchronize(st));
    return b;
}

void write_body(fyzz_hip_world* w, u32 slot, const fyzz_dy_body& b) {
    hipStream_t st = w->ctx->stream;
    FYZZ_HIP_CHECK(hipMemcpyAsync(w->bodies.p + slot, &b, sizeof b, hipMemcpyHostToDevice, st));
    FYZZ_HIP_CHECK(hipStreamSynchronize(st));
}

// The record a setter edits: the host copy before the first step, the
// device's afterwards (written back by commit_body).
fyzz_dy_body& edit_body(fyzz_hip_world* w, u32 id, fyzz_dy_body& scratch_copy, u32* slot_out) {
    FYZZ_REQUIRE(!(w->uploaded && w->native_order), kLogicError,
                 "world: the native solver order takes no body edits after the first step");
    fyzz_dy_body& host = w->body(id);
    const u32 slot = static_cast<u32>(w->slot_of_id[id]);
    if (slot_out) *slot_out = slot;
    if (!w->uploaded) return host;
    scratch_copy = read_body(w, slot);
    return scratch_copy;
}

void commit_body(fyzz_hip_world* w, u32 slot, const fyzz_dy_body& b) {
    if (!w->uploaded) return;
    write_body(w, slot, b);
}

// Scene::set_pose after the first step: the record, then the cache,
// history and peer-wake pass (scene.cpp:589-700).
void author_pose(fyzz_hip_world* w, u32 slot, bool changed) {
    using namespace fyzz_hip::world;
    hipStream_t st = w->ctx->stream;
    const sap::Lists L = current_lists(w);
    const u32 cap = w->pair_cap;
    hipLaunchKernelGGL(k_author_pose, dim3(std::min<u32>(grid_for(std::max(cap, w->mgr_cap), 256), 1024u)), dim3(256),
                       0, st, slot, changed ? 1 : 0, cap, L.cur, L.d_cur, w->cache_bb[w->ccur].p, w->cache_pl.p,
                       w->has_floor ? 1 : 0, w->stamp[w->hcur].p, w->counts_bb.p, w->bodies.p, w->n);
    FYZZ_HIP_CHECK(hipGetLastError());
    FYZZ_HIP_CHECK(hipStreamSynchronize(st));
}

}  // namespace

extern "C" {

int fyzz_hip_world_step_graph(fyzz_hip_world* w, float dt) {
    if (!w) return kInvalidArgument;
    return guarded(w->ctx, [&] {
        FYZZ_REQUIRE(std::isfinite(dt) && dt > 0.0f, kInvalidArgument, "world: dt must be positive");
        FYZZ_REQUIRE(!w->native_order, kLogicError, "world: the native solver order steps eagerly (host island pass)");
        step_graph(w, dt);
    });
}

// Autotuner: picks the Dy partition and PGS workgroup counts (the grid
// barrier's cost against the work per sweep) by timing `trial_steps` real
// steps per candidate, twice per candidate; one parameter at a time. The
// steps advance the simulation (use it in place of warm-up steps); the
// counts do not change results. Writes the chosen counts.
int fyzz_hip_world_autotune(fyzz_hip_world* w, float dt, uint32_t trial_steps, int* partition_blocks,
                            int* pgs_blocks) {
    if (!w) return kInvalidArgument;
    return guarded(w->ctx, [&] {
        FYZZ_REQUIRE(trial_steps > 0, kInvalidArgument, "world: autotune needs at least one step per trial");
        FYZZ_REQUIRE(!w->native_order, kLogicError, "world: the native solver order steps eagerly (host island pass)");
        if (!w->uploaded) upload(w);
        const bool graph = w->n_scenes != 0;
        hipStream_t st = w->ctx->stream;
        hipEvent_t a, b;
        FYZZ_HIP_CHECK(hipEventCreate(&a));
        FYZZ_HIP_CHECK(hipEventCreate(&b));
        auto trial = [&](int pb, int sb) {
            dy::set_coop_blocks(w->dy, pb, sb);
            drop_graphs(w);  // graphs hold their launch shapes
            for (u32 k = 0; k < 2; ++k) graph ? step_graph(w, dt) : step(w, dt);  // settle both graphs
            FYZZ_HIP_CHECK(hipEventRecord(a, st));
            for (u32 k = 0; k < trial_steps; ++k) graph ? step_graph(w, dt) : step(w, dt);
            FYZZ_HIP_CHECK(hipEventRecord(b, st));
            FYZZ_HIP_CHECK(hipEventSynchronize(b));
            float ms = 0.f;
            FYZZ_HIP_CHECK(hipEventElapsedTime(&ms, a, b));
            return ms;
        };
        // The world changes under the trials (a settling scene gains
        // contacts), so each parameter's candidates run twice, in
        // increasing then decreasing order, and their two times are summed:
        // a drift linear in time adds the same to every candidate.
        constexpr int kCand = 5;
        const int cand[kCand] = {16, 32, 64, 128, 256};
        auto pick = [&](auto run) {
            float t[kCand] = {};
            for (int pass = 0; pass < 2; ++pass)
                for (int k = 0; k < kCand; ++k) {
                    const int i = pass ? kCand - 1 - k : k;
                    t[i] += run(cand[i]);
                }
            int best = 0;
            for (int i = 1; i < kCand; ++i)
                if (t[i] < t[best]) best = i;
            return cand[best];
        };
        int pb0 = 0, sb0 = 0;
        dy::coop_blocks_of(w->dy, &pb0, &sb0);
        const int best_p = pick([&](int c) { return trial(c, sb0); });
        const int best_s = pick([&](int c) { return trial(best_p, c); });
        dy::set_coop_blocks(w->dy, best_p, best_s);
        drop_graphs(w);
        (void)hipEventDestroy(a);
        (void)hipEventDestroy(b);
        int pb = 0, sb = 0;
        dy::coop_blocks_of(w->dy, &pb, &sb);
        if (partition_blocks) *partition_blocks = pb;
        if (pgs_blocks) *pgs_blocks = sb;
    });
}

int fyzz_hip_world_set_workgroups(fyzz_hip_world* w, int partition_blocks, int pgs_blocks, int* partition_in_effect,
                                  int* pgs_in_effect) {
    if (!w) return kInvalidArgument;
    return guarded(w->ctx, [&] {
        FYZZ_REQUIRE(partition_blocks >= 0 && pgs_blocks >= 0, kInvalidArgument,
                     "world: workgroup counts must not be negative");
        int pb = 0, sb = 0;
        if (w->uploaded) {
            dy::coop_blocks_of(w->dy, &pb, &sb);
            dy::set_coop_blocks(w->dy, partition_blocks ? partition_blocks : pb, pgs_blocks ? pgs_blocks : sb);
            drop_graphs(w);
            dy::coop_blocks_of(w->dy, &pb, &sb);
        } else {
            // kept until the first step creates the Dy workspace
            if (w->req_partition_blocks) {
                pb = w->req_partition_blocks;
                sb = w->req_pgs_blocks;
            } else {
                dy::effective_blocks(w->ctx->device, 0, 0, &pb, &sb);
            }
            dy::effective_blocks(w->ctx->device, partition_blocks ? partition_blocks : pb, pgs_blocks ? pgs_blocks : sb,
                                 &pb, &sb);
            w->req_partition_blocks = pb;
            w->req_pgs_blocks = sb;
        }
        if (partition_in_effect) *partition_in_effect = pb;
        if (pgs_in_effect) *pgs_in_effect = sb;
    });
}

int fyzz_hip_world_create(fyzz_hip_context* ctx, uint32_t max_bodies, fyzz_hip_world** out) {
    if (!out) return kInvalidArgument;
    *out = nullptr;
    return guarded(ctx, [&] {
        FYZZ_REQUIRE(max_bodies > 0 && max_bodies <= (1u << 26), kInvalidArgument, "world_create: bad max_bodies");
        auto w = std::make_unique<fyzz_hip_world>();
        w->ctx = ctx;
        w->max_bodies = max_bodies;
        w->floor.orientation[3] = 1.0f;
        w->floor.static_friction = 0.5f;
        w->floor.dynamic_friction = 0.5f;
        w->floor.restitution = 0.2f;
        *out = w.release();
    });
}

int fyzz_hip_world_destroy(fyzz_hip_world* w) {
    if (!w) return kInvalidArgument;
    (void)hipSetDevice(w->ctx->device);
    (void)hipStreamSynchronize(w->ctx->stream);
    delete w;
    return kOk;
}

int fyzz_hip_world_add_static_floor(fyzz_hip_world* w, float height, uint32_t* body_id) {
    if (!w) return kInvalidArgument;
    return guarded(w->ctx, [&] {
        w->require_authoring();
        FYZZ_REQUIRE(!w->has_floor, kLogicError, "world: one floor plane per world");
        FYZZ_REQUIRE(std::isfinite(height), kInvalidArgument, "world: floor height must be finite");
        w->has_floor = true;
        w->floor_id = static_cast<u32>(w->slot_of_id.size());
        w->slot_of_id.push_back(-1);
        w->floor.position[1] = height;
        if (body_id) *body_id = w->floor_id;
    });
}

int fyzz_hip_world_add_dynamic_box(fyzz_hip_world* w, const float* half_extents, float density,
                                   uint32_t* body_id) {
    if (!w) return kInvalidArgument;
    return guarded(w->ctx, [&] {
        w->require_authoring();
        FYZZ_REQUIRE(half_extents, kInvalidArgument, "world: null half extents");
        FYZZ_REQUIRE(w->host_bodies.size() < w->max_bodies, kOverflow, "world: max_bodies reached");
        for (int k = 0; k < 3; ++k)
            FYZZ_REQUIRE(std::isfinite(half_extents[k]) && half_extents[k] > 0.0f, kInvalidArgument,
                         "world: half extents must be positive and finite");
        FYZZ_REQUIRE(std::isfinite(density) && density > 0.0f, kInvalidArgument, "world: density must be positive");
        fyzz_dy_body b = default_body();
        world::box_mass(half_extents, density, b.inverse_mass, b.inverse_inertia);
        const u32 id = static_cast<u32>(w->slot_of_id.size());
        w->slot_of_id.push_back(static_cast<int>(w->host_bodies.size()));
        w->host_bodies.push_back(b);
        w->host_half.insert(w->host_half.end(), half_extents, half_extents + 3);
        w->host_sphere.push_back(0);
        if (body_id) *body_id = id;
    });
}

// Scene::add_static_box / add_static_sphere (scene.cpp:441-488): a static
// body with one shape at `position`, identity orientation (set_body_pose
// may turn it before the first step), the default material.
namespace {
int add_static(fyzz_hip_world* w, const float half[3], bool sphere, const float* position, uint32_t* body_id) {
    return guarded(w->ctx, [&] {
        w->require_authoring();
        FYZZ_REQUIRE(!w->n_scenes, kLogicError, "world: batched scenes hold no static bodies but the floor");
        FYZZ_REQUIRE(position, kInvalidArgument, "world: null position");
        for (int k = 0; k < 3; ++k)
            FYZZ_REQUIRE(std::isfinite(half[k]) && half[k] > 0.0f && std::isfinite(position[k]), kInvalidArgument,
                         "world: static extents must be positive and finite, the position finite");
        fyzz_hip_world::StaticBody S{};
        for (int k = 0; k < 3; ++k) S.half[k] = half[k], S.pos[k] = position[k];
        S.quat[3] = 1.0f;
        S.sphere = sphere ? 1 : 0;
        S.static_friction = S.dynamic_friction = 0.5f;
        S.restitution = 0.2f;
        const u32 id = static_cast<u32>(w->slot_of_id.size());
        w->slot_of_id.push_back(-2 - static_cast<int>(w->host_statics.size()));
        w->host_statics.push_back(S);
        if (sphere) w->has_spheres = true;
        if (body_id) *body_id = id;
    });
}
}  // namespace

int fyzz_hip_world_add_static_box(fyzz_hip_world* w, const float* half_extents, const float* position,
                                  uint32_t* body_id) {
    if (!w) return kInvalidArgument;
    if (!half_extents) return kInvalidArgument;
    return add_static(w, half_extents, false, position, body_id);
}

int fyzz_hip_world_add_static_sphere(fyzz_hip_world* w, float radius, const float* position, uint32_t* body_id) {
    if (!w) return kInvalidArgument;
    const float h[3] = {radius, radius, radius};
    return add_static(w, h, true, position, body_id);
}

// Scene::add_dynamic_sphere (scene.cpp:491-525): mass 4/3 pi r^3 density and
// inertia 2/5 m r^2 in double, their inverses rounded to float (an inertia
// whose inverse a float cannot hold leaves the body without rotation).
int fyzz_hip_world_add_dynamic_sphere(fyzz_hip_world* w, float radius, float density, uint32_t* body_id) {
    if (!w) return kInvalidArgument;
    return guarded(w->ctx, [&] {
        w->require_authoring();
        FYZZ_REQUIRE(w->host_bodies.size() < w->max_bodies, kOverflow, "world: max_bodies reached");
        FYZZ_REQUIRE(std::isfinite(radius) && radius > 0.0f && std::isfinite(density) && density > 0.0f,
                     kInvalidArgument, "world: sphere radius and density must be positive and finite");
        const double r = radius;
        const double mass = 4.1887902047863909846168578443727 * r * r * r * density;
        const double inv_mass = 1.0 / mass;
        FYZZ_REQUIRE(std::isfinite(inv_mass) && inv_mass > 0.0 && inv_mass <= 3.402823466e+38 &&
                         static_cast<float>(inv_mass) != 0.0f,
                     kInvalidArgument, "world: sphere mass outside the supported range");
        fyzz_dy_body b = default_body();
        b.inverse_mass = static_cast<float>(inv_mass);
        const double inertia = 0.4 * mass * r * r;
        const double inv_inertia = 1.0 / inertia;
        float inv = 0.0f;
        if (std::isfinite(inv_inertia) && inv_inertia > 0.0 && inv_inertia <= 3.402823466e+38 &&
            static_cast<float>(inv_inertia) > 0.0f)
            inv = static_cast<float>(inv_inertia);
        b.inverse_inertia[0] = b.inverse_inertia[1] = b.inverse_inertia[2] = inv;
        const u32 id = static_cast<u32>(w->slot_of_id.size());
        w->slot_of_id.push_back(static_cast<int>(w->host_bodies.size()));
        w->host_bodies.push_back(b);
        w->host_half.insert(w->host_half.end(), {radius, radius, radius});
        w->host_sphere.push_back(1);
        w->has_spheres = true;
        if (body_id) *body_id = id;
    });
}

int fyzz_hip_world_add_dynamic_boxes(fyzz_hip_world* w, uint32_t n, const float* half_extents, float density,
                                     const float* positions, const float* orientations, const float* linear,
                                     const float* angular, float static_friction, float dynamic_friction,
                                     float restitution, uint32_t* first_id) {
    if (!w) return kInvalidArgument;
    return guarded(w->ctx, [&] {
        w->require_authoring();
        FYZZ_REQUIRE(n == 0 || (half_extents && positions && orientations), kInvalidArgument,
                     "world: null box arrays");
        FYZZ_REQUIRE(w->host_bodies.size() + n <= w->max_bodies, kOverflow, "world: max_bodies reached");
        FYZZ_REQUIRE(std::isfinite(density) && density > 0.0f, kInvalidArgument, "world: density must be positive");
        FYZZ_REQUIRE(static_friction >= 0.0f && dynamic_friction >= 0.0f && restitution >= 0.0f, kInvalidArgument,
                     "world: material values must be non-negative");
        if (first_id) *first_id = static_cast<u32>(w->slot_of_id.size());
        for (u32 i = 0; i < n; ++i) {
            const float* h = half_extents + 3ull * i;
            for (int k = 0; k < 3; ++k)
                FYZZ_REQUIRE(std::isfinite(h[k]) && h[k] > 0.0f, kInvalidArgument,
                             "world: half extents must be positive and finite");
            const float* q = orientations + 4ull * i;
            const float len = std::sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + q[3] * q[3]);
            FYZZ_REQUIRE(std::isfinite(len) && len > 0.0f, kInvalidArgument, "world: orientation must be non-zero");
            fyzz_dy_body b = default_body();
            world::box_mass(h, density, b.inverse_mass, b.inverse_inertia);
            for (int k = 0; k < 3; ++k) {
                b.position[k] = positions[3ull * i + k];
                if (linear) b.linear_velocity[k] = linear[3ull * i + k];
                if (angular) b.angular_velocity[k] = angular[3ull * i + k];
            }
            for (int k = 0; k < 4; ++k) b.orientation[k] = q[k] / len;
            b.static_friction = static_friction;
            b.dynamic_friction = dynamic_friction;
            b.restitution = restitution;
            w->slot_of_id.push_back(static_cast<int>(w->host_bodies.size()));
            w->host_bodies.push_back(b);
            w->host_half.insert(w->host_half.end(), h, h + 3);
            w->host_sphere.push_back(0);
        }
    });
}

// Scene::set_pose (scene.cpp:589-700). Before the first step the host
// record is edited; afterwards the device record, the current pairs' PCM
// caches, their friction history and the touching peers' wake state.
int fyzz_hip_world_set_body_pose(fyzz_hip_world* w, uint32_t body_id, const float* position,
                                 const float* orientation) {
    if (!w) return kInvalidArgument;
    return guarded(w->ctx, [&] {
        FYZZ_REQUIRE(position && orientation, kInvalidArgument, "world: null pose");
        if (w->has_floor && body_id == w->floor_id) {
            // the admitted plane's normal is fixed to +Y and it stays put
            FYZZ_REQUIRE(false, kLogicError, "world: the floor plane cannot be moved");
        }
        const float* q = orientation;
        const float mag = std::sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + q[3] * q[3]);
        bool finite = std::isfinite(mag);
        for (int k = 0; k < 3; ++k) finite = finite && std::isfinite(position[k]);
        for (int k = 0; k < 4; ++k) finite = finite && std::isfinite(q[k]);
        FYZZ_REQUIRE(finite && std::fabs(mag - 1.0f) < 1.0e-2f, kInvalidArgument, "world: rigid pose is not sane");
        if (body_id < w->slot_of_id.size() && w->slot_of_id[body_id] <= -2) {
            FYZZ_REQUIRE(!w->uploaded, kLogicError, "world: static bodies are posed before the first step");
            auto& S = w->host_statics[static_cast<u32>(-2 - w->slot_of_id[body_id])];
            const bool same = q[0] == S.quat[0] && q[1] == S.quat[1] && q[2] == S.quat[2] && q[3] == S.quat[3];
            for (int k = 0; k < 3; ++k) S.pos[k] = position[k];
            if (!same)
                for (int k = 0; k < 4; ++k) S.quat[k] = q[k] / mag;
            return;
        }
        fyzz_dy_body tmp;
        u32 slot = 0;
        fyzz_dy_body& b = edit_body(w, body_id, tmp, &slot);
        float nq[4] = {q[0] / mag, q[1] / mag, q[2] / mag, q[3] / mag};
        // a translation-only edit passes the stored orientation unchanged
        if (q[0] == b.orientation[0] && q[1] == b.orientation[1] && q[2] == b.orientation[2] &&
            q[3] == b.orientation[3])
            for (int k = 0; k < 4; ++k) nq[k] = b.orientation[k];
        bool changed = false;
        for (int k = 0; k < 3; ++k) changed = changed || b.position[k] != position[k];
        for (int k = 0; k < 4; ++k) changed = changed || b.orientation[k] != nq[k];
        world::wake_record(b);
        for (int k = 0; k < 3; ++k) b.position[k] = position[k];
        for (int k = 0; k < 4; ++k) b.orientation[k] = nq[k];
        if (!w->uploaded) return;
        commit_body(w, slot, b);
        author_pose(w, slot, changed);
    });
}

// Scene::set_velocity / set_angular_velocity (scene.cpp:730-781): the
// velocity and the body's own wake; no cache is touched.
int fyzz_hip_world_set_body_velocity(fyzz_hip_world* w, uint32_t body_id, const float* linear,
                                     const float* angular) {
    if (!w) return kInvalidArgument;
    return guarded(w->ctx, [&] {
        FYZZ_REQUIRE(!(w->has_floor && body_id == w->floor_id), kLogicError, "world: static actors have no velocity");
        for (int k = 0; k < 3; ++k) {
            FYZZ_REQUIRE(!linear || std::isfinite(linear[k]), kInvalidArgument, "world: velocity must be finite");
            FYZZ_REQUIRE(!angular || std::isfinite(angular[k]), kInvalidArgument,
                         "world: angular velocity must be finite");
        }
        fyzz_dy_body tmp;
        u32 slot = 0;
        fyzz_dy_body& b = edit_body(w, body_id, tmp, &slot);
        for (int k = 0; k < 3; ++k) {
            if (linear) b.linear_velocity[k] = linear[k];
            if (angular) b.angular_velocity[k] = angular[k];
        }
        if (linear || angular) world::wake_record(b);
        commit_body(w, slot, b);
    });
}

// Scene::set_material -> set_attached_shape_material (scene.cpp:1022,
// 1099-1118): finite, non-negative frictions, restitution at most 1. The
// next step's contact prep reads the new values.
int fyzz_hip_world_set_body_material(fyzz_hip_world* w, uint32_t body_id, float static_friction,
                                     float dynamic_friction, float restitution) {
    if (!w) return kInvalidArgument;
    return guarded(w->ctx, [&] {
        FYZZ_REQUIRE(std::isfinite(static_friction) && std::isfinite(dynamic_friction) && std::isfinite(restitution) &&
                         static_friction >= 0.0f && dynamic_friction >= 0.0f && restitution >= 0.0f &&
                         restitution <= 1.0f,
                     kInvalidArgument, "world: material values must be finite, non-negative, restitution <= 1");
        if (w->has_floor && body_id == w->floor_id) {
            w->floor.static_friction = static_friction;
            w->floor.dynamic_friction = dynamic_friction;
            w->floor.restitution = restitution;
            if (w->uploaded) {
                hipStream_t st = w->ctx->stream;
                FYZZ_HIP_CHECK(hipMemcpyAsync(w->statics.p, &w->floor, sizeof(fyzz_dy_static), hipMemcpyHostToDevice,
                                              st));
                FYZZ_HIP_CHECK(hipStreamSynchronize(st));
            }
            return;
        }
        if (body_id < w->slot_of_id.size() && w->slot_of_id[body_id] <= -2) {
            auto& S = w->host_statics[static_cast<u32>(-2 - w->slot_of_id[body_id])];
            S.static_friction = static_friction;
            S.dynamic_friction = dynamic_friction;
            S.restitution = restitution;
            if (w->uploaded) upload_statics(w);
            return;
        }
        fyzz_dy_body tmp;
        u32 slot = 0;
        fyzz_dy_body& b = edit_body(w, body_id, tmp, &slot);
        b.static_friction = static_friction;
        b.dynamic_friction = dynamic_friction;
        b.restitution = restitution;
        commit_body(w, slot, b);
    });
}

// Scene::set_dynamics (scene.cpp:1119-1147): the reference's validation,
// then the fields the admitted class steps with.
int fyzz_hip_world_set_body_dynamics(fyzz_hip_world* w, uint32_t body_id, const fyzz_body_dynamics* d) {
    if (!w) return kInvalidArgument;
    return guarded(w->ctx, [&] {
        FYZZ_REQUIRE(d, kInvalidArgument, "world: null dynamics");
        FYZZ_REQUIRE(!(w->has_floor && body_id == w->floor_id), kLogicError,
                     "world: static actors have no dynamic solver state");
        const bool valid = std::isfinite(d->linear_damping) && std::isfinite(d->angular_damping) &&
                           std::isfinite(d->max_linear_velocity) && std::isfinite(d->max_angular_velocity) &&
                           std::isfinite(d->max_depenetration_velocity) && std::isfinite(d->max_contact_impulse) &&
                           std::isfinite(d->contact_slop_coefficient) && d->max_depenetration_velocity > 0.0f &&
                           d->max_contact_impulse >= 0.0f && d->contact_slop_coefficient >= 0.0f &&
                           d->linear_damping >= 0.0f && d->angular_damping >= 0.0f &&
                           d->max_linear_velocity >= 0.0f && d->max_angular_velocity >= 0.0f &&
                           std::isfinite(d->max_linear_velocity * d->max_linear_velocity) &&
                           std::isfinite(d->max_angular_velocity * d->max_angular_velocity) &&
                           d->position_iterations != 0 && (d->lock_flags & ~0x3fu) == 0;
        FYZZ_REQUIRE(valid, kInvalidArgument, "world: rigid body dynamics invalid");
        // outside the admitted class: the kernels step without these terms
        FYZZ_REQUIRE(d->max_depenetration_velocity == 1.0e32f && d->max_contact_impulse == 3.40282347e38f &&
                         d->contact_slop_coefficient == 0.0f && d->lock_flags == 0 && !d->gyroscopic_forces_enabled &&
                         !d->retain_accelerations,
                     kLogicError,
                     "world: depenetration/impulse caps, slop, locks, gyroscopic forces and retained accelerations "
                     "are outside the admitted class");
        fyzz_dy_body tmp;
        u32 slot = 0;
        fyzz_dy_body& b = edit_body(w, body_id, tmp, &slot);
        b.linear_damping = d->linear_damping;
        b.angular_damping = d->angular_damping;
        b.max_linear_velocity = d->max_linear_velocity;
        b.max_angular_velocity = d->max_angular_velocity;
        b.flags = d->gravity_enabled ? (b.flags | 4u) : (b.flags & ~4u);
        b.iterations = static_cast<u32>(d->position_iterations) | (static_cast<u32>(d->velocity_iterations) << 8);
        commit_body(w, slot, b);
    });
}

int fyzz_hip_world_step(fyzz_hip_world* w, float dt) {
    if (!w) return kInvalidArgument;
    return guarded(w->ctx, [&] {
        FYZZ_REQUIRE(std::isfinite(dt) && dt > 0.0f, kInvalidArgument, "world: dt must be positive");
        step(w, dt);
    });
}

int fyzz_hip_world_body_records(fyzz_hip_world* w, fyzz_dy_body* out, uint32_t capacity, uint32_t* count) {
    if (!w) return kInvalidArgument;
    return guarded(w->ctx, [&] {
        const u32 n = static_cast<u32>(w->host_bodies.size());
        if (count) *count = n;
        const u32 m = std::min(n, capacity);
        if (!m) return;
        FYZZ_REQUIRE(out, kInvalidArgument, "world: null output");
        if (!w->uploaded) {
            std::copy(w->host_bodies.begin(), w->host_bodies.begin() + m, out);
            return;
        }
        FYZZ_HIP_CHECK(hipMemcpyAsync(out, w->bodies.p, sizeof(fyzz_dy_body) * m, hipMemcpyDeviceToHost,
                                      w->ctx->stream));
        FYZZ_HIP_CHECK(hipStreamSynchronize(w->ctx->stream));
    });
}

int fyzz_hip_world_stats(fyzz_hip_world* w, uint32_t* n_pairs, uint32_t* n_managers, uint32_t* n_partitions,
                         uint32_t* status_flags) {
    if (!w) return kInvalidArgument;
    return guarded(w->ctx, [&] {
        u32 np = 0, nm = 0, parts = 0, dflags = 0, rounds = 0;
        if (w->uploaded && w->n) {
            u32 err = 0, ovf = 0;
            sap::Lists L{};
            if (w->n_scenes) {
                L.d_cur = w->scene_n.p + w->scur;
                L.d_error = w->scene_status.p;
                L.d_overflow = w->scene_status.p;
            } else {
                L = sap::lists(w->sap);
            }
            hipStream_t st = w->ctx->stream;
            FYZZ_HIP_CHECK(hipMemcpyAsync(&np, L.d_cur, 4, hipMemcpyDeviceToHost, st));
            FYZZ_HIP_CHECK(hipMemcpyAsync(&err, L.d_error, 4, hipMemcpyDeviceToHost, st));
            FYZZ_HIP_CHECK(hipMemcpyAsync(&ovf, L.d_overflow, 4, hipMemcpyDeviceToHost, st));
            FYZZ_HIP_CHECK(hipMemcpyAsync(&nm, w->d_nm.p, 4, hipMemcpyDeviceToHost, st));
            (void)fyzz_hip_dy_status(w->dy, &dflags, &parts, &rounds);
            dflags |= (err ? 0x100u : 0u) | (ovf ? 0x200u : 0u);
        }
        if (n_pairs) *n_pairs = np;
        if (n_managers) *n_managers = nm;
        if (n_partitions) *n_partitions = parts;
        if (status_flags) *status_flags = dflags;
    });
}

int fyzz_hip_world_fetch_managers(fyzz_hip_world* w, fyzz_dy_manager* managers, float* impulses,
                                  uint32_t* partitions, uint32_t capacity, uint32_t* count) {
    if (!w) return kInvalidArgument;
    return guarded(w->ctx, [&] {
        u32 nm = 0;
        hipStream_t st = w->ctx->stream;
        if (w->uploaded && w->n) {
            FYZZ_HIP_CHECK(hipMemcpyAsync(&nm, w->d_nm.p, 4, hipMemcpyDeviceToHost, st));
            FYZZ_HIP_CHECK(hipStreamSynchronize(st));
        }
        if (count) *count = nm;
        const u32 m = std::min(nm, capacity);
        if (!m) return;
        if (managers)
            FYZZ_HIP_CHECK(hipMemcpyAsync(managers, w->managers.p, sizeof(fyzz_dy_manager) * m, hipMemcpyDeviceToHost,
                                          st));
        if (impulses)
            FYZZ_HIP_CHECK(hipMemcpyAsync(impulses, w->impulses.p, sizeof(float) * 4 * m, hipMemcpyDeviceToHost, st));
        if (partitions)
            FYZZ_HIP_CHECK(hipMemcpyAsync(partitions, w->parts.p, sizeof(u32) * m, hipMemcpyDeviceToHost, st));
        FYZZ_HIP_CHECK(hipStreamSynchronize(st));
    });
}

int fyzz_hip_world_friction_anchors(fyzz_hip_world* w, fyzz_friction_anchors* out, uint32_t capacity,
                                    uint32_t* count) {
    if (!w) return kInvalidArgument;
    return guarded(w->ctx, [&] {
        u32 nm = 0;
        hipStream_t st = w->ctx->stream;
        if (w->uploaded && w->n) {
            FYZZ_HIP_CHECK(hipMemcpyAsync(&nm, w->d_nm.p, 4, hipMemcpyDeviceToHost, st));
            FYZZ_HIP_CHECK(hipStreamSynchronize(st));
        }
        if (count) *count = nm;
        const u32 m = std::min(nm, capacity);
        if (!m) return;
        FYZZ_REQUIRE(out, kInvalidArgument, "world: null anchor output");
        w->anchors.reserve(w->mgr_cap);  // first use only: outside any step
        const int rc = fyzz_hip_dy_friction_anchors(w->dy, w->anchors.p);
        if (rc) throw Error(rc, w->ctx->last_error);
        FYZZ_HIP_CHECK(hipMemcpyAsync(out, w->anchors.p, sizeof(fyzz_friction_anchors) * m, hipMemcpyDeviceToHost,
                                      st));
        FYZZ_HIP_CHECK(hipStreamSynchronize(st));
    });
}

int fyzz_hip_world_totals(fyzz_hip_world* w, uint64_t* out, int reset) {
    if (!w) return kInvalidArgument;
    return guarded(w->ctx, [&] {
        FYZZ_REQUIRE(out, kInvalidArgument, "world: null output");
        for (u32 k = 0; k < world::kTotals; ++k) out[k] = 0;
        if (!w->uploaded) return;
        hipStream_t st = w->ctx->stream;
        FYZZ_HIP_CHECK(hipMemcpyAsync(out, w->totals.p, world::kTotals * sizeof(u64), hipMemcpyDeviceToHost, st));
        FYZZ_HIP_CHECK(hipStreamSynchronize(st));
        if (reset) FYZZ_HIP_CHECK(hipMemsetAsync(w->totals.p, 0, world::kTotals * sizeof(u64), st));
    });
}

int fyzz_hip_world_set_contact_notifications(fyzz_hip_world* w, int enabled) {
    if (!w) return kInvalidArgument;
    return guarded(w->ctx, [&] {
        FYZZ_REQUIRE(!(enabled && w->n_scenes), kLogicError,
                     "world: contact notifications are published for single-scene worlds");
        const bool on = enabled != 0;
        if (on == w->notify) return;
        w->notify = on;
        if (!w->uploaded) return;
        // Captured steps keep the publication state they were recorded
        // with; the buffers are reserved here, never inside a step.
        FYZZ_HIP_CHECK(hipStreamSynchronize(w->ctx->stream));
        drop_graphs(w);
        // touch flags are kept every step, so the first published step
        // still tells found from persists
        if (on) reserve_events(w);
    });
}

int fyzz_hip_world_set_solver_order(fyzz_hip_world* w, int native) {
    if (!w) return kInvalidArgument;
    return guarded(w->ctx, [&] {
        w->require_authoring();
        FYZZ_REQUIRE(native == 0 || native == 1, kInvalidArgument, "world: solver order is 0 (device) or 1 (native)");
        FYZZ_REQUIRE(native == 0 || w->n_scenes == 0, kLogicError,
                     "world: the native solver order is for single-scene worlds");
        w->native_order = native == 1;
    });
}

int fyzz_hip_world_set_scenes(fyzz_hip_world* w, uint32_t n_scenes, const uint32_t* box_counts) {
    if (!w) return kInvalidArgument;
    return guarded(w->ctx, [&] {
        w->require_authoring();
        FYZZ_REQUIRE(n_scenes == 0 || box_counts, kInvalidArgument, "world: null scene box counts");
        FYZZ_REQUIRE(n_scenes == 0 || !w->native_order, kLogicError,
                     "world: the native solver order is for single-scene worlds");
        FYZZ_REQUIRE(n_scenes == 0 || w->host_statics.empty(), kLogicError,
                     "world: batched scenes hold no static bodies but the floor");
        FYZZ_REQUIRE(!w->notify, kLogicError, "world: contact notifications are published for single-scene worlds");
        w->scene_counts.assign(box_counts, box_counts + n_scenes);
        w->n_scenes = n_scenes;
    });
}

int fyzz_hip_world_contact_events(fyzz_hip_world* w, fyzz_contact_event* events, uint32_t capacity,
                                  uint32_t* n_events, fyzz_contact_event_point* points, uint32_t point_capacity,
                                  uint32_t* n_points) {
    if (!w) return kInvalidArgument;
    return guarded(w->ctx, [&] {
        u32 counts[4] = {0, 0, 0, 0}, np = 0;
        hipStream_t st = w->ctx->stream;
        if (w->notify && w->ev_cap && w->uploaded) {
            FYZZ_HIP_CHECK(hipMemcpyAsync(counts, w->ev_counts.p, sizeof counts, hipMemcpyDeviceToHost, st));
            FYZZ_HIP_CHECK(hipMemcpyAsync(&np, w->d_npts.p, 4, hipMemcpyDeviceToHost, st));
            FYZZ_HIP_CHECK(hipStreamSynchronize(st));
        }
        FYZZ_REQUIRE(np <= w->pt_cap || !w->notify, kOverflow, "world: contact event points exceed capacity");
        if (n_events) *n_events = counts[3];
        if (n_points) *n_points = np;
        const u32 ne = std::min(counts[3], capacity), mp = std::min(np, point_capacity);
        if (events && ne)
            FYZZ_HIP_CHECK(hipMemcpyAsync(events, w->events.p, sizeof(fyzz_contact_event) * ne, hipMemcpyDeviceToHost,
                                          st));
        if (points && mp)
            FYZZ_HIP_CHECK(hipMemcpyAsync(points, w->points.p, sizeof(fyzz_contact_event_point) * mp,
                                          hipMemcpyDeviceToHost, st));
        FYZZ_HIP_CHECK(hipStreamSynchronize(st));
    });
}

int fyzz_hip_world_set_materials(fyzz_hip_world* w, uint32_t first_id, uint32_t n, const float* static_friction,
                                 const float* dynamic_friction, const float* restitution) {
    if (!w) return kInvalidArgument;
    return guarded(w->ctx, [&] {
        w->require_authoring();  // bulk form: before the first step (set_body_material works at any time)
        FYZZ_REQUIRE(n == 0 || (static_friction && dynamic_friction && restitution), kInvalidArgument,
                     "world: null material arrays");
        for (u32 k = 0; k < n; ++k) {
            FYZZ_REQUIRE(static_friction[k] >= 0.0f && dynamic_friction[k] >= 0.0f && restitution[k] >= 0.0f,
                         kInvalidArgument, "world: material values must be non-negative");
            fyzz_dy_body& b = w->body(first_id + k);
            b.static_friction = static_friction[k];
            b.dynamic_friction = dynamic_friction[k];
            b.restitution = restitution[k];
        }
    });
}

int fyzz_hip_world_record_dy_inputs(fyzz_hip_world* w, fyzz_dy_manager* managers, fyzz_dy_patch* history,
                                    uint32_t capacity, uint32_t* n_managers) {
    if (!w) return kInvalidArgument;
    return guarded(w->ctx, [&] {
        FYZZ_REQUIRE(w->uploaded, kLogicError, "world: no step taken");
        hipStream_t st = w->ctx->stream;
        u32 nm = 0;
        if (w->n) {
            FYZZ_HIP_CHECK(hipMemcpyAsync(&nm, w->d_nm.p, 4, hipMemcpyDeviceToHost, st));
            FYZZ_HIP_CHECK(hipStreamSynchronize(st));
        }
        if (n_managers) *n_managers = nm;
        if (!managers || !history || !nm) return;
        FYZZ_REQUIRE(capacity >= nm, kOverflow, "world: record capacity below the manager count");
        hipLaunchKernelGGL(world::k_record_dy, dim3(std::min<u32>(grid_for(nm, 256), 2048u)), dim3(256), 0, st,
                           w->d_nm.p, w->mgr_cap, w->managers.p, w->hist[w->hcur ^ 1].p, managers, history);
        FYZZ_HIP_CHECK(hipGetLastError());
    });
}

int fyzz_hip_world_append_dy_inputs(fyzz_hip_world* w, fyzz_dy_manager* managers, fyzz_dy_patch* history,
                                    uint64_t capacity, uint64_t* cursor, uint32_t* count) {
    if (!w) return kInvalidArgument;
    return guarded(w->ctx, [&] {
        FYZZ_REQUIRE(w->uploaded, kLogicError, "world: no step taken");
        FYZZ_REQUIRE(cursor && count && (capacity == 0 || (managers && history)), kInvalidArgument,
                     "world: null arena, cursor or count");
        hipStream_t st = w->ctx->stream;
        const u32 grid = std::min<u32>(grid_for(std::max<u32>(w->mgr_cap, 1u), 256), 2048u);
        if (w->n)
            hipLaunchKernelGGL(world::k_append_dy, dim3(grid), dim3(256), 0, st, w->d_nm.p, w->mgr_cap,
                               w->managers.p, w->hist[w->hcur ^ 1].p, managers, history, capacity, cursor);
        hipLaunchKernelGGL(world::k_append_advance, dim3(1), dim3(1), 0, st, w->n ? w->d_nm.p : w->d_total.p,
                           w->n ? w->mgr_cap : 0u, capacity, cursor, count);
        FYZZ_HIP_CHECK(hipGetLastError());
    });
}

int fyzz_hip_world_copy_bodies(fyzz_hip_world* w, fyzz_dy_body* dst) {
    if (!w) return kInvalidArgument;
    return guarded(w->ctx, [&] {
        if (!w->uploaded) upload(w);
        FYZZ_REQUIRE(dst || !w->n, kInvalidArgument, "world: null destination");
        if (w->n)
            FYZZ_HIP_CHECK(hipMemcpyAsync(dst, w->bodies.p, sizeof(fyzz_dy_body) * w->n, hipMemcpyDeviceToDevice,
                                          w->ctx->stream));
    });
}

int fyzz_hip_world_broadphase_reports(fyzz_hip_world* w, int which, uint32_t* out_pairs, uint64_t capacity,
                                      uint64_t* written) {
    if (!w) return kInvalidArgument;
    return guarded(w->ctx, [&] {
        FYZZ_REQUIRE(which >= 0 && which <= 2, kInvalidArgument, "world: which must be 0, 1 or 2");
        FYZZ_REQUIRE(which == 0 || !w->n_scenes, kLogicError,
                     "world: found/lost reports belong to the incremental broadphase (no scenes)");
        if (written) *written = 0;
        if (!w->uploaded || !w->n) return;
        if (!w->n_scenes) {
            const int rc = fyzz_hip_sap_fetch(w->sap, which, out_pairs, capacity, written);
            if (rc) throw Error(rc, w->ctx->last_error);
            return;
        }
        hipStream_t st = w->ctx->stream;
        const sap::Lists L = current_lists(w);
        u32 np = 0;
        FYZZ_HIP_CHECK(hipMemcpyAsync(&np, L.d_cur, 4, hipMemcpyDeviceToHost, st));
        FYZZ_HIP_CHECK(hipStreamSynchronize(st));
        np = std::min(np, w->pair_cap);
        if (written) *written = np;
        const u64 m = std::min<u64>(np, capacity);
        if (!m) return;
        FYZZ_REQUIRE(out_pairs, kInvalidArgument, "world: null output");
        std::vector<u64> keys(m);
        FYZZ_HIP_CHECK(hipMemcpyAsync(keys.data(), L.cur, 8 * m, hipMemcpyDeviceToHost, st));
        FYZZ_HIP_CHECK(hipStreamSynchronize(st));
        for (u64 i = 0; i < m; ++i) {
            out_pairs[2 * i] = static_cast<u32>(keys[i] >> 32);
            out_pairs[2 * i + 1] = static_cast<u32>(keys[i] & 0xffffffffu);
        }
    });
}

int fyzz_hip_world_fetch_touch(fyzz_hip_world* w, uint32_t* touch_bb, uint32_t pair_capacity, uint32_t* touch_pl) {
    if (!w) return kInvalidArgument;
    return guarded(w->ctx, [&] {
        if (!w->uploaded || !w->n) return;
        hipStream_t st = w->ctx->stream;
        const sap::Lists L = current_lists(w);
        u32 np = 0;
        FYZZ_HIP_CHECK(hipMemcpyAsync(&np, L.d_cur, 4, hipMemcpyDeviceToHost, st));
        FYZZ_HIP_CHECK(hipStreamSynchronize(st));
        np = std::min(np, w->pair_cap);
        FYZZ_REQUIRE(!touch_bb || pair_capacity >= np, kOverflow, "world: pair capacity below the pair count");
        if (touch_bb && np)
            FYZZ_HIP_CHECK(hipMemcpyAsync(touch_bb, w->touch_bb[w->ccur].p, 4ull * np, hipMemcpyDeviceToHost, st));
        if (touch_pl) FYZZ_HIP_CHECK(hipMemcpyAsync(touch_pl, w->touch_pl.p, 4ull * w->n, hipMemcpyDeviceToHost, st));
        FYZZ_HIP_CHECK(hipStreamSynchronize(st));
    });
}

int fyzz_hip_world_fetch_persistent(fyzz_hip_world* w, void* caches_bb, uint32_t pair_capacity, uint32_t* n_pairs,
                                    void* caches_pl, fyzz_dy_patch* history, uint32_t* stamps, uint32_t* step) {
    if (!w) return kInvalidArgument;
    return guarded(w->ctx, [&] {
        if (n_pairs) *n_pairs = 0;
        if (step) *step = 0;
        if (!w->uploaded || !w->n) return;
        hipStream_t st = w->ctx->stream;
        const sap::Lists L = current_lists(w);
        u32 np = 0, dstep = 0;
        FYZZ_HIP_CHECK(hipMemcpyAsync(&np, L.d_cur, 4, hipMemcpyDeviceToHost, st));
        FYZZ_HIP_CHECK(hipMemcpyAsync(&dstep, w->d_step.p, 4, hipMemcpyDeviceToHost, st));
        FYZZ_HIP_CHECK(hipStreamSynchronize(st));
        np = std::min(np, w->pair_cap);
        if (n_pairs) *n_pairs = np;
        if (step) *step = dstep;
        FYZZ_REQUIRE(!(caches_bb || history || stamps) || pair_capacity >= np, kOverflow,
                     "world: pair capacity below the pair count");
        const u32 n = w->n, cap = w->pair_cap;
        if (caches_bb && np)
            FYZZ_HIP_CHECK(hipMemcpyAsync(caches_bb, w->cache_bb[w->ccur].p, sizeof(world::CacheRec) * np,
                                          hipMemcpyDeviceToHost, st));
        if (caches_pl)
            FYZZ_HIP_CHECK(hipMemcpyAsync(caches_pl, w->cache_pl.p, sizeof(world::CacheRec) * n,
                                          hipMemcpyDeviceToHost, st));
        const int h = w->hcur;
        if (history) {
            if (np)
                FYZZ_HIP_CHECK(hipMemcpyAsync(history, w->hist[h].p, sizeof(fyzz_dy_patch) * np,
                                              hipMemcpyDeviceToHost, st));
            FYZZ_HIP_CHECK(hipMemcpyAsync(history + np, w->hist[h].p + cap, sizeof(fyzz_dy_patch) * n,
                                          hipMemcpyDeviceToHost, st));
        }
        if (stamps) {
            if (np) FYZZ_HIP_CHECK(hipMemcpyAsync(stamps, w->stamp[h].p, 4ull * np, hipMemcpyDeviceToHost, st));
            FYZZ_HIP_CHECK(hipMemcpyAsync(stamps + np, w->stamp[h].p + cap, 4ull * n, hipMemcpyDeviceToHost, st));
        }
        FYZZ_HIP_CHECK(hipStreamSynchronize(st));
    });
}

// Device pointers of the step's manager list and Dy outputs (valid until
// the next step): managers[*n_managers], impulses[.][4], partitions[.],
// and the body records.
int fyzz_hip_world_device_views(fyzz_hip_world* w, const fyzz_dy_manager** managers, const uint32_t** d_n_managers,
                                const float** impulses, const uint32_t** partitions, fyzz_dy_body** bodies) {
    if (!w) return kInvalidArgument;
    return guarded(w->ctx, [&] {
        FYZZ_REQUIRE(w->uploaded, kLogicError, "world: no step taken");
        if (managers) *managers = w->managers.p;
        if (d_n_managers) *d_n_managers = w->d_nm.p;
        if (impulses) *impulses = w->impulses.p;
        if (partitions) *partitions = w->parts.p;
        if (bodies) *bodies = w->bodies.p;
    });
}

}  // extern "C"
