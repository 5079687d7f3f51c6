// Internal hooks between the stage sources and the step pipeline: device
// pair lists of the broadphase, list-driven narrowphase launches and a Dy
// step whose manager count stays on the device. Nothing here crosses the
// C-ABI.
#pragma once
#include "../../include/fyzz_hip.h"
#include "common.hpp"
#include "context.hpp"

namespace fyzz_hip {

namespace sap {
// Current and previous sorted pair lists (u64 lo slot << 32 | hi slot) of a
// broadphase after its last advance, with their device counts.
struct Lists {
    const u64* cur;
    const u32* d_cur;
    const u64* prev;
    const u32* d_prev;
    u32 cap;
    const u32* d_error;     // non-zero: a box bound was not finite
    const u32* d_overflow;  // non-zero: pair or delta capacity exceeded
    const u64* lost;        // pairs that left the list this frame (sorted)
    const u32* d_lost;
    const u32* d_found;     // pairs that entered the list this frame (count; may be null)
    const u32* src = nullptr;  // previous-list index of each current pair (kNone: new), when known
};
Lists lists(fyzz_hip_sap* h);
// Whether a frame has been advanced (the next one is incremental).
bool built(fyzz_hip_sap* h);
// The host-side bookkeeping of one incremental advance (frame stamp,
// current buffer), for a replayed graph that recorded such an advance.
void replay_advance(fyzz_hip_sap* h);
// Looks up this device's copy of the fixed sort workspace (call with the
// device current; idempotent).
void bind_fixed_workspace(int device);
// Per-slot sphere flags (device, n entries; null: all boxes): a sphere's
// bounds are its radius half[3 slot] on every axis, whatever its
// orientation. The pointer must stay valid while the broadphase advances.
void set_spheres(fyzz_hip_sap* h, const uint8_t* sphere);
// Stable LSD radix sort of (u64 key, u32 value) pairs in place: the low
// `bits` key bits (a multiple of 16, so the result lands back in `keys`),
// *d_count (device) entries of at most cap; scratch of
// sort_keys_scratch_words(cap) words.
std::size_t sort_keys_scratch_words(u32 cap);
void sort_keys(hipStream_t st, u64* keys, u32* vals, u64* tmp_keys, u32* tmp_vals, u32 cap, const u32* d_count,
               int bits, u32* scratch);
}  // namespace sap

namespace pcm {
constexpr u32 kNoSource = 0xffffffffu;
// Box/box pairs of a device list: pair t (t < min(*d_n, cap)) reads its
// cache from cache_src[src_index[t]] (a zero cache when kNoSource) and
// writes it to cache_dst[t].
struct ListPairs {
    u32 cap;
    const u32* d_n;
    const u32* first;
    const u32* second;
    const void* cache_src;
    const u32* src_index;
    void* cache_dst;
};
// sphere: per slot, non-zero for a sphere of radius half[3 slot] (its
// pairs take the stateless sphere contacts), or null when all are boxes.
void box_box_list(fyzz_hip_context* ctx, const ListPairs& L, const float* pos, const float* quat,
                  const float* half, float cd, float* contacts, u32* counts, u32* flags, u32* scratch,
                  const uint8_t* sphere = nullptr);
// Every box slot t < n against one plane; caches in place, indexed by slot.
void plane_all(fyzz_hip_context* ctx, u32 n, const float* pos, const float* quat, const float* half,
               const float plane_position[3], float cd, void* cache, float* contacts, u32* counts, u32* flags,
               const uint8_t* sphere = nullptr);
}  // namespace pcm

namespace dy {
// Device counters of the last Dy step: [4] partitions, [5] position and
// [6] velocity iterations, [7] partition rounds, [8] contact rows, [9]
// friction rows.
const u32* counters(fyzz_hip_dy* d);
enum : u32 { kCtrParts = 4, kCtrPos = 5, kCtrVel = 6, kCtrRounds = 7, kCtrContacts = 8, kCtrFriction = 9 };
// Launch the grid-barrier kernels as plain kernels (for graph capture).
void set_plain_launch(fyzz_hip_dy* d, bool plain);
// Workgroups of the partition and PGS launches (clamped to residency).
void set_coop_blocks(fyzz_hip_dy* d, int partition_blocks, int pgs_blocks);
void coop_blocks_of(fyzz_hip_dy* d, int* partition_blocks, int* pgs_blocks);
// the counts a Dy workspace on `device` would take for the requested ones
// (0: the default), clamped as set_coop_blocks clamps
void effective_blocks(int device, int partition_blocks, int pgs_blocks, int* partition_out, int* pgs_out);
// fyzz_hip_dy_step with the manager count read from d_nm (at most nm_cap).
void step_device(fyzz_hip_dy* d, float dt, u32 nb, fyzz_dy_body* bodies, u32 ns, const fyzz_dy_static* statics,
                 u32 nm_cap, const u32* d_nm, const fyzz_dy_manager* managers, const fyzz_dy_patch* history,
                 u32 nh, fyzz_dy_patch* patches, float* impulses, u32* partitions);
}  // namespace dy

}  // namespace fyzz_hip
