// Dy stage of the GPU step on gfx950: unconstrained pre-integration, solver
// rows, exact first-fit conflict partitions, contact preparation with
// correlated friction and anchor growth, partitioned projected Gauss-Seidel
// (normal rows, then the patch's friction rows under the Coulomb clamp),
// integration and the wake-counter sleep check.
//
// Reference (src/native/scene.cpp):
//   prepare_unconstrained_velocities            10383-10507   k_pre
//   prepare_solver_bodies                       6940-7112     k_rows
//   partition_constraints                       7113-7489     k_part_*
//   prepare_constraints (contact managers)      7490-7515, 7763-8511   k_prep
//   cached_friction_patch_valid                 3981-4010     history_valid
//   solve_constraint_task                       8853-9148     k_pgs
//   publish_solver_bodies, integrate_solver_task, integrate_quaternion
//                                               9509-9549, 10508-10707, 111-146   k_integrate
//
// Admitted scene class: dynamic boxes and static planes/boxes, PCM
// manifolds (one normal per manager, 1-4 contacts), non-compliant
// materials, no joints, locks, slop or gyroscopic terms; one solver task.
//
// Exactness. Every FP32 expression keeps the reference's operation order
// (the library builds with -ffp-contract=off). Partitions reproduce the
// sequential first-fit over 32-bit masks exactly: a manager's partition
// depends only on the masks its two bodies carry after the earlier managers
// incident to them, so a manager is coloured as soon as its predecessor on
// each body is (one cooperative launch, no global rounds). Within a
// partition no two managers share a
// dynamic body and the static world's solver body keeps a zero state, so
// solving a partition's managers concurrently gives the sequential sweep's
// values bit for bit.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <memory>
#include <string>

#include "../../include/fyzz_hip.h"
#include "context.hpp"
#include "select.hpp"
#include "stages.hpp"
#include "vecmath.hpp"

namespace fyzz_hip::dy {
using namespace fyzz_hip::vm;

constexpr u32 kNone = 0xffffffffu;
constexpr u32 kPending = 0xfffffffeu;
constexpr u32 kMaxParts = 4096;       // partitions per step
constexpr u32 kMaxPending = 65536;    // managers left for overflow waves
constexpr int kBlock = 256;
constexpr int kCoopBlocks = 64;       // workgroups of the cooperative launches
constexpr int kPgsBlocks = 256;       // workgroups of the sweep (no grid barrier there)

// device counters
enum : u32 { kFront0 = 0, kFront1, kFront2, kPend, kNParts, kPosIters, kVelIters, kRounds, kContacts, kFrictionRows,
             kNbWord, kCtrWords = 16 };

// one contact row (SolverContactPoint fields the solve reads), 64 B
struct alignas(16) CRow {
    float nx, ny, nz, vm;
    float rax, ray, raz, biased;
    float rbx, rby, rbz, unbiased;
    float applied, max_impulse, pad0, pad1;
};
// one friction row (SolverContactFriction), 64 B
struct alignas(16) FRow {
    float tx, ty, tz, vm;
    float rax, ray, raz, bias;
    float rbx, rby, rbz, target;
    float applied, pad0, pad1, pad2;
};

struct Ws {
    // step inputs
    float dt;
    u32 nb, ns, nm, nh;
    fyzz_dy_body* bodies;
    const fyzz_dy_static* statics;
    const fyzz_dy_manager* mg;
    const fyzz_dy_patch* history;
    fyzz_dy_patch* patches;
    float* impulses;
    u32* parts_out;
    const u32* d_nm;  // device manager count (at most nm), or null
    u32 with_rows;    // contact prep ran: order, rows and patches are valid
    // per body
    float4 *pv, *pw;          // pre-integrated velocities
    float4 *lin, *ang;        // solver deltas: lin_delta|inv_mass, ang_state
    float4 *mlin, *mang;      // motion snapshot after the last position iteration
    float4 *rpos, *rq;        // mass pose of the solver row
    float4* sii;              // 3 columns of the world sqrt inverse inertia
    u32* has_row;
    u32 *deg, *off, *fill, *max_dyn, *body_mask;
    // per manager
    u32 *inc, *pred, *part, *pend, *sorted, *order;
    uint8_t* dd;  // [nm] the manager touches two dynamic bodies (a compact copy for the per-body passes)
    u64* entry;  // [2 nm] partition colouring: ready | level << 32 | mask after the manager, per side
    int4* sinfo; // [nm] per partition-sorted position: body0, body1, manager, contact rows | anchors << 8
    u32* pos_of;    // [nm] partition-sorted position of each manager in a partition
    uint4* srank;   // [nm] per position: rank of the manager among body0's / body1's (by partition), their degrees
    u32* erank;     // [2 nm] per incidence entry: rank among the body's managers by partition
    u64* vst;       // [8 nb] sweep state, a 64-byte line per body: (component bits, updates applied) x 6, pad
    CRow* crow;
    FRow* frow;
    u32 *part_count, *part_off, *part_fill;
    u32* blocksums;
    u32* ctr;
    u32* flags;
    u32* bar;  // grid barrier: arrival count, generation
    // adjoint support (tape mode)
    u32 tape;             // record branches and raw row vectors
    u32 tape_iters;       // iterations the tape holds
    uint8_t* tape_c;      // [iter][4 nm] normal rows: 1 free update, 0 clamped to zero, 2 at the maximum
    uint8_t* tape_f;      // [iter][4 nm] friction rows: 0 free, 1 clamped inside, 2 at +max, 3 at -max
    float* tape_ni;       // [iter][nm] the manager's normal impulse sum of the iteration
    float4* craw;         // [4 nm][2] contact rows: (ra x n, kv), (rb x n, 0); d err / d vrel = kv vm
    float4* fraw;         // [4 nm][2] friction rows: (ra x t, 0), (rb x t, 0) unscaled
    // adjoint inputs / outputs and state
    const float* adj_v_out;
    const float* adj_w_out;
    const float* adj_p_out;
    float* adj_v_in;
    float* adj_w_in;
    float* adj_friction;  // per body: d loss / d dynamic friction
    float4 *a_lin, *a_ang;
    float4* a_side;  // [4 nm] per manager and side: (velocity adjoint, friction adjoint), (angular adjoint, 0)
    float *a_capp, *a_fapp, *a_err, *a_tgt, *a_df;
};

__device__ inline u32 ld(const u32* p) { return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }
__device__ inline void raise_flag(const Ws& W, u32 f) { atomicOr(W.flags, f); }
__device__ inline u32 mcount(const Ws& W) { return W.d_nm ? min(ld(W.d_nm), W.nm) : W.nm; }
// Grid-wide barrier of a cooperative launch (every workgroup resident),
// two levels: workgroups arrive on their group's counter (groups of
// kBarGroup), the last of each group arrives on the root counter, and the
// last root arrival publishes the next generation. bar[0] root count,
// bar[1] generation, bar[2 + g] group counts. Relaxed agent-scope atomics;
// every workgroup releases its writes (agent scope) before arriving, the
// releasing workgroup acquires before publishing, and every workgroup
// acquires after leaving, so the grid's global-memory writes before the
// barrier are visible after it. Thread 0 of each workgroup keeps the
// generation it waits on. The spin is bounded; exceeding it (workgroups
// not co-resident) raises a flag.
constexpr u32 kSpinLimit = 1u << 20;
constexpr u32 kBarGroup = 8;
constexpr u32 kBarWords = 64;  // root, generation, up to 62 groups
__device__ inline u32 barrier_start(const Ws& W) {
    return threadIdx.x == 0 ? __hip_atomic_load(&W.bar[1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : 0u;
}
__device__ inline void grid_sync(const Ws& W, u32& gen) {
    __syncthreads();
    if (threadIdx.x == 0) {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        const u32 g0 = gen;
        const u32 grp = blockIdx.x / kBarGroup;
        const u32 ngroups = (gridDim.x + kBarGroup - 1) / kBarGroup;
        const u32 gsize = min(kBarGroup, gridDim.x - grp * kBarGroup);
        bool last = false;
        if (__hip_atomic_fetch_add(&W.bar[2 + grp], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == gsize - 1) {
            __hip_atomic_store(&W.bar[2 + grp], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (__hip_atomic_fetch_add(&W.bar[0], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == ngroups - 1) {
                __hip_atomic_store(&W.bar[0], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                last = true;
            }
        }
        if (last) {
            __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "agent");
            __hip_atomic_store(&W.bar[1], g0 + 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        } else {
            u32 spins = 0;
            while (__hip_atomic_load(&W.bar[1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == g0) {
                if (++spins > kSpinLimit) {
                    atomicOr(W.flags, FYZZ_DY_LOGIC);
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
// A wave whose lanes all wait backs off (64, 256, then 1024 cycles per
// poll), so polling leaves the memory system to the waves doing work.
__device__ inline void wait_backoff(u32 idle) {
    if (idle == 0) return;
    if (idle < 8)
        __builtin_amdgcn_s_sleep(1);
    else if (idle < 64)
        __builtin_amdgcn_s_sleep(4);
    else
        __builtin_amdgcn_s_sleep(16);
}
__device__ inline bool finite3(V v) { return isfinite(v.x) && isfinite(v.y) && isfinite(v.z); }
__device__ inline bool finite4(Q q) { return isfinite(q.x) && isfinite(q.y) && isfinite(q.z) && isfinite(q.w); }
__device__ inline V v3(const float* p) { return {p[0], p[1], p[2]}; }
__device__ inline Q q4(const float* p) { return {p[0], p[1], p[2], p[3]}; }
__device__ inline V f3(float4 f) { return {f.x, f.y, f.z}; }
__device__ inline float len2(V v) { return v.x * v.x + v.y * v.y + v.z * v.z; }
__device__ inline void put3(float* d, V s) {
    d[0] = s.x;
    d[1] = s.y;
    d[2] = s.z;
}
// mass pose of a box: identity mass-local frame (scene.cpp:359-365)
__device__ inline V mass_position(V p, Q q) { return vadd(p, srot(q, V{0.0f, 0.0f, 0.0f})); }
__device__ inline Q mass_orientation(Q q) { return qmul(q, Q{0.0f, 0.0f, 0.0f, 1.0f}); }

// ---------------------------------------------------------------- bodies

__device__ inline bool manager_ok(const Ws& W, const fyzz_dy_manager& M) {
    return !((M.body0 < 0 && M.body1 < 0) || M.body0 == M.body1 || M.body0 >= int(W.nb) || M.body1 >= int(W.nb) ||
             (M.body0 < 0 && u32(-1 - M.body0) >= W.ns) || (M.body1 < 0 && u32(-1 - M.body1) >= W.ns) ||
             M.count == 0 || M.count > 4);
}

// status flags and counters of the step (a fill kernel, so a captured
// step records no memset node)
__global__ __launch_bounds__(64) void k_clear(Ws W) {
    if (threadIdx.x < kCtrWords) W.ctr[threadIdx.x] = 0;
    if (threadIdx.x == 0) *W.flags = 0;
    if (threadIdx.x < kBarWords && threadIdx.x != 1) W.bar[threadIdx.x] = 0;  // counts, not the generation
}

// scene.cpp:10383-10507 for default dynamics (no pending velocity mods)
__global__ __launch_bounds__(kBlock) void k_pre(Ws W) {
    const u32 i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= W.nb) return;
    W.has_row[i] = 0;
    const fyzz_dy_body& b = W.bodies[i];
    if (b.flags & 1u) return;
    const float dt = W.dt;
    const float gravity = -9.81f;
    float vx = b.linear_velocity[0] + 0.0f + 0.0f * dt;
    float vy = b.linear_velocity[1] + 0.0f + 0.0f * dt;
    float vz = b.linear_velocity[2] + 0.0f + 0.0f * dt;
    vy = vy + ((b.flags & 4u) ? gravity * dt : 0.0f);
    const float lm = smax(1.0f - b.linear_damping * dt, 0.0f);
    vx *= lm;
    vy *= lm;
    vz *= lm;
    V a{b.angular_velocity[0] + 0.0f + 0.0f * dt, b.angular_velocity[1] + 0.0f + 0.0f * dt,
        b.angular_velocity[2] + 0.0f + 0.0f * dt};
    const float am = smax(1.0f - b.angular_damping * dt, 0.0f);
    a.x *= am;
    a.y *= am;
    a.z *= am;
    const float s2 = vx * vx + vy * vy + vz * vz;
    const float ml2 = b.max_linear_velocity * b.max_linear_velocity;
    const float ma2 = b.max_angular_velocity * b.max_angular_velocity;
    if (s2 > ml2) {
        const float f = sqrtf(ml2 / s2);
        vx *= f;
        vy *= f;
        vz *= f;
    }
    const float a2 = a.x * a.x + a.y * a.y + a.z * a.z;
    if (a2 > ma2) {
        const float f = sqrtf(ma2 / a2);
        a.x *= f;
        a.y *= f;
        a.z *= f;
    }
    if (!finite3(V{vx, vy, vz}) || !finite3(a)) raise_flag(W, FYZZ_DY_NONFINITE);
    W.pv[i] = make_float4(vx, vy, vz, 0.0f);
    W.pw[i] = make_float4(a.x, a.y, a.z, 0.0f);
}

// bodies of touching managers get a solver row (scene.cpp:6987-7013)
__global__ __launch_bounds__(kBlock) void k_mark(Ws W) {
    const u32 m = blockIdx.x * kBlock + threadIdx.x;
    if (m >= mcount(W)) return;
    const fyzz_dy_manager M = W.mg[m];
    if (!manager_ok(W, M)) {
        raise_flag(W, FYZZ_DY_BAD_INPUT);
        return;
    }
    if (M.body0 >= 0 && !(W.bodies[M.body0].flags & 1u)) W.has_row[M.body0] = 1;
    if (M.body1 >= 0 && !(W.bodies[M.body1].flags & 1u)) W.has_row[M.body1] = 1;
}

// world sqrt inverse inertia R diag(sqrt I^-1) R^T (scene.cpp:7092-7110)
__global__ __launch_bounds__(kBlock) void k_rows(Ws W) {
    // iteration counts (maximum over bodies with rows): one atomic per
    // workgroup and counter
    __shared__ u32 pos_it, vel_it;
    if (threadIdx.x == 0) pos_it = vel_it = 0;
    __syncthreads();
    const u32 i = blockIdx.x * kBlock + threadIdx.x;
    if (i < W.nb && W.has_row[i]) {
        const u32 it = W.bodies[i].iterations;
        atomicMax(&pos_it, it & 0xffu);
        atomicMax(&vel_it, (it >> 8) & 0xffu);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        if (pos_it) atomicMax(&W.ctr[kPosIters], pos_it);
        if (vel_it) atomicMax(&W.ctr[kVelIters], vel_it);
    }
    if (i >= W.nb) return;
    float4 c0 = make_float4(0.f, 0.f, 0.f, 0.f), c1 = c0, c2 = c0;
    W.ang[i] = c0;
    W.mlin[i] = c0;
    W.mang[i] = c0;
    if (!W.has_row[i]) {
        W.lin[i] = c0;
        W.rpos[i] = c0;
        W.rq[i] = make_float4(0.f, 0.f, 0.f, 1.f);
        W.sii[3 * i] = c0;
        W.sii[3 * i + 1] = c0;
        W.sii[3 * i + 2] = c0;
        return;
    }
    const fyzz_dy_body& b = W.bodies[i];
    const Q q = mass_orientation(q4(b.orientation));
    const V p = mass_position(v3(b.position), q4(b.orientation));
    W.rpos[i] = make_float4(p.x, p.y, p.z, 0.f);
    W.rq[i] = make_float4(q.x, q.y, q.z, q.w);
    W.lin[i] = make_float4(0.f, 0.f, 0.f, b.inverse_mass);
    const V l{sqrtf(b.inverse_inertia[0]), sqrtf(b.inverse_inertia[1]), sqrtf(b.inverse_inertia[2])};
    if (q.x == 0.0f && q.y == 0.0f && q.z == 0.0f && q.w == 1.0f) {
        c0.x = l.x;
        c1.y = l.y;
        c2.z = l.z;
    } else {
        const V rx = srot(q, V{1.f, 0.f, 0.f});
        const V ry = srot(q, V{0.f, 1.f, 0.f});
        const V rz = srot(q, V{0.f, 0.f, 1.f});
        const float xx = l.x * rx.x * rx.x + l.y * ry.x * ry.x + l.z * rz.x * rz.x;
        const float yy = l.x * rx.y * rx.y + l.y * ry.y * ry.y + l.z * rz.y * rz.y;
        const float zz = l.x * rx.z * rx.z + l.y * ry.z * ry.z + l.z * rz.z * rz.z;
        const float xy = l.x * rx.x * rx.y + l.y * ry.x * ry.y + l.z * rz.x * rz.y;
        const float xz = l.x * rx.x * rx.z + l.y * ry.x * ry.z + l.z * rz.x * rz.z;
        const float yz = l.x * rx.y * rx.z + l.y * ry.y * ry.z + l.z * rz.y * rz.z;
        c0 = make_float4(xx, xy, xz, 0.f);
        c1 = make_float4(xy, yy, yz, 0.f);
        c2 = make_float4(xz, yz, zz, 0.f);
    }
    W.sii[3 * i] = c0;
    W.sii[3 * i + 1] = c1;
    W.sii[3 * i + 2] = c2;
}

// ------------------------------------------------------------ partitions

__device__ inline bool dyn_dyn(const fyzz_dy_manager& M) { return M.body0 >= 0 && M.body1 >= 0; }

// First-fit partitions over 32-bit masks (scene.cpp:7352-7445) for one task
// spanning the manager list; stable scatter by partition (7447-7489).
//
// The data-parallel phases are plain kernels over bodies or managers:
// degrees, their scan, incidence lists (CSR by body), each body's list in
// manager order with the predecessor of every dynamic/dynamic entry,
// static managers after the body's dynamic partitions, sizes, offsets and
// the scatter. The colouring is one cooperative kernel: a manager's
// partition depends only on the masks its two bodies carry after their
// earlier managers, so each thread walks its managers in increasing order
// (m = tid, tid + threads, ...) and colours one as soon as the entries of
// its two predecessors are published. An entry is one 64-bit word (ready
// bit, chain level, mask after the manager) written and read as a single
// relaxed agent-scope atomic, cleared by the degree kernel every step. The
// smallest uncoloured manager is always ready and its thread is resident,
// so the colouring always advances; its waits are bounded regardless.

constexpr u64 kEntryReady = 1ull << 63;

// per body: degree, CSR fill cursor, highest dynamic partition
__global__ __launch_bounds__(kBlock) void k_part_clear(Ws W) {
    const u32 i = blockIdx.x * kBlock + threadIdx.x;
    if (i < W.nb) {
        W.deg[i] = 0;
        W.fill[i] = 0;
        W.max_dyn[i] = 0;
#pragma unroll
        for (u32 k = 0; k < 8; ++k) W.vst[8 * i + k] = 0ull;  // 0.0f, no updates yet
    }
    if (i < kMaxParts) {
        W.part_count[i] = 0;
        W.part_fill[i] = 0;
    }
    if (i == 0) W.ctr[kNbWord] = W.nb;
}

// incidence degrees (dynamic sides); clears the manager's entries
__global__ __launch_bounds__(kBlock) void k_part_degrees(Ws W) {
    const u32 m = blockIdx.x * kBlock + threadIdx.x;
    if (m >= mcount(W)) return;
    const fyzz_dy_manager& M = W.mg[m];
    W.part[m] = kNone;
    W.entry[2 * m] = 0ull;
    W.entry[2 * m + 1] = 0ull;
    const bool ok = manager_ok(W, M);
    W.dd[m] = ok && dyn_dyn(M) ? 1u : 0u;
    if (!ok) return;
    if (M.body0 >= 0) atomicAdd(&W.deg[M.body0], 1u);
    if (M.body1 >= 0) atomicAdd(&W.deg[M.body1], 1u);
}

struct DegreeOf {
    const u32* deg;
    __device__ u32 operator()(u32 i) const { return deg[i]; }
};

// incidence lists, entry = manager << 1 | side
__global__ __launch_bounds__(kBlock) void k_part_fill(Ws W) {
    const u32 m = blockIdx.x * kBlock + threadIdx.x;
    if (m >= mcount(W)) return;
    const fyzz_dy_manager& M = W.mg[m];
    if (!manager_ok(W, M)) return;
    if (M.body0 >= 0) W.inc[W.off[M.body0] + atomicAdd(&W.fill[M.body0], 1u)] = m << 1;
    if (M.body1 >= 0) W.inc[W.off[M.body1] + atomicAdd(&W.fill[M.body1], 1u)] = (m << 1) | 1u;
}

// each body's list in manager order; predecessor links of the
// dynamic/dynamic entries
__global__ __launch_bounds__(kBlock) void k_part_chain(Ws W) {
    const u32 b = blockIdx.x * kBlock + threadIdx.x;
    if (b >= W.nb) return;
    const u32 lo = W.off[b], hi = W.off[b + 1];
    for (u32 i = lo + 1; i < hi; ++i) {
        const u32 e = W.inc[i];
        u32 j = i;
        while (j > lo && W.inc[j - 1] > e) {
            W.inc[j] = W.inc[j - 1];
            --j;
        }
        W.inc[j] = e;
    }
    u32 prev = kNone;
    for (u32 i = lo; i < hi; ++i) {
        const u32 e = W.inc[i];
        if (!W.dd[e >> 1]) continue;
        W.pred[e] = prev;
        prev = e;
    }
}

// Ascending sort of a[0, p) by one workgroup (every thread calls it): a
// bitonic network over the next power of two, in the form whose every
// compare-exchange puts the smaller key at the lower index (each merge
// starts with a flipped half-cleaner), so the virtual padding (+inf keys
// beyond p) never moves and its exchanges are skipped. The pending list
// holds distinct manager indices.
__device__ void sort_pending(u32* a, u32 p) {
    u32 p2 = 1;
    while (p2 < p) p2 <<= 1;
    for (u32 k = 2; k <= p2; k <<= 1) {
        for (u32 j = k >> 1; j > 0; j >>= 1) {
            for (u32 t = threadIdx.x; t < (p2 >> 1); t += blockDim.x) {
                u32 l, r;
                if (j == (k >> 1)) {  // flipped half-cleaner: l pairs with its mirror in the block of k
                    const u32 blk = t / j, o = t % j;
                    l = blk * k + o;
                    r = blk * k + k - 1 - o;
                } else {
                    l = (t / j) * 2 * j + t % j;
                    r = l + j;
                }
                if (r < p) {
                    const u32 x = a[l], y = a[r];
                    if (x > y) {
                        a[l] = y;
                        a[r] = x;
                    }
                }
            }
            __threadfence_block();
            __syncthreads();
        }
    }
}

// The colouring (cooperative) and the overflow waves.
__global__ __launch_bounds__(kBlock) void k_part_color(Ws W) {
    u32 gen = barrier_start(W);
    const u32 nthreads = gridDim.x * kBlock;
    const u32 nm = mcount(W);
    u32 m = blockIdx.x * kBlock + threadIdx.x;
    bool have = false;
    u32 p0 = kNone, p1 = kNone, spins = 0, depth = 0, idle = 0;
    for (;;) {
        bool progress = false;
        if (m < nm && !have) {
            if (!W.dd[m]) {  // not a valid manager between two dynamic bodies
                m += nthreads;
                progress = true;
            } else {
                p0 = W.pred[2 * m];
                p1 = W.pred[2 * m + 1];
                have = true;
            }
        }
        if (m < nm && have) {
            const u64 a0 = p0 != kNone ? __hip_atomic_load(&W.entry[p0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)
                                       : kEntryReady;
            const u64 a1 = p1 != kNone ? __hip_atomic_load(&W.entry[p1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)
                                       : kEntryReady;
            if ((a0 & kEntryReady) && (a1 & kEntryReady)) {
                const u32 before0 = static_cast<u32>(a0), before1 = static_cast<u32>(a1);
                const u32 level = max(static_cast<u32>(a0 >> 32) & 0x7fffffffu,
                                      static_cast<u32>(a1 >> 32) & 0x7fffffffu) + 1u;
                depth = max(depth, level);
                const u32 avail = ~(before0 | before1);
                u32 after0 = before0, after1 = before1;
                if (avail) {
                    const u32 bit = __builtin_ctz(avail);
                    W.part[m] = bit;
                    after0 |= 1u << bit;
                    after1 |= 1u << bit;
                } else {
                    W.part[m] = kPending;
                    const u32 k = atomicAdd(&W.ctr[kPend], 1u);
                    if (k < kMaxPending) W.pend[k] = m;
                }
                const u64 hi = kEntryReady | (static_cast<u64>(level) << 32);
                __hip_atomic_store(&W.entry[2 * m], hi | after0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                __hip_atomic_store(&W.entry[2 * m + 1], hi | after1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                have = false;
                m += nthreads;
                progress = true;
                spins = 0;  // the budget bounds one wait
            } else if (++spins > kSpinLimit) {
                raise_flag(W, FYZZ_DY_LOGIC);
                m = nm;  // give up on this thread's managers
            }
        }
        const bool busy = m < nm;
        if (!__ballot(busy)) break;
        idle = __ballot(progress) ? 0u : idle + 1u;
        wait_backoff(idle);
    }
    // wave maximum first: one atomic per wave on the shared word
    for (int d = kWave / 2; d > 0; d >>= 1) depth = max(depth, static_cast<u32>(__shfl_xor(static_cast<int>(depth), d, kWave)));
    if (lane_id() == 0 && depth) atomicMax(&W.ctr[kRounds], depth);
    grid_sync(W, gen);
    // overflow waves: the unassigned managers, in order, with fresh masks
    if (blockIdx.x != 0) return;
    __shared__ u32 s_pending;
    if (threadIdx.x == 0) {
        u32 p = ld(&W.ctr[kPend]);
        if (p > kMaxPending) {
            raise_flag(W, FYZZ_DY_OVERFLOW);
            p = 0;
        }
        s_pending = p;
    }
    __syncthreads();
    sort_pending(W.pend, s_pending);
    if (threadIdx.x == 0) {
        u32 p = s_pending;
        for (u32 wave = 32; p; wave += 32) {
            for (u32 k = 0; k < p; ++k) {
                const fyzz_dy_manager& M = W.mg[W.pend[k]];
                W.body_mask[M.body0] = 0;
                W.body_mask[M.body1] = 0;
            }
            u32 over = 0;
            for (u32 k = 0; k < p; ++k) {
                const u32 mm = W.pend[k];
                const fyzz_dy_manager& M = W.mg[mm];
                const u32 avail = ~(W.body_mask[M.body0] | W.body_mask[M.body1]);
                if (!avail) {
                    W.pend[over++] = mm;
                    continue;
                }
                const u32 bit = __builtin_ctz(avail);
                W.body_mask[M.body0] |= 1u << bit;
                W.body_mask[M.body1] |= 1u << bit;
                W.part[mm] = wave + bit;
            }
            if (over == p) {
                raise_flag(W, FYZZ_DY_LOGIC);
                break;
            }
            p = over;
        }
    }
}

// highest dynamic partition per body
__global__ __launch_bounds__(kBlock) void k_part_maxdyn(Ws W) {
    const u32 m = blockIdx.x * kBlock + threadIdx.x;
    if (m >= mcount(W)) return;
    const fyzz_dy_manager& M = W.mg[m];
    if (!dyn_dyn(M)) return;
    const u32 pm = W.part[m];
    if (pm >= kPending) return;
    atomicMax(&W.max_dyn[M.body0], pm + 1);
    atomicMax(&W.max_dyn[M.body1], pm + 1);
}

// Dynamic/static managers follow the body's dynamic partitions; then each
// of the body's managers is ranked by partition (the order in which a
// sweep reaches them; a body's managers have distinct partitions). A
// static manager touches one dynamic body, so this thread owns its
// partition; the dynamic ones are final after the colouring.
__global__ __launch_bounds__(kBlock) void k_part_static(Ws W) {
    const u32 b = blockIdx.x * kBlock + threadIdx.x;
    if (b >= W.nb) return;
    const u32 lo = W.off[b], hi = W.off[b + 1];
    u32 k = W.max_dyn[b];
    // Bodies with at most kCache managers (all of a pile's) keep the
    // partitions in registers: one load per manager instead of one per
    // pair of managers. A static manager has this body as its only dynamic
    // body, so this thread alone writes its partition.
    constexpr u32 kCache = 16;
    if (hi - lo <= kCache) {
        u32 pc[kCache], ec[kCache];
#pragma unroll
        for (u32 q = 0; q < kCache; ++q) {
            pc[q] = kNone;  // absent entries never count as lower
            ec[q] = 0;
            if (lo + q < hi) {
                const u32 e = W.inc[lo + q];
                const u32 m = e >> 1;
                ec[q] = e;
                if (!W.dd[m]) {
                    W.part[m] = k;
                    pc[q] = k++;
                } else {
                    pc[q] = W.part[m];
                }
            }
        }
#pragma unroll
        for (u32 q = 0; q < kCache; ++q) {
            if (lo + q >= hi) continue;
            u32 r = 0;
#pragma unroll
            for (u32 t = 0; t < kCache; ++t) r += pc[t] < pc[q] ? 1u : 0u;
            W.erank[ec[q]] = r;
        }
        return;
    }
    for (u32 i = lo; i < hi; ++i) {
        const u32 m = W.inc[i] >> 1;
        if (!W.dd[m]) W.part[m] = k++;
    }
    for (u32 i = lo; i < hi; ++i) {
        const u32 e = W.inc[i];
        const u32 pe = W.part[e >> 1];
        u32 r = 0;
        for (u32 j = lo; j < hi; ++j) r += W.part[W.inc[j] >> 1] < pe ? 1u : 0u;
        W.erank[e] = r;
    }
}

// Partition sizes and count. A pile has a few tens of partitions, so one
// global counter per manager would serialize ~10^4 atomics per address:
// each workgroup counts in LDS (its managers' partition range) and adds
// one value per partition it holds.
__global__ __launch_bounds__(kBlock) void k_part_sizes(Ws W) {
    __shared__ u32 cnt[kMaxParts];
    __shared__ u32 top;
    const u32 m = blockIdx.x * kBlock + threadIdx.x;
    for (u32 i = threadIdx.x; i < kMaxParts; i += kBlock) cnt[i] = 0;
    if (threadIdx.x == 0) top = 0;
    __syncthreads();
    const u32 pm = m < mcount(W) ? W.part[m] : kNone;
    if (pm != kNone) {  // kNone: malformed record, already flagged
        if (pm >= kMaxParts) {
            raise_flag(W, FYZZ_DY_OVERFLOW);
        } else {
            atomicAdd(&cnt[pm], 1u);
            atomicMax(&top, pm + 1);
        }
    }
    __syncthreads();
    const u32 np = top;
    for (u32 i = threadIdx.x; i < np; i += kBlock)
        if (cnt[i]) atomicAdd(&W.part_count[i], cnt[i]);
    if (threadIdx.x == 0 && np) atomicMax(&W.ctr[kNParts], np);
}

// partition offsets (one workgroup)
__global__ __launch_bounds__(kBlock) void k_part_offsets(Ws W) {
    __shared__ u32 scratch[kBlock / kWave + 1];
    const u32 np = W.ctr[kNParts];
    u32 carry = 0, total;
    for (u32 t0 = 0; t0 < np; t0 += kBlock) {
        const u32 t = t0 + threadIdx.x;
        const u32 v = t < np ? W.part_count[t] : 0u;
        const u32 ex = block_exclusive_scan<kBlock>(v, scratch, &total);
        if (t < np) W.part_off[t] = carry + ex;
        carry += total;
    }
    if (threadIdx.x == 0) W.part_off[np] = carry;
}

// Managers grouped by partition (order inside a partition is free): each
// workgroup ranks its managers per partition in LDS, reserves one range
// per partition it holds, then writes.
__global__ __launch_bounds__(kBlock) void k_part_scatter(Ws W) {
    __shared__ u32 cnt[kMaxParts];
    __shared__ u32 base[kMaxParts];
    __shared__ u32 top;
    const u32 m = blockIdx.x * kBlock + threadIdx.x;
    for (u32 i = threadIdx.x; i < kMaxParts; i += kBlock) cnt[i] = 0;
    if (threadIdx.x == 0) top = 0;
    __syncthreads();
    const u32 pm = m < mcount(W) ? W.part[m] : kNone;
    const bool live = pm < kMaxParts;
    u32 local = 0;
    if (live) {
        local = atomicAdd(&cnt[pm], 1u);
        atomicMax(&top, pm + 1);
    }
    __syncthreads();
    const u32 np = top;
    for (u32 i = threadIdx.x; i < np; i += kBlock)
        if (cnt[i]) base[i] = W.part_off[i] + atomicAdd(&W.part_fill[i], cnt[i]);
    __syncthreads();
    if (live) {
        const fyzz_dy_manager& M = W.mg[m];
        const u32 pos = base[pm] + local;
        W.sorted[pos] = m;
        W.pos_of[m] = pos;
        W.sinfo[pos] = make_int4(M.body0, M.body1, static_cast<int>(m), 0);
        W.srank[pos] = make_uint4(M.body0 >= 0 ? W.erank[2 * m] : 0u, M.body1 >= 0 ? W.erank[2 * m + 1] : 0u,
                                  M.body0 >= 0 ? W.deg[M.body0] : 0u, M.body1 >= 0 ? W.deg[M.body1] : 0u);
    }
}

// ----------------------------------------------------- contact preparation

struct Row {
    V pos;
    Q q;
    V v, w;
    float inv_mass;
    V c0, c1, c2;
};
__device__ inline Row load_row(const Ws& W, int id) {
    Row r{};
    r.q = Q{0.f, 0.f, 0.f, 1.f};
    if (id < 0 || !W.has_row[id]) return r;
    r.pos = f3(W.rpos[id]);
    const float4 q = W.rq[id];
    r.q = Q{q.x, q.y, q.z, q.w};
    r.v = f3(W.pv[id]);
    r.w = f3(W.pw[id]);
    r.inv_mass = W.lin[id].w;
    r.c0 = f3(W.sii[3 * id]);
    r.c1 = f3(W.sii[3 * id + 1]);
    r.c2 = f3(W.sii[3 * id + 2]);
    return r;
}
__device__ inline V transform(const Row& d, V v) {
    return {d.c0.x * v.x + d.c1.x * v.y + d.c2.x * v.z, d.c0.y * v.x + d.c1.y * v.y + d.c2.y * v.z,
            d.c0.z * v.x + d.c1.z * v.y + d.c2.z * v.z};
}
struct BodyView {
    V pos;
    Q q;
};
__device__ inline BodyView view_of(const Ws& W, int id) {
    if (id >= 0) return {v3(W.bodies[id].position), q4(W.bodies[id].orientation)};
    const fyzz_dy_static& s = W.statics[-1 - id];
    return {v3(s.position), q4(s.orientation)};
}
__device__ inline V local_point(const BodyView& a, V world) {
    const Q q = mass_orientation(a.q);
    const V p = mass_position(a.pos, a.q);
    return srot(Q{-q.x, -q.y, -q.z, q.w}, vsub(world, p));
}
__device__ inline V local_normal(const BodyView& a, V n) {
    const Q q = mass_orientation(a.q);
    return srot(Q{-q.x, -q.y, -q.z, q.w}, n);
}
__device__ inline V world_point(const BodyView& a, V local) {
    const V r = srot(mass_orientation(a.q), local);
    const V p = mass_position(a.pos, a.q);
    return vadd(r, p);
}

struct Patch {
    V normal, n0, n1;
    float sf, df, rest, damping;
    V a0[2], a1[2];
    V bmin, bmax;
    u32 contacts, anchors;
    bool broken, disable_strong;
};

// Field-by-field copy (a whole-struct copy from a runtime-indexed slot
// would keep the destination in memory as well).
__device__ inline void copy_patch(Patch& d, const Patch& s) {
    d.normal = V{s.normal.x, s.normal.y, s.normal.z};
    d.n0 = V{s.n0.x, s.n0.y, s.n0.z};
    d.n1 = V{s.n1.x, s.n1.y, s.n1.z};
    d.sf = s.sf;
    d.df = s.df;
    d.rest = s.rest;
    d.damping = s.damping;
#pragma unroll
    for (int k = 0; k < 2; ++k) {
        d.a0[k] = V{s.a0[k].x, s.a0[k].y, s.a0[k].z};
        d.a1[k] = V{s.a1[k].x, s.a1[k].y, s.a1[k].z};
    }
    d.bmin = V{s.bmin.x, s.bmin.y, s.bmin.z};
    d.bmax = V{s.bmax.x, s.bmax.y, s.bmax.z};
    d.contacts = s.contacts;
    d.anchors = s.anchors;
    d.broken = s.broken;
    d.disable_strong = s.disable_strong;
}

// scene.cpp:3981-4010
__device__ inline bool history_valid(const Patch& old, const BodyView& a0, const BodyView& a1) {
    if (old.broken || old.disable_strong || old.anchors == 0u) return false;
    const float same_normal = 0.999f;
    const float corr = 0.025f;
    const Q q0 = mass_orientation(a0.q), q1 = mass_orientation(a1.q);
    const V n10 = srot_inv(q0, srot(q1, old.n1));
    if (vdot(old.n0, n10) <= same_normal) return false;
    const V p0 = mass_position(a0.pos, a0.q), p1 = mass_position(a1.pos, a1.q);
    // constant anchor indices (at most two), so the patch stays in registers
#pragma unroll
    for (u32 k = 0; k < 2; ++k) {
        if (k >= old.anchors) break;
        const V l = srot(q1, old.a1[k]);
        const V world = vadd(l, p1);
        const V in0 = srot_inv(q0, vsub(world, p0));
        const V d = vsub(old.a0[k], in0);
        if (fabsf(vdot(d, old.n0)) >= corr) return false;
    }
    return true;
}

struct Run {
    u32 start, count;
    V normal, bmin, bmax;
    int next;
};

// One thread per manager: materials, same-normal runs, friction-patch
// correlation with the previous step, contact rows, anchor growth and
// friction rows (scene.cpp:7794-8505).
__global__ __launch_bounds__(128) void k_prep(Ws W) {
    const u32 m = blockIdx.x * 128 + threadIdx.x;
    if (m >= mcount(W)) return;
    const fyzz_dy_manager& M = W.mg[m];
    if (W.part[m] >= kMaxParts) return;
    const float inv_dt = 1.0f / W.dt;
    const BodyView a0 = view_of(W, M.body0), a1 = view_of(W, M.body1);
    const Row a = load_row(W, M.body0);
    const Row b = load_row(W, M.body1);
    float sf0, df0, r0, sf1, df1, r1;
    if (M.body0 >= 0) {
        sf0 = W.bodies[M.body0].static_friction;
        df0 = W.bodies[M.body0].dynamic_friction;
        r0 = W.bodies[M.body0].restitution;
    } else {
        sf0 = W.statics[-1 - M.body0].static_friction;
        df0 = W.statics[-1 - M.body0].dynamic_friction;
        r0 = W.statics[-1 - M.body0].restitution;
    }
    if (M.body1 >= 0) {
        sf1 = W.bodies[M.body1].static_friction;
        df1 = W.bodies[M.body1].dynamic_friction;
        r1 = W.bodies[M.body1].restitution;
    } else {
        sf1 = W.statics[-1 - M.body1].static_friction;
        df1 = W.statics[-1 - M.body1].dynamic_friction;
        r1 = W.statics[-1 - M.body1].restitution;
    }
    if (r0 < 0.0f || r1 < 0.0f) {  // compliant contacts are not admitted
        raise_flag(W, FYZZ_DY_LOGIC);
        return;
    }
    const float rigid_contact_bias = 0.8f;
    const float bounce_threshold = -2.0f;
    const float same_normal = 0.999f;
    const float friction_offset_threshold = 0.04f;
    const float restitution = 0.5f * (r0 + r1);
    const float damping = 0.5f * (0.0f + 0.0f);
    const float dynamic_friction = smax(0.0f, 0.5f * (df0 + df1));
    const float static_friction = smax(dynamic_friction, 0.5f * (sf0 + sf1));
    const float max_pen_bias = -1.0e32f;
    const float rest_distance = 0.0f;
    // the previous step's patch when it correlates (scene.cpp:8007-8030)
    bool have_old = false;
    Patch c{};
    if (M.history != kNone && M.history < W.nh && W.history[M.history].anchor_count > 2u) {
        raise_flag(W, FYZZ_DY_BAD_INPUT);
    } else if (M.history != kNone && M.history < W.nh) {
        const fyzz_dy_patch& h = W.history[M.history];
        Patch old{};
        old.n0 = v3(h.body0_normal);
        old.n1 = v3(h.body1_normal);
        for (int k = 0; k < 2; ++k) {
            old.a0[k] = v3(h.body0_anchors[k]);
            old.a1[k] = v3(h.body1_anchors[k]);
        }
        old.sf = h.static_friction;
        old.df = h.dynamic_friction;
        old.rest = h.restitution;
        old.damping = h.damping;
        old.anchors = h.anchor_count;
        old.broken = h.flags & 1u;
        old.disable_strong = (h.flags >> 1) & 1u;
        if (history_valid(old, a0, a1)) {
            c = old;
            c.normal = srot(mass_orientation(a0.q), old.n0);
            c.contacts = 0;
            c.bmin = {kFltMax, kFltMax, kFltMax};
            c.bmax = {-kFltMax, -kFltMax, -kFltMax};
            have_old = true;
        }
    }
    // The kept patch, its header normal and the contact order (2-bit fields,
    // correlation-list order). A manager whose contacts all continue the
    // first contact's normal (every PCM manager: one published normal)
    // forms one run: the previous patch if it correlates, else one new
    // patch, keeps all contacts in their own order. Only other managers
    // take the general path over runs and candidate patches.
    Patch P;
    V header;
    u32 order_bits = 0, no = 0;
    bool one_run = M.count > 0 && M.count <= 4;
    {
        const V n0c = v3(M.contacts[0].normal);
#pragma unroll
        for (u32 jj = 1; jj < 4; ++jj)
            if (jj < M.count && vdot(n0c, v3(M.contacts[jj].normal)) < same_normal) one_run = false;
    }
    if (one_run) {
        // the run (scene.cpp:7964-7997) and its correlation (8007-8099)
        const V rn = v3(M.contacts[0].normal);
        V rbmin = v3(M.contacts[0].position), rbmax = rbmin;
#pragma unroll
        for (u32 jj = 1; jj < 4; ++jj)
            if (jj < M.count) {
                const V pp = v3(M.contacts[jj].position);
                rbmin = {smin(rbmin.x, pp.x), smin(rbmin.y, pp.y), smin(rbmin.z, pp.z)};
                rbmax = {smax(rbmax.x, pp.x), smax(rbmax.y, pp.y), smax(rbmax.z, pp.z)};
            }
        if (have_old && vdot(c.normal, rn) >= same_normal && c.rest == restitution && c.damping == damping &&
            c.sf == static_friction && c.df == dynamic_friction) {
            P = c;
        } else {
            P = Patch{};
            P.normal = rn;
            P.sf = static_friction;
            P.df = dynamic_friction;
            P.rest = restitution;
            P.damping = damping;
            P.n0 = local_normal(a0, rn);
            P.n1 = local_normal(a1, rn);
        }
        P.bmin = rbmin;
        P.bmax = rbmax;
        P.contacts = M.count;
        header = rn;
        no = M.count;
        order_bits = 0u | (1u << 2) | (2u << 4) | (3u << 6);
    } else {
        // contiguous same-normal runs (scene.cpp:7964-7997)
        Run runs[4];
        u32 nr = 0;
        for (u32 j = 0; j < M.count; ++j) {
            const V pn = v3(M.contacts[j].normal);
            const V pp = v3(M.contacts[j].position);
            if (nr == 0 || vdot(runs[nr - 1].normal, pn) < same_normal) {
                runs[nr++] = Run{j, 1, pn, pp, pp, -1};
            } else {
                Run& r = runs[nr - 1];
                ++r.count;
                r.bmin = {smin(r.bmin.x, pp.x), smin(r.bmin.y, pp.y), smin(r.bmin.z, pp.z)};
                r.bmax = {smax(r.bmax.x, pp.x), smax(r.bmax.y, pp.y), smax(r.bmax.z, pp.z)};
            }
        }
        // correlation against the previous step's patch (scene.cpp:8007-8099)
        Patch local[5];
        u32 nl = 0;
        if (have_old) local[nl++] = c;
        int heads[5] = {-1, -1, -1, -1, -1};
        for (u32 ri = 0; ri < nr; ++ri) {
            Run& cp = runs[ri];
            u32 lp = 0;
            for (; lp < nl; ++lp) {
                const Patch& p = local[lp];
                if (vdot(p.normal, cp.normal) >= same_normal && p.rest == restitution && p.damping == damping &&
                    p.sf == static_friction && p.df == dynamic_friction)
                    break;
            }
            if (lp == nl) {
                Patch p{};
                p.normal = cp.normal;
                p.sf = static_friction;
                p.df = dynamic_friction;
                p.rest = restitution;
                p.damping = damping;
                p.n0 = local_normal(a0, cp.normal);
                p.n1 = local_normal(a1, cp.normal);
                local[nl++] = p;
            }
            Patch& p = local[lp];
            if (p.contacts == 0) {
                p.bmin = cp.bmin;
                p.bmax = cp.bmax;
            } else {
                p.bmin = {smin(p.bmin.x, cp.bmin.x), smin(p.bmin.y, cp.bmin.y), smin(p.bmin.z, cp.bmin.z)};
                p.bmax = {smax(p.bmax.x, cp.bmax.x), smax(p.bmax.y, cp.bmax.y), smax(p.bmax.z, cp.bmax.z)};
            }
            p.contacts += cp.count;
            cp.next = heads[lp];
            heads[lp] = int(ri);
        }
        // exactly one patch keeps contacts: PCM managers publish one normal
        int kept = -1, nkept = 0;
        for (u32 j = 0; j < nl; ++j)
            if (local[j].contacts) {
                if (kept < 0) kept = int(j);
                ++nkept;
            }
        if (nkept != 1) {
            raise_flag(W, FYZZ_DY_LOGIC);
            return;
        }
        copy_patch(P, local[kept]);
        header = v3(M.contacts[runs[heads[kept]].start].normal);
        for (int cp = heads[kept]; cp >= 0; cp = runs[cp].next)
            for (u32 c = runs[cp].start; c < runs[cp].start + runs[cp].count; ++c) {
                order_bits |= c << (2 * no);
                ++no;
            }
    }
    // contact rows in the correlation-list order (scene.cpp:8204-8225, 8237-8330)
    bool bad = false;
    for (u32 q = 0; q < no; ++q) {
        const fyzz_dy_contact& C = M.contacts[(order_bits >> (2 * q)) & 3u];
        const V n = header;
        const V cpos = v3(C.position);
        const V ra_x_n = vcross(vsub(cpos, a.pos), n);
        const V rb_x_n = vcross(vsub(cpos, b.pos), n);
        const float lin_vrel = vdot(n, vsub(a.v, b.v));
        const V ra = transform(a, ra_x_n);
        const V rb = transform(b, rb_x_n);
        const float ang_vrel = vdot(ra_x_n, a.w) - vdot(rb_x_n, b.w);
        const float vrel = lin_vrel + ang_vrel;
        const float response = (a.inv_mass + vdot(ra, ra)) + (b.inv_mass + vdot(rb, rb));
        const float vm = response > 0.0f ? 1.0f / response : 0.0f;
        const float pen = C.separation - rest_distance;
        const float pen_inv_dt = pen * inv_dt;
        const bool separated = pen >= 0.0f;
        const bool closing = -vrel > pen_inv_dt;
        const bool bouncing = restitution > 0.0f && bounce_threshold > vrel && closing;
        const float target_velocity = bouncing ? -vrel * restitution : 0.0f;
        const float target = target_velocity - vrel;
        const float pen_scaled = separated ? pen_inv_dt : pen * (inv_dt * rigid_contact_bias);
        const float scaled_bias = bouncing ? 0.0f : vm * smax(max_pen_bias, pen_scaled);
        CRow r;
        r.nx = n.x;
        r.ny = n.y;
        r.nz = n.z;
        r.vm = vm;
        r.rax = ra.x;
        r.ray = ra.y;
        r.raz = ra.z;
        r.biased = target * vm - scaled_bias;
        r.rbx = rb.x;
        r.rby = rb.y;
        r.rbz = rb.z;
        r.unbiased = target * vm - (bouncing ? 0.0f : smax(scaled_bias, 0.0f));
        r.applied = 0.0f;
        r.max_impulse = kFltMax;
        r.pad0 = r.pad1 = 0.0f;
        bad |= !isfinite(response) || !isfinite(r.biased) || !isfinite(r.unbiased);
        W.crow[4 * m + q] = r;
        if (W.tape) {
            // err = target vm - bias with target = -(1 + e) vrel when bouncing, -vrel otherwise
            W.craw[2 * (4 * m + q)] = make_float4(ra_x_n.x, ra_x_n.y, ra_x_n.z, bouncing ? -(1.0f + restitution) : -1.0f);
            W.craw[2 * (4 * m + q) + 1] = make_float4(rb_x_n.x, rb_x_n.y, rb_x_n.z, 0.0f);
        }
    }
    if (bad) raise_flag(W, FYZZ_DY_NONFINITE);
    // anchor growth (scene.cpp:8359-8433)
    {
        const V extent = vsub(P.bmax, P.bmin);
        const float diag2 = len2(extent);
        bool keep_pair = false;
        if (P.anchors == 2) {
            const V d = vsub(P.a0[0], P.a0[1]);
            keep_pair = len2(d) * 4.0f >= diag2;
            if (!keep_pair) P.anchors = 0;
        }
        V anchors[2] = {V{0.f, 0.f, 0.f}, V{0.f, 0.f, 0.f}};
        bool changed[2] = {false, false};
        u32 count = P.anchors;
        if (!keep_pair) {
            if (count == 1) anchors[0] = world_point(a0, P.a0[0]);
            float pd2 = 0.0f;
#pragma unroll
            for (u32 q = 0; q < 4; ++q) {
                if (q >= no) continue;
                const u32 j = (order_bits >> (2 * q)) & 3u;
                if (M.contacts[j].separation >= friction_offset_threshold + rest_distance) continue;
                const V wp = v3(M.contacts[j].position);
                if (count == 0) {
                    anchors[0] = wp;
                    changed[0] = true;
                    ++count;
                } else if (count == 1) {
                    const float d = len2(vsub(wp, anchors[0]));
                    if (d > 1.0e-8f) {
                        anchors[1] = wp;
                        changed[1] = true;
                        pd2 = d;
                        ++count;
                    }
                } else {
                    const float d0 = len2(vsub(wp, anchors[0]));
                    const float d1 = len2(vsub(wp, anchors[1]));
                    if (d0 > d1 && d0 > pd2) {
                        anchors[1] = wp;
                        changed[1] = true;
                        pd2 = d0;
                    } else if (d1 >= d0 && d1 > pd2) {
                        anchors[0] = wp;
                        changed[0] = true;
                        pd2 = d1;
                    }
                }
            }
#pragma unroll
            for (u32 k = 0; k < 2; ++k) {
                if (k >= count || !changed[k]) continue;
                P.a0[k] = local_point(a0, anchors[k]);
                P.a1[k] = local_point(a1, anchors[k]);
            }
            if (count == 0) {
                P.a0[0] = V{0.f, 0.f, 0.f};
                P.a1[0] = V{0.f, 0.f, 0.f};
            }
        }
        P.anchors = count;
        if (count) {
            const V n = header;
            const V rel = vsub(a.v, b.v);
            V t0 = vsub(rel, V{n.x * vdot(rel, n), n.y * vdot(rel, n), n.z * vdot(rel, n)});
            if (len2(t0) <= 0.0001f) t0 = fabsf(n.x) < 0.70710678f ? V{0.0f, -n.z, n.y} : V{-n.y, n.x, 0.0f};
            const float il = 1.0f / sqrtf(len2(t0));
            t0 = {t0.x * il, t0.y * il, t0.z * il};
            const V t1 = vcross(n, t0);
#pragma unroll
            for (u32 k = 0; k < 2; ++k) {
                if (k >= count) continue;
                const V ra = srot(mass_orientation(a0.q), P.a0[k]);
                const V rb = srot(mass_orientation(a1.q), P.a1[k]);
                const V c0 = mass_position(a0.pos, a0.q), c1 = mass_position(a1.pos, a1.q);
                const V err = vsub(vadd(ra, c0), vadd(rb, c1));
                for (u32 ti = 0; ti < 2; ++ti) {
                    const V tangent = ti ? t1 : t0;
                    const V ax0 = vcross(ra, tangent);
                    const V ax1 = vcross(rb, tangent);
                    const V j0 = transform(a, ax0);
                    const V j1 = transform(b, ax1);
                    const float response = (a.inv_mass + vdot(j0, j0)) + (b.inv_mass + vdot(j1, j1));
                    FRow fr;
                    fr.tx = tangent.x;
                    fr.ty = tangent.y;
                    fr.tz = tangent.z;
                    fr.vm = response > 0.0f ? rigid_contact_bias / response : 0.0f;
                    fr.rax = j0.x;
                    fr.ray = j0.y;
                    fr.raz = j0.z;
                    fr.bias = vdot(tangent, err) * inv_dt;
                    fr.rbx = j1.x;
                    fr.rby = j1.y;
                    fr.rbz = j1.z;
                    fr.target = -(vdot(tangent, rel) + vdot(ax0, a.w) - vdot(ax1, b.w));
                    fr.applied = 0.0f;
                    fr.pad0 = fr.pad1 = fr.pad2 = 0.0f;
                    W.frow[4 * m + 2 * k + ti] = fr;
                    if (W.tape) {
                        W.fraw[2 * (4 * m + 2 * k + ti)] = make_float4(ax0.x, ax0.y, ax0.z, 0.0f);
                        W.fraw[2 * (4 * m + 2 * k + ti) + 1] = make_float4(ax1.x, ax1.y, ax1.z, 0.0f);
                    }
                }
            }
        }
    }
    fyzz_dy_patch o;
    put3(o.body0_normal, P.n0);
    put3(o.body1_normal, P.n1);
    for (int k = 0; k < 2; ++k) {
        put3(o.body0_anchors[k], P.a0[k]);
        put3(o.body1_anchors[k], P.a1[k]);
    }
    o.static_friction = P.sf;
    o.dynamic_friction = P.df;
    o.restitution = P.rest;
    o.damping = P.damping;
    o.anchor_count = P.anchors;
    o.flags = (P.broken ? 1u : 0u) | (P.disable_strong ? 2u : 0u);
    W.patches[m] = o;
    W.order[m] = no | ((order_bits & 3u) << 8) | ((no > 1 ? (order_bits >> 2) & 3u : 0u) << 12) |
                 ((no > 2 ? (order_bits >> 4) & 3u : 0u) << 16) | ((no > 3 ? (order_bits >> 6) & 3u : 0u) << 20);
    // the sweep reads the row counts with the position record
    W.sinfo[W.pos_of[m]].w = static_cast<int>(no | (P.anchors << 8));
}

// ------------------------------------------------------------------ solve

// scene.cpp:8985-9010 (normal rows), 9050-9120 (friction rows)
// Coherent body state: during the sweep a body's velocity change lives in
// one 64-byte line of three 16-byte chunks, each chunk two components'
// bits, each followed by the number of updates applied to the body, moved
// between workgroups on different XCDs by device-coherent (sc1) 16-byte
// loads and stores (one request per chunk). A reader takes the state when
// all six counts carry the count it waits for: a count is written once
// per body, so matching counts give the values written with it, and a
// writer needs no wait between its state stores and a separate counter.
struct BodyState {
    V v, w;
};
typedef u32 Chunk __attribute__((ext_vector_type(4)));
constexpr u32 kStateBytes = 64;
constexpr int kDeviceCoherent = 16;  // cache policy: sc1 (device scope)
__device__ inline __amdgpu_buffer_rsrc_t state_rsrc(const Ws& W) {
    // raw buffer over the state lines, range check off (offsets < 2^32)
    return __builtin_amdgcn_make_buffer_rsrc(W.vst, 0, 0xFFFFFFFF, 0x00020000);
}
__device__ inline Chunk load_chunk(__amdgpu_buffer_rsrc_t R, u32 off) {
    return __builtin_amdgcn_raw_buffer_load_b128(R, off, 0, kDeviceCoherent);
}
// the state of three chunks as loaded, and whether all counts are `need`
__device__ inline bool take_state(Chunk a, Chunk b, Chunk c, u32 need, BodyState& s) {
    s.v = {__uint_as_float(a.x), __uint_as_float(a.z), __uint_as_float(b.x)};
    s.w = {__uint_as_float(b.z), __uint_as_float(c.x), __uint_as_float(c.z)};
    return a.y == need && a.w == need && b.y == need && b.w == need && c.y == need && c.w == need;
}
__device__ inline void publish_state(__amdgpu_buffer_rsrc_t R, u32 body, u32 count, const V& v, const V& w) {
    const u32 off = body * kStateBytes;
    const Chunk a = {__float_as_uint(v.x), count, __float_as_uint(v.y), count};
    const Chunk b = {__float_as_uint(v.z), count, __float_as_uint(w.x), count};
    const Chunk c = {__float_as_uint(w.y), count, __float_as_uint(w.z), count};
    __builtin_amdgcn_raw_buffer_store_b128(a, R, off, 0, kDeviceCoherent);
    __builtin_amdgcn_raw_buffer_store_b128(b, R, off + 16u, 0, kDeviceCoherent);
    __builtin_amdgcn_raw_buffer_store_b128(c, R, off + 32u, 0, kDeviceCoherent);
}
// Clock cycles before a body still (need - seen) updates short of a
// waiting manager is polled again: about one solve hop per missing update
// (kPollHopCycles), at most 32 hops; a body one update short (or a torn
// read, seen == need) is polled on the next pass.
constexpr u32 kPollHopCycles = 1024;
__device__ inline u64 poll_delay(u32 need, u32 seen) {
    const u32 d = need - seen;  // counts only grow up to need
    return d <= 1u ? 0ull : static_cast<u64>(min(d - 1u, 32u)) * kPollHopCycles;
}

// One manager's rows of one iteration (scene.cpp:8853-9148). info: (body0,
// body1, manager) of a partition-sorted position. Rows and patches are read
// with plain loads: the same thread solves the same manager in every
// iteration, and nothing else writes them during the sweep.
// s0 / s1: the bodies' states as taken by the sweep (zero for a static
// side); ima / imb: inverse masses. The updated states are published with
// counts c0 / c1; fin0 / fin1: the body's last update of the sweep, whose
// state also goes to the plain rows read after the sweep.
// Loads the contact and friction rows in use (row counts come with the
// position record), all at once.
__device__ inline void load_rows(const Ws& W, int4 info, CRow (&rows)[4], FRow (&frows)[4]) {
    const u32 m = static_cast<u32>(info.z);
    const u32 no = static_cast<u32>(info.w) & 0xffu;
    const u32 anchors = (static_cast<u32>(info.w) >> 8) & 0xffu;
#pragma unroll
    for (u32 q = 0; q < 4; ++q) {
        if (q < no) rows[q] = W.crow[4 * m + q];
        if (q < 2 * anchors) frows[q] = W.frow[4 * m + q];
    }
}
// rows / frows: the manager's rows as loaded by load_rows.
__device__ void solve_manager(const Ws& W, int4 info, CRow (&rows)[4], FRow (&frows)[4], float4 fric, const BodyState& s0,
                              const BodyState& s1, float ima, float imb, u32 c0, u32 c1, bool snap0, bool snap1,
                              bool fin0, bool fin1, u32 iter, u32 pos_iters, u32 total_iters) {
    const int i0 = info.x, i1 = info.y;
    const u32 m = static_cast<u32>(info.z);
    const u32 no = static_cast<u32>(info.w) & 0xffu;
    const u32 anchors = (static_cast<u32>(info.w) >> 8) & 0xffu;
    V va = s0.v, vb = s1.v, wa = s0.w, wb = s1.w;
    bool bad = false;
    float normal_impulse = 0.0f;
#pragma unroll
    for (u32 q = 0; q < 4; ++q) {
        if (q >= no) break;
        CRow& r = rows[q];
        const V n{r.nx, r.ny, r.nz};
        const V ra{r.rax, r.ray, r.raz}, rb{r.rbx, r.rby, r.rbz};
        const float nv = vdot(n, va) - vdot(n, vb) + vdot(wa, ra) - vdot(wb, rb);
        const float err = iter < pos_iters ? r.biased : r.unbiased;
        const float x = -nv * r.vm + err;
        const float dc = smax(x, -r.applied);
        const float next = smin(r.max_impulse, 1.0f * r.applied + dc);
        if (W.tape && iter < W.tape_iters)
            W.tape_c[static_cast<std::size_t>(iter) * 4 * W.nm + 4 * m + q] =
                r.max_impulse < 1.0f * r.applied + dc ? 2 : (x < -r.applied ? 0 : 1);
        const float delta = next - r.applied;
        const float da = delta * ima;
        const float db = delta * imb;
        va = {va.x + n.x * da, va.y + n.y * da, va.z + n.z * da};
        vb = {vb.x - n.x * db, vb.y - n.y * db, vb.z - n.z * db};
        wa = {wa.x + ra.x * delta, wa.y + ra.y * delta, wa.z + ra.z * delta};
        wb = {wb.x - rb.x * delta, wb.y - rb.y * delta, wb.z - rb.z * delta};
        r.applied = next;
        W.crow[4 * m + q].applied = next;
        normal_impulse += next;
    }
    if (W.tape && iter < W.tape_iters) W.tape_ni[static_cast<std::size_t>(iter) * W.nm + m] = normal_impulse;
    if ((iter >= pos_iters || pos_iters - iter <= 3u) && anchors) {
        const float coef = anchors == 2 ? 0.5f : 1.0f;
        const float max_static = fric.x * coef * normal_impulse;
        const float max_dynamic = fric.y * coef * normal_impulse;
        bool broken = false;
#pragma unroll
        for (u32 k = 0; k < 4; ++k) {
            if (k >= anchors * 2) break;
            FRow& fr = frows[k];
            const V t{fr.tx, fr.ty, fr.tz};
            const V ra{fr.rax, fr.ray, fr.raz}, rb{fr.rbx, fr.rby, fr.rbz};
            const float td = vdot(t, va) - vdot(t, vb) + vdot(wa, ra) - vdot(wb, rb);
            const float bias = iter < pos_iters ? fr.bias : 0.0f;
            const float cand = fr.applied + (fr.target - bias - td) * fr.vm;
            const bool clamped = fabsf(cand) > max_static;
            const float next = clamped ? smin(max_dynamic, smax(-max_dynamic, cand)) : cand;
            if (W.tape && iter < W.tape_iters)
                W.tape_f[static_cast<std::size_t>(iter) * 4 * W.nm + 4 * m + k] =
                    !clamped ? 0 : (max_dynamic < cand ? 2 : (cand < -max_dynamic ? 3 : 1));
            const float delta = next - fr.applied;
            const float da = delta * ima;
            const float db = delta * imb;
            va = {va.x + t.x * da, va.y + t.y * da, va.z + t.z * da};
            vb = {vb.x - t.x * db, vb.y - t.y * db, vb.z - t.z * db};
            wa = {wa.x + ra.x * delta, wa.y + ra.y * delta, wa.z + ra.z * delta};
            wb = {wb.x - rb.x * delta, wb.y - rb.y * delta, wb.z - rb.z * delta};
            fr.applied = next;
            W.frow[4 * m + k].applied = next;
            broken |= clamped;
        }
        if (iter + 1 == total_iters) W.patches[m].flags = (__float_as_uint(fric.z) & ~1u) | (broken ? 1u : 0u);
    }
    // a velocity component once non-finite stays so through later row
    // updates (sums with finite or non-finite terms), so the final check
    // equals a check after every row
    bad = !finite3(va) || !finite3(vb) || !finite3(wa) || !finite3(wb);
    if (bad) raise_flag(W, FYZZ_DY_NONFINITE);
    if (i0 >= 0) {
        publish_state(state_rsrc(W), static_cast<u32>(i0), c0, va, wa);
        const float4 l = make_float4(va.x, va.y, va.z, ima), a = make_float4(wa.x, wa.y, wa.z, 0.f);
        if (snap0) {
            W.mlin[i0] = l;
            W.mang[i0] = a;
        }
        if (fin0) {
            W.lin[i0] = l;
            W.ang[i0] = a;
        }
    }
    if (i1 >= 0) {
        publish_state(state_rsrc(W), static_cast<u32>(i1), c1, vb, wb);
        const float4 l = make_float4(vb.x, vb.y, vb.z, imb), a = make_float4(wb.x, wb.y, wb.z, 0.f);
        if (snap1) {
            W.mlin[i1] = l;
            W.mang[i1] = a;
        }
        if (fin1) {
            W.lin[i1] = l;
            W.ang[i1] = a;
        }
    }
}

// Position then velocity iterations over the partitions in order, motion
// snapshotted after the last position iteration (scene.cpp:9136-9146),
// without grid barriers. The sweep is the sequence (iteration, sorted
// position); within a partition no two managers share a dynamic body, so
// the Gauss-Seidel result depends only on each body seeing its managers'
// updates in that order. A body's state words count the updates applied
// to it; the manager at rank r among the body's d managers (by partition)
// runs in iteration t once all six words carry t*d + r, and publishes its
// result with t*d + r + 1. Thread k solves positions k, k + T, ...
// in every iteration (T threads), in sequence order, each as soon as it is
// ready (a lane waits without holding back the rest of its wave). The
// earliest unsolved (iteration, position) always has its predecessors done
// and its thread resident (cooperative launch), so the sweep always
// advances; waits are bounded regardless (a status flag, not a hang).
__global__ __launch_bounds__(kBlock) void k_pgs(Ws W) {
    const u32 tid = blockIdx.x * kBlock + threadIdx.x;
    const u32 nthreads = gridDim.x * kBlock;
    const u32 np = min(ld(&W.ctr[kNParts]), kMaxParts);
    const u32 pos_iters = ld(&W.ctr[kPosIters]);
    const u32 vel_iters = max(ld(&W.ctr[kVelIters]), 1u);
    const u32 total = pos_iters + vel_iters;
    const u32 n = np ? W.part_off[np] : 0u;
    u32 iter = 0, i = tid, spins = 0, idle = 0;
    const __amdgpu_buffer_rsrc_t R = state_rsrc(W);
    u64 wake0 = 0, wake1 = 0;  // earliest clock at which each body is polled again
    bool have = false, ready0 = false, ready1 = false;
    int4 info = make_int4(-1, -1, 0, 0);
    uint4 rk = make_uint4(0u, 0u, 0u, 0u);
    BodyState s0{}, s1{};
    float ima = 0.f, imb = 0.f;
    CRow rows[4];
    FRow frows[4];
    float4 fric = make_float4(0.f, 0.f, 0.f, 0.f);  // static and dynamic friction, patch flags
    // the position record of the thread's next position, loaded while the
    // current manager waits (positions repeat in every iteration)
    u32 next_pos = kNone;
    int4 next_info = info;
    uint4 next_rk = rk;
    for (;;) {
        bool progress = false;
        while (iter < total && i >= n) {
            ++iter;
            i = tid;
        }
        if (iter < total) {
            if (!have) {
                if (next_pos == i) {
                    info = next_info;
                    rk = next_rk;
                } else {
                    info = W.sinfo[i];
                    rk = W.srank[i];
                }
                have = true;
                ready0 = info.x < 0;
                ready1 = info.y < 0;
                wake0 = wake1 = 0;
                s0 = BodyState{};
                s1 = BodyState{};
                // inverse masses: constant through the sweep
                ima = info.x >= 0 ? W.lin[info.x].w : 0.f;
                imb = info.y >= 0 ? W.lin[info.y].w : 0.f;
                // the rows are loaded while the bodies are awaited (only
                // this thread writes them during the sweep)
                load_rows(W, info, rows, frows);
                {
                    // the patch's friction coefficients and flags (only
                    // this thread writes them during the sweep)
                    const fyzz_dy_patch& P = W.patches[info.z];
                    fric = make_float4(P.static_friction, P.dynamic_friction, __uint_as_float(P.flags), 0.f);
                }
                const u32 nx = i + nthreads < n ? i + nthreads : tid;
                if (nx < n) {
                    next_pos = nx;
                    next_info = W.sinfo[nx];
                    next_rk = W.srank[nx];
                }
            }
            // a body whose count reached the awaited value stays there
            // until this manager runs: its state is kept, not polled again.
            // A body still `need - seen` updates short is polled again only
            // after about that many solve hops: waiting lanes would
            // otherwise keep the memory system busy with polls that cannot
            // succeed, slowing the hops they wait for.
            const u32 need0 = iter * rk.z + rk.x, need1 = iter * rk.w + rk.y;
            const u64 now = clock64();
            // both bodies' first chunks in flight together, then the other
            // chunks of the bodies whose first chunk carries the awaited
            // count (the count it carries: the updates the body has had)
            const bool p0 = !ready0 && now >= wake0, p1 = !ready1 && now >= wake1;
            const bool polled = p0 || p1;
            const u32 o0 = static_cast<u32>(info.x) * kStateBytes, o1 = static_cast<u32>(info.y) * kStateBytes;
            Chunk a0, b0, c0, a1, b1, c1;
            if (p0) a0 = load_chunk(R, o0);
            if (p1) a1 = load_chunk(R, o1);
            const bool m0 = p0 && a0.y == need0, m1 = p1 && a1.y == need1;
            if (m0) {
                b0 = load_chunk(R, o0 + 16u);
                c0 = load_chunk(R, o0 + 32u);
            }
            if (m1) {
                b1 = load_chunk(R, o1 + 16u);
                c1 = load_chunk(R, o1 + 32u);
            }
            if (p0) {
                ready0 = m0 && take_state(a0, b0, c0, need0, s0);
                if (!ready0) wake0 = now + poll_delay(need0, a0.y);
            }
            if (p1) {
                ready1 = m1 && take_state(a1, b1, c1, need1, s1);
                if (!ready1) wake1 = now + poll_delay(need1, a1.y);
            }
            if (ready0 && ready1) {
                const bool snap = iter + 1 == pos_iters, last = iter + 1 == total;
                const bool end0 = rk.x + 1 == rk.z, end1 = rk.y + 1 == rk.w;
                solve_manager(W, info, rows, frows, fric, s0, s1, ima, imb, need0 + 1u, need1 + 1u, snap && end0,
                              snap && end1, last && end0, last && end1, iter, pos_iters, total);
                have = false;
                i += nthreads;
                progress = true;
                spins = 0;  // the budget bounds one wait, not the thread's whole sweep
            } else if (polled && ++spins > kSpinLimit) {
                raise_flag(W, FYZZ_DY_LOGIC);
                iter = total;  // give up on this thread's work
            }
        }
        if (!__ballot(iter < total)) break;
        idle = __ballot(progress) ? 0u : idle + 1u;
        wait_backoff(idle);
    }
}

// --------------------------------------------------------------- adjoint
//
// Reverse mode through one step's PGS and integration with the step's
// contact rows frozen (normals, lever arms, effective masses, biases and
// friction frames are constants): every row update is then piecewise
// linear in the velocities, with the branch recorded in the tape. Inputs:
// adjoints of the step's outputs (velocity, angular velocity, position);
// outputs: adjoints of its input velocities and of each body's dynamic
// friction. Orientation is not differentiated (it reaches the outputs only
// through the frozen rows).

__device__ inline V ld3(const float* p, u32 i) { return {p[3 * i], p[3 * i + 1], p[3 * i + 2]}; }

// adjoint state at the end of the sweeps
__global__ __launch_bounds__(kBlock) void k_adj_init(Ws W) {
    const u32 t = blockIdx.x * kBlock + threadIdx.x;
    if (t < W.nb) {
        const V av = ld3(W.adj_v_out, t), aw = ld3(W.adj_w_out, t);
        const float4 c0 = W.sii[3 * t], c1 = W.sii[3 * t + 1], c2 = W.sii[3 * t + 2];
        // v_out = v_pre + lin; w_out = w_pre + S ang (S symmetric)
        W.a_lin[t] = make_float4(av.x, av.y, av.z, 0.f);
        W.a_ang[t] = make_float4(c0.x * aw.x + c0.y * aw.y + c0.z * aw.z, c1.x * aw.x + c1.y * aw.y + c1.z * aw.z,
                                 c2.x * aw.x + c2.y * aw.y + c2.z * aw.z, 0.f);
        W.adj_friction[t] = 0.0f;
    }
    const u32 nm = mcount(W);
    if (t < 4 * nm) {
        W.a_capp[t] = 0.f;
        W.a_fapp[t] = 0.f;
        W.a_err[t] = 0.f;
        W.a_tgt[t] = 0.f;
    }
    if (t < nm) W.a_df[t] = 0.f;
}

__device__ void adjoint_manager(const Ws& W, u32 m, u32 iter, u32 pos_iters) {
    const fyzz_dy_manager& M = W.mg[m];
    const int i0 = M.body0, i1 = M.body1;
    const float4 z4 = make_float4(0.f, 0.f, 0.f, 0.f);
    const float ima = i0 >= 0 ? W.lin[i0].w : 0.f, imb = i1 >= 0 ? W.lin[i1].w : 0.f;
    V ava = f3(i0 >= 0 ? W.a_lin[i0] : z4), avb = f3(i1 >= 0 ? W.a_lin[i1] : z4);
    V awa = f3(i0 >= 0 ? W.a_ang[i0] : z4), awb = f3(i1 >= 0 ? W.a_ang[i1] : z4);
    const fyzz_dy_patch& P = W.patches[m];
    const u32 anchors = P.anchor_count;
    const std::size_t row_base = static_cast<std::size_t>(iter) * 4 * W.nm + 4 * m;
    float a_ni = 0.0f;
    if ((iter >= pos_iters || pos_iters - iter <= 3u) && anchors) {
        const float coef = anchors == 2 ? 0.5f : 1.0f;
        const float ni = W.tape_ni[static_cast<std::size_t>(iter) * W.nm + m];
        const float df = P.dynamic_friction;
        float a_maxd = 0.0f;
        for (int k = static_cast<int>(anchors * 2) - 1; k >= 0; --k) {
            const FRow& fr = W.frow[4 * m + k];
            const V t{fr.tx, fr.ty, fr.tz};
            const V ra{fr.rax, fr.ray, fr.raz}, rb{fr.rbx, fr.rby, fr.rbz};
            const float a_delta = vdot(t, ava) * ima - vdot(t, avb) * imb + vdot(ra, awa) - vdot(rb, awb);
            const float a_next = W.a_fapp[4 * m + k] + a_delta;
            float a_applied = -a_delta;
            const uint8_t br = W.tape_f[row_base + k];
            float a_cand = 0.0f;
            if (br <= 1)
                a_cand = a_next;
            else if (br == 2)
                a_maxd += a_next;
            else
                a_maxd -= a_next;
            a_applied += a_cand;
            W.a_tgt[4 * m + k] += fr.vm * a_cand;
            const float a_td = -fr.vm * a_cand;
            ava = vadd(ava, vmul(t, a_td));
            avb = vsub(avb, vmul(t, a_td));
            awa = vadd(awa, vmul(ra, a_td));
            awb = vsub(awb, vmul(rb, a_td));
            W.a_fapp[4 * m + k] = a_applied;
        }
        // max_dynamic = df * coef * ni
        W.a_df[m] += coef * ni * a_maxd;
        a_ni = df * coef * a_maxd;
    }
    const u32 no = W.order[m] & 0xffu;
    for (int q = static_cast<int>(no) - 1; q >= 0; --q) {
        const CRow& r = W.crow[4 * m + q];
        const V n{r.nx, r.ny, r.nz};
        const V ra{r.rax, r.ray, r.raz}, rb{r.rbx, r.rby, r.rbz};
        const float a_delta = vdot(n, ava) * ima - vdot(n, avb) * imb + vdot(ra, awa) - vdot(rb, awb);
        const float a_next = W.a_capp[4 * m + q] + a_delta + a_ni;
        float a_applied = -a_delta;
        if (W.tape_c[row_base + q] == 1) {
            a_applied += a_next;
            const float a_nv = -r.vm * a_next;
            W.a_err[4 * m + q] += a_next;
            ava = vadd(ava, vmul(n, a_nv));
            avb = vsub(avb, vmul(n, a_nv));
            awa = vadd(awa, vmul(ra, a_nv));
            awb = vsub(awb, vmul(rb, a_nv));
        }
        W.a_capp[4 * m + q] = a_applied;
    }
    if (i0 >= 0) {
        W.a_lin[i0] = make_float4(ava.x, ava.y, ava.z, 0.f);
        W.a_ang[i0] = make_float4(awa.x, awa.y, awa.z, 0.f);
    }
    if (i1 >= 0) {
        W.a_lin[i1] = make_float4(avb.x, avb.y, avb.z, 0.f);
        W.a_ang[i1] = make_float4(awb.x, awb.y, awb.z, 0.f);
    }
}

// The sweeps in reverse: iterations and partitions backwards. The motion
// snapshot taken after the last position iteration feeds the position
// update, so its adjoint (dt x adjoint of the position) joins the
// velocity-delta adjoint when the reverse reaches that point.
__global__ __launch_bounds__(kBlock) void k_pgs_adjoint(Ws W) {
    u32 gen = barrier_start(W);
    const u32 tid = blockIdx.x * kBlock + threadIdx.x;
    const u32 nthreads = gridDim.x * kBlock;
    const u32 np = ld(&W.ctr[kNParts]);
    const u32 pos_iters = ld(&W.ctr[kPosIters]);
    const u32 vel_iters = max(ld(&W.ctr[kVelIters]), 1u);
    const u32 total = pos_iters + vel_iters;
    if (total > W.tape_iters) {  // the tape does not hold every iteration
        if (tid == 0) raise_flag(W, FYZZ_DY_OVERFLOW);
        return;
    }
    for (u32 it = total; it-- > 0;) {
        if (it + 1 == pos_iters) {
            for (u32 i = tid; i < W.nb; i += nthreads)
                if (W.has_row[i]) {
                    const V ap = ld3(W.adj_p_out, i);
                    float4 a = W.a_lin[i];
                    a.x += W.dt * ap.x;
                    a.y += W.dt * ap.y;
                    a.z += W.dt * ap.z;
                    W.a_lin[i] = a;
                }
            grid_sync(W, gen);
        }
        for (u32 p = np; p-- > 0;) {
            const u32 lo = W.part_off[p], hi = W.part_off[p + 1];
            for (u32 i = lo + tid; i < hi; i += nthreads) adjoint_manager(W, W.sorted[i], it, pos_iters);
            grid_sync(W, gen);
        }
    }
}

// Rows' velocity terms back to the bodies' pre-integrated velocities; the
// combined dynamic friction back to each body's (max(0, (df0 + df1) / 2)).
__global__ __launch_bounds__(kBlock) void k_adj_rows(Ws W) {
    const u32 m = blockIdx.x * kBlock + threadIdx.x;
    if (m >= mcount(W) || W.part[m] >= kMaxParts) return;
    const fyzz_dy_manager& M = W.mg[m];
    const int i0 = M.body0, i1 = M.body1;
    V ava{0.f, 0.f, 0.f}, avb{0.f, 0.f, 0.f}, awa{0.f, 0.f, 0.f}, awb{0.f, 0.f, 0.f};
    const u32 no = W.order[m] & 0xffu;
    for (u32 q = 0; q < no; ++q) {
        const CRow& r = W.crow[4 * m + q];
        const float4 x0 = W.craw[2 * (4 * m + q)], x1 = W.craw[2 * (4 * m + q) + 1];
        const float a_vrel = x0.w * r.vm * W.a_err[4 * m + q];
        const V n{r.nx, r.ny, r.nz};
        ava = vadd(ava, vmul(n, a_vrel));
        avb = vsub(avb, vmul(n, a_vrel));
        awa = vadd(awa, vmul(f3(x0), a_vrel));
        awb = vsub(awb, vmul(f3(x1), a_vrel));
    }
    const fyzz_dy_patch& P = W.patches[m];
    for (u32 k = 0; k < 2 * P.anchor_count; ++k) {
        const FRow& fr = W.frow[4 * m + k];
        const float a_t = W.a_tgt[4 * m + k];
        const V t{fr.tx, fr.ty, fr.tz};
        // target = -(t . (va - vb) + (ra x t) . wa - (rb x t) . wb)
        ava = vsub(ava, vmul(t, a_t));
        avb = vadd(avb, vmul(t, a_t));
        awa = vsub(awa, vmul(f3(W.fraw[2 * (4 * m + k)]), a_t));
        awb = vadd(awb, vmul(f3(W.fraw[2 * (4 * m + k) + 1]), a_t));
    }
    const float a_df = P.dynamic_friction > 0.0f ? 0.5f * W.a_df[m] : 0.0f;
    // each side's terms, summed per body by k_adj_bodies in the body's
    // incidence order (no atomics: the gradients are the same every run)
    W.a_side[4 * m] = make_float4(ava.x, ava.y, ava.z, a_df);
    W.a_side[4 * m + 1] = make_float4(awa.x, awa.y, awa.z, 0.f);
    W.a_side[4 * m + 2] = make_float4(avb.x, avb.y, avb.z, a_df);
    W.a_side[4 * m + 3] = make_float4(awb.x, awb.y, awb.z, 0.f);
}

// Pre-integration backwards: v_pre = lm (v_in + g dt), w_pre = am w_in, and
// v_pre reaches the outputs directly (v_out, and p_out through dt).
__global__ __launch_bounds__(kBlock) void k_adj_bodies(Ws W) {
    const u32 i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= W.nb) return;
    const fyzz_dy_body& b = W.bodies[i];
    const V av = ld3(W.adj_v_out, i), aw = ld3(W.adj_w_out, i), ap = ld3(W.adj_p_out, i);
    if (b.flags & 1u) {  // asleep: the step leaves it unchanged
        for (int k = 0; k < 3; ++k) {
            W.adj_v_in[3 * i + k] = (&av.x)[k];
            W.adj_w_in[3 * i + k] = (&aw.x)[k];
        }
        return;
    }
    // the rows' terms of the body's managers, in incidence order (sorted by
    // manager when the partitions were built)
    V a_rows_v{0.f, 0.f, 0.f}, a_rows_w{0.f, 0.f, 0.f};
    float a_fric = 0.0f;
    if (mcount(W) && W.has_row[i]) {
        for (u32 k = W.off[i]; k < W.off[i + 1]; ++k) {
            const u32 e = W.inc[k], m = e >> 1, side = e & 1u;
            if (W.part[m] >= kMaxParts) continue;
            const float4 v = W.a_side[4 * m + 2 * side], w = W.a_side[4 * m + 2 * side + 1];
            a_rows_v = vadd(a_rows_v, f3(v));
            a_rows_w = vadd(a_rows_w, f3(w));
            a_fric += v.w;
        }
    }
    W.adj_friction[i] = a_fric;
    const V a_vpre = vadd(vadd(av, vmul(ap, W.dt)), a_rows_v);
    const V a_wpre = vadd(aw, a_rows_w);
    const float lm = smax(1.0f - b.linear_damping * W.dt, 0.0f);
    const float am = smax(1.0f - b.angular_damping * W.dt, 0.0f);
    W.adj_v_in[3 * i] = lm * a_vpre.x;
    W.adj_v_in[3 * i + 1] = lm * a_vpre.y;
    W.adj_v_in[3 * i + 2] = lm * a_vpre.z;
    W.adj_w_in[3 * i] = am * a_wpre.x;
    W.adj_w_in[3 * i + 1] = am * a_wpre.y;
    W.adj_w_in[3 * i + 2] = am * a_wpre.z;
}

// ------------------------------------------------------------- integrate

// sinf and cosf as the C library the native scene links computes them
// (glibc 2.28 and later ship the ARM optimized-routines sinf/cosf): the
// float argument widened to double; below 0.75 (a test on the top 12 bits
// against pi/4) a double polynomial in x^2 rounded once to float, below
// 2^-12 the argument itself (sin) or 1 (cos); up to 120 a reduction by the
// nearest multiple of pi/2 first. The published minimax coefficients are
// restated here; every float in [2^-13, 120) gives the host library's
// result but 17 arguments above 17 (the nearest-multiple reduction of a
// value close to a multiple of pi/2). Beyond 120 (angular speeds above
// 28,800 rad/s at 120 Hz) the device's own functions are used.
namespace libm {
constexpr double kS1 = -0x1.555545995a603p-3, kS2 = 0x1.1107605230bc4p-7, kS3 = -0x1.994eb3774cf24p-13;
constexpr double kC0 = 0x1p0, kC1 = -0x1.ffffffd0c621cp-2, kC2 = 0x1.55553e1068f19p-5,
                 kC3 = -0x1.6c087e89a359dp-10, kC4 = 0x1.99343027bf8c3p-16;
constexpr double kHalfPiInv = 0x1.45F306DC9C883p+23, kHalfPi = 0x1.921FB54442D18p0;

__device__ inline uint32_t top12(float f) { return (__float_as_uint(f) >> 20) & 0x7ffu; }

// cos_branch: the even polynomial (cos), else the odd one (sin); negate:
// the coefficients' sign flipped (the odd quadrants of the reduction)
__device__ inline double poly(double x, double x2, bool cos_branch, bool negate) {
    if (!cos_branch) {
        const double x3 = x * x2;
        const double s1 = kS2 + x2 * kS3;
        const double x7 = x3 * x2;
        const double s = x + x3 * kS1;
        return s + x7 * s1;
    }
    const double g = negate ? -1.0 : 1.0;
    const double x4 = x2 * x2;
    const double c2 = g * kC3 + x2 * (g * kC4);
    const double c1 = g * kC0 + x2 * (g * kC1);
    const double x6 = x4 * x2;
    const double c = c1 + x4 * (g * kC2);
    return c + x6 * c2;
}

__device__ inline float sin_cos(float y, bool want_cos) {
    const double x = y;
    if (top12(y) < top12(0x1.921FB6p-1f)) {
        if (top12(y) < top12(0x1p-12f)) return want_cos ? 1.0f : y;
        return static_cast<float>(poly(x, x * x, want_cos, false));
    }
    if (top12(y) < top12(120.0f)) {
        const double r = x * kHalfPiInv;
        const int n = (static_cast<int>(r) + 0x800000) >> 24;
        const double xr = x - n * kHalfPi;
        const double sgn = ((n & 3) == 1 || (n & 3) == 2) ? -1.0 : 1.0;
        const int q = want_cos ? n + 1 : n;
        return static_cast<float>(poly(xr * sgn, xr * xr, (q & 1) != 0, (q & 2) != 0));
    }
    return want_cos ? static_cast<float>(cos(static_cast<double>(y))) : sinf(y);
}
}  // namespace libm

// scene.cpp:111-146
__device__ bool integrate_quaternion(Q pose, V w, float dt, Q& out) {
    float ww = w.x * w.x + w.y * w.y + w.z * w.z;
    if (ww != 0.0f) {
        ww = sqrtf(ww);
        const float max_w = 1.0e7f;
        if (ww > max_w) {
            const float s = max_w / ww;
            w = {w.x * s, w.y * s, w.z * s};
            ww = max_w;
        }
        const float v = dt * ww * 0.5f;
        const float s = libm::sin_cos(v, false) / ww;
        const float c = libm::sin_cos(v, true);
        const Q dq = qmul(Q{w.x * s, w.y * s, w.z * s, 0.0f}, pose);
        const Q nx{dq.x + pose.x * c, dq.y + pose.y * c, dq.z + pose.z * c, dq.w + pose.w * c};
        const float len = sqrtf(nx.x * nx.x + nx.y * nx.y + nx.z * nx.z + nx.w * nx.w);
        if (!(len > 0.0f) || !isfinite(len)) return false;
        pose = {nx.x / len, nx.y / len, nx.z / len, nx.w / len};
    }
    out = pose;
    return true;
}

__device__ inline V apply_sii(V w, float4 c0, float4 c1, float4 c2, V s) {
    return {w.x + c0.x * s.x + c1.x * s.y + c2.x * s.z, w.y + c0.y * s.x + c1.y * s.y + c2.y * s.z,
            w.z + c0.z * s.x + c1.z * s.y + c2.z * s.z};
}

// publish (scene.cpp:9509-9549), integrate and sleep check (10508-10707)
__global__ __launch_bounds__(kBlock) void k_integrate(Ws W) {
    const u32 i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= W.nb) return;
    fyzz_dy_body& B = W.bodies[i];
    if (B.flags & 1u) return;
    const float dt = W.dt;
    V vel = f3(W.pv[i]), ang = f3(W.pw[i]);
    V motion_lin = vel, motion_ang = ang;
    if (W.has_row[i]) {
        const V rv = vel, rw = ang;
        const float4 c0 = W.sii[3 * i], c1 = W.sii[3 * i + 1], c2 = W.sii[3 * i + 2];
        const V ld = f3(W.lin[i]), as = f3(W.ang[i]);
        const V ml = f3(W.mlin[i]), ma = f3(W.mang[i]);
        vel = vadd(rv, ld);
        ang = apply_sii(rw, c0, c1, c2, as);
        motion_lin = vadd(rv, ml);
        motion_ang = apply_sii(rw, c0, c1, c2, ma);
        if (!finite3(vel) || !finite3(ang)) raise_flag(W, FYZZ_DY_NONFINITE);
    }
    const V p0 = v3(B.position);
    const Q q0 = q4(B.orientation);
    const V center = mass_position(p0, q0);
    V p{center.x + motion_lin.x * dt, center.y + motion_lin.y * dt, center.z + motion_lin.z * dt};
    Q q;
    if (!integrate_quaternion(mass_orientation(q0), motion_ang, dt, q)) {
        raise_flag(W, FYZZ_DY_NONFINITE);
        q = q0;
    }
    q = qmul(q, Q{-0.0f, -0.0f, -0.0f, 1.0f});
    const V off = srot(q, V{0.f, 0.f, 0.f});
    p = vsub(p, off);
    if (!finite3(p) || !finite4(q)) raise_flag(W, FYZZ_DY_NONFINITE);
    // sleep check (scene.cpp:10640-10705)
    const float wake_reset = 20.0f * 0.02f;
    const float energy_threshold = 5.0e-5f * 10.0f * 10.0f;
    float wc = B.wake_counter;
    V lin_acc = v3(B.sleep_linear_acc), ang_acc = v3(B.sleep_angular_acc);
    if (wc < wake_reset * 0.5f || wc < dt) {
        const V la = srot_inv(q, motion_ang);
        const V l2 = vadd(lin_acc, motion_lin);
        const V a2 = vadd(ang_acc, la);
        const V ii = v3(B.inverse_inertia);
        const V inertia{ii.x > 0.0f ? 1.0f / ii.x : 1.0f, ii.y > 0.0f ? 1.0f / ii.y : 1.0f,
                        ii.z > 0.0f ? 1.0f / ii.z : 1.0f};
        const float ang_e =
            (a2.x * a2.x * inertia.x + a2.y * a2.y * inertia.y + a2.z * a2.z * inertia.z) * B.inverse_mass;
        const float lin_e = l2.x * l2.x + l2.y * l2.y + l2.z * l2.z;
        const float ne = 0.5f * (ang_e + lin_e);
        const float cluster = 1.0f + float(B.counted_interactions);
        const float th = cluster * energy_threshold;
        if (ne >= th) {
            lin_acc = {0.f, 0.f, 0.f};
            ang_acc = {0.f, 0.f, 0.f};
            const float factor = th == 0.0f ? 2.0f : smin(ne / th, 2.0f);
            wc = factor * 0.5f * wake_reset + dt * (cluster - 1.0f);
        } else {
            lin_acc = l2;
            ang_acc = a2;
            wc = smax(wc - dt, 0.0f);
        }
    } else {
        wc = smax(wc - dt, 0.0f);
    }
    const bool ready = wc == 0.0f;
    if (ready) {
        lin_acc = {0.f, 0.f, 0.f};
        ang_acc = {0.f, 0.f, 0.f};
    }
    put3(B.position, p);
    B.orientation[0] = q.x;
    B.orientation[1] = q.y;
    B.orientation[2] = q.z;
    B.orientation[3] = q.w;
    put3(B.linear_velocity, vel);
    put3(B.angular_velocity, ang);
    B.wake_counter = wc;
    put3(B.sleep_linear_acc, lin_acc);
    put3(B.sleep_angular_acc, ang_acc);
    B.flags = (B.flags & ~2u) | (ready ? 2u : 0u);
}

// impulses per manifold contact and the partition of each manager; the
// step's contact and friction row counts
// (the impulses land in their contact slots through selects, and the row
// totals are summed per workgroup in LDS: one global atomic per workgroup
// instead of one per wave on the same two words)
__global__ __launch_bounds__(kBlock) void k_outputs(Ws W) {
    __shared__ u32 s_rows, s_frows;
    if (threadIdx.x == 0) s_rows = s_frows = 0;
    __syncthreads();
    const u32 m = blockIdx.x * kBlock + threadIdx.x;
    u32 rows = 0, frows = 0;
    if (m < mcount(W)) {
        if (W.parts_out) W.parts_out[m] = W.part[m];
        float out[4] = {0.f, 0.f, 0.f, 0.f};
        if (W.with_rows && W.part[m] < kMaxParts) {
            const u32 o = W.order[m];
            const u32 no = min(o & 0xffu, 4u);
#pragma unroll
            for (u32 q = 0; q < 4; ++q) {
                if (q >= no) continue;
                const u32 slot = (o >> (8 + 4 * q)) & 0xfu;
                const float v = W.crow[4 * m + q].applied;
#pragma unroll
                for (u32 k = 0; k < 4; ++k) out[k] = slot == k ? v : out[k];
            }
            rows = no;
            frows = 2u * W.patches[m].anchor_count;
        }
        if (W.impulses)
            for (u32 q = 0; q < 4; ++q) W.impulses[4 * m + q] = out[q];
    }
    for (int d = kWave / 2; d > 0; d >>= 1) {
        rows += __shfl_xor(rows, d, kWave);
        frows += __shfl_xor(frows, d, kWave);
    }
    if (lane_id() == 0 && (rows || frows)) {
        atomicAdd(&s_rows, rows);
        atomicAdd(&s_frows, frows);
    }
    __syncthreads();
    if (threadIdx.x == 0 && (s_rows || s_frows)) {
        atomicAdd(&W.ctr[kContacts], s_rows);
        atomicAdd(&W.ctr[kFrictionRows], s_frows);
    }
}

// Scene::write_back_contacts' friction part (scene.cpp:9339-9388) for
// the last step's managers: each anchor's world position on body0 (the
// manager's first body, whose frame holds the patch's body0 anchors: the
// mass pose the rows were prepared with, centre + rotated local anchor)
// and its world impulse, the two tangent rows' tangent x applied force
// summed in double.
__global__ __launch_bounds__(kBlock) void k_anchors(Ws W, fyzz_friction_anchors* out) {
    const u32 m = blockIdx.x * kBlock + threadIdx.x;
    if (m >= mcount(W)) return;
    fyzz_friction_anchors r{};
    if (W.with_rows && W.part[m] < kMaxParts) {
        const fyzz_dy_patch& P = W.patches[m];
        const u32 na = min(P.anchor_count, 2u);
        r.anchor_count = na;
        r.broken = P.flags & 1u;
        const int b0 = W.mg[m].body0;
        const float4 cp = W.rpos[b0 >= 0 ? b0 : 0];
        const float4 cq = W.rq[b0 >= 0 ? b0 : 0];
        const Q q{cq.x, cq.y, cq.z, cq.w};
        for (u32 k = 0; k < na; ++k) {
            const V local{P.body0_anchors[k][0], P.body0_anchors[k][1], P.body0_anchors[k][2]};
            const V rot = srot(q, local);
            r.position[k][0] = cp.x + rot.x;
            r.position[k][1] = cp.y + rot.y;
            r.position[k][2] = cp.z + rot.z;
            const FRow& x = W.frow[4 * m + 2 * k];
            const FRow& y = W.frow[4 * m + 2 * k + 1];
            r.impulse[k][0] = static_cast<double>(x.tx) * x.applied + static_cast<double>(y.tx) * y.applied;
            r.impulse[k][1] = static_cast<double>(x.ty) * x.applied + static_cast<double>(y.ty) * y.applied;
            r.impulse[k][2] = static_cast<double>(x.tz) * x.applied + static_cast<double>(y.tz) * y.applied;
        }
    }
    out[m] = r;
}

}  // namespace fyzz_hip::dy

using namespace fyzz_hip;

struct fyzz_hip_dy {
    fyzz_hip_context* ctx = nullptr;
    u32 body_cap = 0, mgr_cap = 0;
    int coop_blocks_partition = 0, coop_blocks_pgs = 0, coop_blocks_adjoint = 0;
    bool plain_launch = false;
    DevBuf<float4> pv, pw, lin, ang, mlin, mang, rpos, rq, sii;
    DevBuf<u32> has_row, deg, off, fill, max_dyn, body_mask;
    DevBuf<u32> inc, pred, part, pend, sorted, order;
    DevBuf<uint8_t> dd;
    DevBuf<u64> entry;
    DevBuf<int4> sinfo;
    DevBuf<uint4> srank;
    DevBuf<u32> erank, pos_of;
    DevBuf<u64> vst;
    DevBuf<dy::CRow> crow;
    DevBuf<dy::FRow> frow;
    DevBuf<u32> part_count, part_off, part_fill, blocksums, ctr, flags, bar;
    // tape mode
    bool tape = false;
    bool taped_step = false;  // the last solved step recorded a tape
    DevBuf<uint8_t> tape_c, tape_f;
    DevBuf<float> tape_ni, a_capp, a_fapp, a_err, a_tgt, a_df;
    DevBuf<float4> craw, fraw, a_lin, a_ang, a_side;
    dy::Ws ws{};
    bool prepared = false;
};

namespace {

int resident_blocks(const void* kernel, int device) {
    int per_cu = 0, cus = 0;
    FYZZ_HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, dy::kBlock, 0));
    FYZZ_HIP_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device));
    FYZZ_REQUIRE(per_cu > 0 && cus > 0, kDeviceError, "dy: cooperative kernel does not fit a compute unit");
    return per_cu * cus;
}
int coop_blocks(const void* kernel, int device, int wanted = dy::kCoopBlocks) {
    // 64 resident workgroups: a grid barrier costs ~3 us there, against
    // ~17 us at 512 (the barrier is the Dy kernels' critical path)
    // resident capacity, and the grid barrier's group counters
    return std::min({resident_blocks(kernel, device), wanted, static_cast<int>(dy::kBarGroup * (dy::kBarWords - 2))});
}

// Cooperative launch (the runtime checks that every workgroup is resident).
// Inside a captured graph the kernels are launched as plain kernels with
// the same small grid (far below the device's resident capacity); the grid
// barrier's bounded spin turns a non-resident grid into a status flag.
void launch_coop(const void* kernel, int blocks, dy::Ws& ws, hipStream_t st, bool plain) {
    void* args[] = {&ws};
    if (plain)
        FYZZ_HIP_CHECK(hipLaunchKernel(kernel, dim3(blocks), dim3(dy::kBlock), args, 0, st));
    else
        FYZZ_HIP_CHECK(hipLaunchCooperativeKernel(kernel, dim3(blocks), dim3(dy::kBlock), args, 0, st));
}

void dy_prepare(fyzz_hip_dy* d, float dt, uint32_t nb, fyzz_dy_body* bodies, uint32_t ns,
                const fyzz_dy_static* statics, uint32_t nm, const fyzz_dy_manager* managers,
                const fyzz_dy_patch* history, uint32_t nh, fyzz_dy_patch* patches, bool partition_only,
                const u32* d_nm = nullptr) {
    FYZZ_REQUIRE(nb <= d->body_cap, kOverflow, "dy: body count exceeds the capacity given at creation");
    FYZZ_REQUIRE(nm <= d->mgr_cap, kOverflow, "dy: manager count exceeds the capacity given at creation");
    FYZZ_REQUIRE(dt > 0.0f, kInvalidArgument, "dy: dt must be positive");
    FYZZ_REQUIRE(nb == 0 || bodies, kInvalidArgument, "dy: null bodies");
    FYZZ_REQUIRE(nm == 0 || managers, kInvalidArgument, "dy: null managers");
    FYZZ_REQUIRE(ns == 0 || statics, kInvalidArgument, "dy: null statics");
    FYZZ_REQUIRE(nh == 0 || history, kInvalidArgument, "dy: null history");
    FYZZ_REQUIRE(partition_only || nm == 0 || patches, kInvalidArgument, "dy: null patch output");
    dy::Ws& w = d->ws;
    w.dt = dt;
    w.nb = nb;
    w.ns = partition_only ? 0xffffffffu : ns;  // partitions never read statics
    w.nm = nm;
    w.nh = nh;
    w.bodies = bodies;
    w.statics = statics;
    w.mg = managers;
    w.history = history;
    w.patches = patches;
    w.d_nm = d_nm;
    w.with_rows = partition_only ? 0u : 1u;
    w.tape = (d->tape && !partition_only) ? 1u : 0u;
    d->taped_step = false;
    hipStream_t st = d->ctx->stream;
    auto& prof = d->ctx->prof;
    {
        ProfScope ps(&prof, "dy.bodies", st);
        hipLaunchKernelGGL(dy::k_clear, dim3(1), dim3(64), 0, st, w);
        if (nb) hipLaunchKernelGGL(dy::k_pre, dim3(grid_for(nb, dy::kBlock)), dim3(dy::kBlock), 0, st, w);
        if (nm) hipLaunchKernelGGL(dy::k_mark, dim3(grid_for(nm, dy::kBlock)), dim3(dy::kBlock), 0, st, w);
        if (nb) hipLaunchKernelGGL(dy::k_rows, dim3(grid_for(nb, dy::kBlock)), dim3(dy::kBlock), 0, st, w);
        FYZZ_HIP_CHECK(hipGetLastError());
    }
    {
        ProfScope ps(&prof, "dy.partition", st);
        const dim3 blk(dy::kBlock);
        const dim3 gb(grid_for(std::max<u32>(nb, dy::kMaxParts), dy::kBlock)), gm(grid_for(nm, dy::kBlock));
        hipLaunchKernelGGL(dy::k_part_clear, gb, blk, 0, st, w);
        if (nm) {
            hipLaunchKernelGGL(dy::k_part_degrees, gm, blk, 0, st, w);
            if (nb)
                sel::exclusive_scan(st, nb, d->ctr.get() + dy::kNbWord, dy::DegreeOf{d->deg.get()}, d->off.get(),
                                    d->blocksums.get(), d->off.get() + nb);
            hipLaunchKernelGGL(dy::k_part_fill, gm, blk, 0, st, w);
            if (nb) hipLaunchKernelGGL(dy::k_part_chain, dim3(grid_for(nb, dy::kBlock)), blk, 0, st, w);
            FYZZ_HIP_CHECK(hipGetLastError());
            launch_coop(reinterpret_cast<const void*>(&dy::k_part_color), d->coop_blocks_partition, w, st,
                        d->plain_launch || !d->ctx->cooperative);
            hipLaunchKernelGGL(dy::k_part_maxdyn, gm, blk, 0, st, w);
            if (nb) hipLaunchKernelGGL(dy::k_part_static, dim3(grid_for(nb, dy::kBlock)), blk, 0, st, w);
            hipLaunchKernelGGL(dy::k_part_sizes, gm, blk, 0, st, w);
        }
        hipLaunchKernelGGL(dy::k_part_offsets, dim3(1), blk, 0, st, w);
        if (nm) hipLaunchKernelGGL(dy::k_part_scatter, gm, blk, 0, st, w);
        FYZZ_HIP_CHECK(hipGetLastError());
    }
    if (!partition_only && nm) {
        ProfScope ps(&prof, "dy.prep", st);
        hipLaunchKernelGGL(dy::k_prep, dim3(grid_for(nm, 128)), dim3(128), 0, st, w);
        FYZZ_HIP_CHECK(hipGetLastError());
    }
    d->prepared = !partition_only;
}

void dy_solve(fyzz_hip_dy* d) {
    FYZZ_REQUIRE(d->prepared, kLogicError, "dy_solve: no prepared step");
    d->taped_step = d->ws.tape != 0;
    if (!d->ws.nm) return;
    ProfScope ps(&d->ctx->prof, "dy.pgs", d->ctx->stream);
    launch_coop(reinterpret_cast<const void*>(&dy::k_pgs), d->coop_blocks_pgs, d->ws, d->ctx->stream,
                d->plain_launch || !d->ctx->cooperative);
}

void dy_integrate(fyzz_hip_dy* d, float* impulses, uint32_t* partitions) {
    FYZZ_REQUIRE(d->prepared, kLogicError, "dy_integrate: no prepared step");
    hipStream_t st = d->ctx->stream;
    d->ws.impulses = impulses;
    d->ws.parts_out = partitions;
    ProfScope ps(&d->ctx->prof, "dy.integrate", st);
    if (d->ws.nb)
        hipLaunchKernelGGL(dy::k_integrate, dim3(grid_for(d->ws.nb, dy::kBlock)), dim3(dy::kBlock), 0, st, d->ws);
    if (d->ws.nm)
        hipLaunchKernelGGL(dy::k_outputs, dim3(grid_for(d->ws.nm, dy::kBlock)), dim3(dy::kBlock), 0, st, d->ws);
    FYZZ_HIP_CHECK(hipGetLastError());
    d->prepared = false;
}

void dy_set_tape(fyzz_hip_dy* d, bool enabled) {
    if (enabled && !d->tape_c.p) {
        const std::size_t nm = std::max<u32>(d->mgr_cap, 1), nb = std::max<u32>(d->body_cap, 1);
        constexpr u32 kTapeIters = 8;
        d->tape_c.reserve(kTapeIters * 4 * nm);
        d->tape_f.reserve(kTapeIters * 4 * nm);
        d->tape_ni.reserve(kTapeIters * nm);
        for (auto* b : {&d->a_capp, &d->a_fapp, &d->a_err, &d->a_tgt}) b->reserve(4 * nm);
        d->a_df.reserve(nm);
        d->craw.reserve(8 * nm);
        d->fraw.reserve(8 * nm);
        for (auto* b : {&d->a_lin, &d->a_ang}) b->reserve(nb);
        d->a_side.reserve(4 * nm);
        dy::Ws& w = d->ws;
        w.tape_iters = kTapeIters;
        w.tape_c = d->tape_c.p;
        w.tape_f = d->tape_f.p;
        w.tape_ni = d->tape_ni.p;
        w.craw = d->craw.p;
        w.fraw = d->fraw.p;
        w.a_capp = d->a_capp.p;
        w.a_fapp = d->a_fapp.p;
        w.a_err = d->a_err.p;
        w.a_tgt = d->a_tgt.p;
        w.a_df = d->a_df.p;
        w.a_lin = d->a_lin.p;
        w.a_ang = d->a_ang.p;
        w.a_side = d->a_side.p;
    }
    d->tape = enabled;
}

void dy_adjoint(fyzz_hip_dy* d, const float* adj_v_out, const float* adj_w_out, const float* adj_p_out,
                float* adj_v_in, float* adj_w_in, float* adj_friction) {
    FYZZ_REQUIRE(d->taped_step, kLogicError, "dy_adjoint: the last step was not solved with the tape on");
    FYZZ_REQUIRE(!d->prepared, kLogicError, "dy_adjoint: the last step has not been integrated");
    dy::Ws& w = d->ws;
    FYZZ_REQUIRE(w.nb == 0 || (adj_v_out && adj_w_out && adj_p_out && adj_v_in && adj_w_in && adj_friction),
                 kInvalidArgument, "dy_adjoint: null device pointer");
    w.adj_v_out = adj_v_out;
    w.adj_w_out = adj_w_out;
    w.adj_p_out = adj_p_out;
    w.adj_v_in = adj_v_in;
    w.adj_w_in = adj_w_in;
    w.adj_friction = adj_friction;
    hipStream_t st = d->ctx->stream;
    ProfScope ps(&d->ctx->prof, "dy.adjoint", st);
    const u32 span = std::max<u32>(w.nb, 4 * w.nm);
    if (span) hipLaunchKernelGGL(dy::k_adj_init, dim3(grid_for(span, dy::kBlock)), dim3(dy::kBlock), 0, st, w);
    if (w.nm) {
        launch_coop(reinterpret_cast<const void*>(&dy::k_pgs_adjoint), d->coop_blocks_adjoint, w, st,
                    d->plain_launch || !d->ctx->cooperative);
        hipLaunchKernelGGL(dy::k_adj_rows, dim3(grid_for(w.nm, dy::kBlock)), dim3(dy::kBlock), 0, st, w);
    }
    if (w.nb) hipLaunchKernelGGL(dy::k_adj_bodies, dim3(grid_for(w.nb, dy::kBlock)), dim3(dy::kBlock), 0, st, w);
    FYZZ_HIP_CHECK(hipGetLastError());
}

}  // namespace

namespace fyzz_hip::dy {
const u32* counters(fyzz_hip_dy* d) { return d->ctr.get(); }
void set_plain_launch(fyzz_hip_dy* d, bool plain) { d->plain_launch = plain; }
void set_coop_blocks(fyzz_hip_dy* d, int partition_blocks, int pgs_blocks) {
    const int cap_p = coop_blocks(reinterpret_cast<const void*>(&k_part_color), d->ctx->device, 1 << 30);
    // the sweep has no grid barrier: its cap is the resident capacity
    const int cap_s = resident_blocks(reinterpret_cast<const void*>(&k_pgs), d->ctx->device);
    FYZZ_REQUIRE(partition_blocks > 0 && pgs_blocks > 0, kInvalidArgument, "dy: workgroup counts must be positive");
    d->coop_blocks_partition = std::min(partition_blocks, cap_p);
    d->coop_blocks_pgs = std::min(pgs_blocks, cap_s);
    FYZZ_REQUIRE(2 * static_cast<std::size_t>(std::max(d->coop_blocks_partition, d->coop_blocks_pgs)) + 2 <=
                     d->blocksums.cap,
                 kInvalidArgument, "dy: workgroup count beyond the workspace");
}
void effective_blocks(int device, int partition_blocks, int pgs_blocks, int* partition_out, int* pgs_out) {
    const void* color = reinterpret_cast<const void*>(&k_part_color);
    const void* pgs = reinterpret_cast<const void*>(&k_pgs);
    *partition_out = partition_blocks > 0 ? coop_blocks(color, device, partition_blocks) : coop_blocks(color, device, 128);
    *pgs_out = pgs_blocks > 0 ? std::min(pgs_blocks, resident_blocks(pgs, device)) : coop_blocks(pgs, device, kPgsBlocks);
}
void coop_blocks_of(fyzz_hip_dy* d, int* partition_blocks, int* pgs_blocks) {
    *partition_blocks = d->coop_blocks_partition;
    *pgs_blocks = d->coop_blocks_pgs;
}

void step_device(fyzz_hip_dy* d, float dt, u32 nb, fyzz_dy_body* bodies, u32 ns, const fyzz_dy_static* statics,
                 u32 nm_cap, const u32* d_nm, const fyzz_dy_manager* managers, const fyzz_dy_patch* history,
                 u32 nh, fyzz_dy_patch* patches, float* impulses, u32* partitions) {
    dy_prepare(d, dt, nb, bodies, ns, statics, nm_cap, managers, history, nh, patches, false, d_nm);
    dy_solve(d);
    dy_integrate(d, impulses, partitions);
}
}  // namespace fyzz_hip::dy

extern "C" {

int fyzz_hip_dy_create(fyzz_hip_context* ctx, uint32_t body_capacity, uint32_t manager_capacity,
                       fyzz_hip_dy** out) {
    return guarded(ctx, [&] {
        FYZZ_REQUIRE(out, kInvalidArgument, "dy_create: null out");
        FYZZ_REQUIRE(manager_capacity < (1u << 30), kInvalidArgument, "dy_create: manager capacity too large");
        // the sweep addresses a body's 64-byte state line with a 32-bit offset
        FYZZ_REQUIRE(body_capacity <= (1u << 26), kInvalidArgument, "dy_create: body capacity too large");
        auto d = std::make_unique<fyzz_hip_dy>();
        d->ctx = ctx;
        d->body_cap = body_capacity;
        d->mgr_cap = manager_capacity;
        const std::size_t nb = std::max<u32>(body_capacity, 1), nm = std::max<u32>(manager_capacity, 1);
        for (auto* b : {&d->pv, &d->pw, &d->lin, &d->ang, &d->mlin, &d->mang, &d->rpos, &d->rq}) b->reserve(nb);
        d->sii.reserve(3 * nb);
        for (auto* b : {&d->has_row, &d->deg, &d->fill, &d->max_dyn, &d->body_mask}) b->reserve(nb);
        d->off.reserve(nb + 1);
        for (auto* b : {&d->inc, &d->pred}) b->reserve(2 * nm);
        d->entry.reserve(2 * nm);
        d->sinfo.reserve(nm);
        d->srank.reserve(nm);
        d->erank.reserve(2 * nm);
        d->vst.reserve(8ull * nb);
        d->pos_of.reserve(nm);
        for (auto* b : {&d->part, &d->pend, &d->sorted, &d->order}) b->reserve(nm);
        d->dd.reserve(nm);
        d->crow.reserve(4 * nm);
        d->frow.reserve(4 * nm);
        d->part_count.reserve(dy::kMaxParts + 1);
        d->part_off.reserve(dy::kMaxParts + 1);
        d->part_fill.reserve(dy::kMaxParts);
        d->ctr.reserve(dy::kCtrWords);
        d->flags.reserve(1);
        d->bar.reserve(dy::kBarWords);
        FYZZ_HIP_CHECK(hipMemset(d->bar.get(), 0, dy::kBarWords * sizeof(u32)));
        // 128 colouring workgroups measured best on the pile (64: +7 us, 256: +2 us)
        d->coop_blocks_partition = coop_blocks(reinterpret_cast<const void*>(&dy::k_part_color), ctx->device, 128);
        d->coop_blocks_pgs = coop_blocks(reinterpret_cast<const void*>(&dy::k_pgs), ctx->device, dy::kPgsBlocks);
        // the reverse sweep keeps grid barriers between partitions
        d->coop_blocks_adjoint = coop_blocks(reinterpret_cast<const void*>(&dy::k_pgs_adjoint), ctx->device);
        // room for the largest grid the autotuner may pick, and for the
        // tile sums of the degree scan
        d->blocksums.reserve(std::max<std::size_t>(2 * 1024 + 2, sel::tiles_for(static_cast<u32>(nb)) + 1));
        FYZZ_HIP_CHECK(hipMemset(d->flags.get(), 0, sizeof(u32)));
        FYZZ_HIP_CHECK(hipMemset(d->ctr.get(), 0, dy::kCtrWords * sizeof(u32)));
        dy::Ws& w = d->ws;
        w.pv = d->pv.get();
        w.pw = d->pw.get();
        w.lin = d->lin.get();
        w.ang = d->ang.get();
        w.mlin = d->mlin.get();
        w.mang = d->mang.get();
        w.rpos = d->rpos.get();
        w.rq = d->rq.get();
        w.sii = d->sii.get();
        w.has_row = d->has_row.get();
        w.deg = d->deg.get();
        w.off = d->off.get();
        w.fill = d->fill.get();
        w.max_dyn = d->max_dyn.get();
        w.body_mask = d->body_mask.get();
        w.inc = d->inc.get();
        w.pred = d->pred.get();
        w.part = d->part.get();
        w.entry = d->entry.get();
        w.sinfo = d->sinfo.get();
        w.srank = d->srank.get();
        w.erank = d->erank.get();
        w.vst = d->vst.get();
        w.pos_of = d->pos_of.get();
        w.pend = d->pend.get();
        w.sorted = d->sorted.get();
        w.order = d->order.get();
        w.dd = d->dd.get();
        w.crow = d->crow.get();
        w.frow = d->frow.get();
        w.part_count = d->part_count.get();
        w.part_off = d->part_off.get();
        w.part_fill = d->part_fill.get();
        w.blocksums = d->blocksums.get();
        w.ctr = d->ctr.get();
        w.flags = d->flags.get();
        w.bar = d->bar.get();
        // every workspace array the step kernels touch must exist
        const void* need[] = {w.pv, w.pw, w.lin, w.ang, w.mlin, w.mang, w.rpos, w.rq, w.sii, w.has_row,
                              w.deg, w.off, w.fill, w.max_dyn, w.body_mask, w.inc, w.pred, w.part, w.pend,
                              w.sorted, w.order, w.entry, w.sinfo, w.srank, w.erank, w.vst, w.pos_of, w.crow,
                              w.frow,
                              w.part_count, w.part_off, w.part_fill, w.blocksums, w.ctr, w.flags, w.bar};
        for (const void* q : need) FYZZ_REQUIRE(q != nullptr, kDeviceError, "dy_create: workspace array missing");
        *out = d.release();
    });
}

int fyzz_hip_dy_destroy(fyzz_hip_dy* d) {
    if (!d) return kInvalidArgument;
    (void)hipSetDevice(d->ctx->device);
    (void)hipStreamSynchronize(d->ctx->stream);
    delete d;
    return kOk;
}

int fyzz_hip_dy_partition(fyzz_hip_dy* d, uint32_t n_bodies, fyzz_dy_body* bodies, uint32_t n_managers,
                          const fyzz_dy_manager* managers, uint32_t* partitions) {
    if (!d) return kInvalidArgument;
    return guarded(d->ctx, [&] {
        FYZZ_REQUIRE(n_managers == 0 || partitions, kInvalidArgument, "dy_partition: null output");
        dy_prepare(d, 1.0f, n_bodies, bodies, 0, nullptr, n_managers, managers, nullptr, 0, nullptr, true);
        d->ws.parts_out = partitions;
        d->ws.impulses = nullptr;
        if (n_managers)
            hipLaunchKernelGGL(dy::k_outputs, dim3(grid_for(n_managers, dy::kBlock)), dim3(dy::kBlock), 0,
                               d->ctx->stream, d->ws);
        FYZZ_HIP_CHECK(hipGetLastError());
    });
}

int fyzz_hip_dy_prepare(fyzz_hip_dy* d, float dt, uint32_t n_bodies, fyzz_dy_body* bodies, uint32_t n_statics,
                        const fyzz_dy_static* statics, uint32_t n_managers, const fyzz_dy_manager* managers,
                        const fyzz_dy_patch* history, uint32_t n_history, fyzz_dy_patch* patches) {
    if (!d) return kInvalidArgument;
    return guarded(d->ctx, [&] {
        dy_prepare(d, dt, n_bodies, bodies, n_statics, statics, n_managers, managers, history, n_history, patches,
                   false);
    });
}

int fyzz_hip_dy_solve(fyzz_hip_dy* d) {
    if (!d) return kInvalidArgument;
    return guarded(d->ctx, [&] { dy_solve(d); });
}

int fyzz_hip_dy_integrate(fyzz_hip_dy* d, float* impulses, uint32_t* partitions) {
    if (!d) return kInvalidArgument;
    return guarded(d->ctx, [&] { dy_integrate(d, impulses, partitions); });
}

int fyzz_hip_dy_step(fyzz_hip_dy* d, float dt, uint32_t n_bodies, fyzz_dy_body* bodies, uint32_t n_statics,
                     const fyzz_dy_static* statics, uint32_t n_managers, const fyzz_dy_manager* managers,
                     const fyzz_dy_patch* history, uint32_t n_history, fyzz_dy_patch* patches, float* impulses,
                     uint32_t* partitions) {
    int rc = fyzz_hip_dy_prepare(d, dt, n_bodies, bodies, n_statics, statics, n_managers, managers, history,
                                 n_history, patches);
    if (rc == kOk) rc = fyzz_hip_dy_solve(d);
    if (rc == kOk) rc = fyzz_hip_dy_integrate(d, impulses, partitions);
    return rc;
}

int fyzz_hip_dy_set_tape(fyzz_hip_dy* d, int enabled) {
    if (!d) return kInvalidArgument;
    return guarded(d->ctx, [&] { dy_set_tape(d, enabled != 0); });
}

int fyzz_hip_dy_adjoint(fyzz_hip_dy* d, const float* adj_v_out, const float* adj_w_out, const float* adj_p_out,
                        float* adj_v_in, float* adj_w_in, float* adj_friction) {
    if (!d) return kInvalidArgument;
    return guarded(d->ctx, [&] { dy_adjoint(d, adj_v_out, adj_w_out, adj_p_out, adj_v_in, adj_w_in, adj_friction); });
}

int fyzz_hip_dy_friction_anchors(fyzz_hip_dy* d, fyzz_friction_anchors* out) {
    if (!d) return kInvalidArgument;
    return guarded(d->ctx, [&] {
        FYZZ_REQUIRE(out || d->ws.nm == 0, kInvalidArgument, "dy: null anchor output");
        if (!d->ws.nm) return;
        hipLaunchKernelGGL(dy::k_anchors, dim3(grid_for(d->ws.nm, dy::kBlock)), dim3(dy::kBlock), 0, d->ctx->stream,
                           d->ws, out);
        FYZZ_HIP_CHECK(hipGetLastError());
    });
}

int fyzz_hip_dy_status(fyzz_hip_dy* d, uint32_t* flags, uint32_t* n_partitions, uint32_t* rounds) {
    if (!d) return kInvalidArgument;
    return guarded(d->ctx, [&] {
        u32 f = 0, c[dy::kCtrWords];
        FYZZ_HIP_CHECK(hipMemcpyAsync(&f, d->flags.get(), sizeof f, hipMemcpyDeviceToHost, d->ctx->stream));
        FYZZ_HIP_CHECK(hipMemcpyAsync(c, d->ctr.get(), sizeof c, hipMemcpyDeviceToHost, d->ctx->stream));
        FYZZ_HIP_CHECK(hipStreamSynchronize(d->ctx->stream));
        if (flags) *flags = f;
        if (n_partitions) *n_partitions = c[dy::kNParts];
        if (rounds) *rounds = c[dy::kRounds];
        FYZZ_REQUIRE(!(f & FYZZ_DY_BAD_INPUT), kInvalidArgument, "dy: malformed manager record");
        FYZZ_REQUIRE(!(f & FYZZ_DY_OVERFLOW), kOverflow, "dy: partition or overflow-wave capacity exceeded");
        FYZZ_REQUIRE(!(f & FYZZ_DY_NONFINITE), kOverflow, "dy: non-finite solver value");
        FYZZ_REQUIRE(!(f & FYZZ_DY_LOGIC), kLogicError, "dy: manager outside the admitted class");
    });
}

}  // extern "C"
