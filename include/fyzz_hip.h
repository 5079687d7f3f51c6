/*
 * fyzz_hip.h - C-ABI of the MI355X (gfx950) FyzzXYZZY step pipeline.
 *
 * Plain pointers and sizes only. Every function returns 0 on success or a
 * negative status (FYZZ_HIP_E_*); the message of the last failure on a
 * context is available from fyzz_hip_last_error(). No C++ exception
 * crosses this boundary. Pointers documented as "device" must be device
 * allocations visible to the context's device; all work is enqueued on
 * the context's stream (fyzz_hip_set_stream) and is asynchronous unless
 * the function says it synchronizes.
 *
 * Reference interfaces replaced (paths relative to the reference root):
 *   GpuBroadphaseProducer plug-in           src/native/scene.h:383-505
 *   GpuBroadphaseReports {found,lost,...}   src/native/scene.h:78-91
 *   Scene::update_bounds / update_pairs     src/native/scene.cpp:1336-2004, 2564-2704
 *   Scene::generate_pcm_box_box / plane_box src/native/scene.cpp:3466-3654, 3774-3955
 *   Dy prepare / partition / solve / integrate  src/native/scene.cpp:6940-8511, 8853-9148, 10383-10707
 * Each declaration below names the reference function it stands in for.
 */
#ifndef FYZZ_HIP_H
#define FYZZ_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FYZZ_HIP_OK 0
#define FYZZ_HIP_E_INVALID_ARGUMENT (-1) /* std::invalid_argument */
#define FYZZ_HIP_E_DEVICE (-2)           /* std::runtime_error (device) */
#define FYZZ_HIP_E_OVERFLOW (-3)         /* std::overflow_error */
#define FYZZ_HIP_E_LOGIC (-4)            /* std::logic_error */
#define FYZZ_HIP_E_OUT_OF_MEMORY (-5)

typedef struct fyzz_hip_context fyzz_hip_context;
typedef struct fyzz_hip_sap fyzz_hip_sap;
typedef struct fyzz_hip_dy fyzz_hip_dy;
typedef struct fyzz_hip_world fyzz_hip_world;

/* ---- Dy records (plain data, shared by the device entries and hosts) --
 * fyzz_dy_body: one dynamic rigid body (Scene::Actor's Dy fields,
 * scene.h:1109-1165, with its single shape's material). flags: bit0
 * sleeping, bit1 ready for sleep, bit2 gravity enabled. iterations:
 * position | velocity << 8 (RigidBodyDynamics, world.h:115-140). */
typedef struct fyzz_dy_body {
    float position[3];
    float orientation[4];
    float linear_velocity[3];
    float angular_velocity[3];
    float inverse_mass;
    float inverse_inertia[3];
    float wake_counter;
    float sleep_linear_acc[3];
    float sleep_angular_acc[3];
    uint32_t counted_interactions;
    uint32_t flags;
    float linear_damping, angular_damping, max_linear_velocity, max_angular_velocity;
    float static_friction, dynamic_friction, restitution;
    uint32_t iterations;
    float pad[2];
} fyzz_dy_body; /* 144 bytes */

/* A static body (the floor): pose and material. */
typedef struct fyzz_dy_static {
    float position[3];
    float orientation[4];
    float static_friction, dynamic_friction, restitution;
} fyzz_dy_static; /* 40 bytes */

/* One published Gu contact (ContactPoint, scene.h:510-514). */
typedef struct fyzz_dy_contact {
    float position[3];
    float normal[3];
    float separation;
} fyzz_dy_contact; /* 28 bytes */

/* One touching contact manager in solver (descriptor) order: Gu-ordered
 * bodies (body0 first; a negative value -1-k names static body k), its
 * 1..4 published contacts, and the index of its previous-frame friction
 * patch (0xffffffff: none). */
typedef struct fyzz_dy_manager {
    int32_t body0, body1;
    uint32_t count;
    uint32_t history;
    fyzz_dy_contact contacts[4];
} fyzz_dy_manager; /* 128 bytes */

/* A persistent correlated friction patch (Scene::FrictionPatch,
 * scene.h:1258-1280): body-local normals and anchors. flags: bit0 broken,
 * bit1 disable strong friction. */
typedef struct fyzz_dy_patch {
    float body0_normal[3], body1_normal[3];
    float body0_anchors[2][3], body1_anchors[2][3];
    float static_friction, dynamic_friction, restitution, damping;
    uint32_t anchor_count, flags;
} fyzz_dy_patch; /* 96 bytes */

/* fyzz::ContactEvent (include/fyzzxyzzy/world.h:178-190), same layout:
 * body0 < body1, kind (0 found, 1 persists, 2 lost), contact count, offset
 * of the event's window in the point array, summed normal impulse, shapes
 * in body order. */
typedef struct fyzz_contact_event {
    uint32_t body0, body1;
    uint8_t kind, contact_count;
    uint8_t pad[6];
    uint64_t point_offset;
    double normal_impulse;
    uint32_t shape0, shape1;
} fyzz_contact_event; /* 40 bytes */

/* fyzz::ContactEventPoint (world.h:194-199), same layout: position,
 * outward normal, separation, per-point normal impulse. */
typedef struct fyzz_contact_event_point {
    float position[3];
    float normal[3];
    float separation;
    float pad;
    double normal_impulse;
} fyzz_contact_event_point; /* 40 bytes */

/* Scene::FrictionPatchWriteback (scene.h:522-528), per touching manager:
 * its patch's anchor count and broken flag, each anchor's world position
 * on body0 (the manager's first body) and world friction impulse (the two
 * tangent rows' tangent x applied force, summed in double), as
 * Scene::write_back_contacts publishes them (scene.cpp:9339-9388). */
typedef struct fyzz_friction_anchors {
    uint32_t anchor_count;
    uint32_t broken;
    float position[2][3];
    double impulse[2][3];
} fyzz_friction_anchors; /* 80 bytes */

/* ---- context -------------------------------------------------------- */
/* Stands in for the producer's owned compute::Context + stream
 * (source_sc_owned_gpu_broadphase_native.h:66-162). */
int fyzz_hip_context_create(int device, fyzz_hip_context** out);
int fyzz_hip_context_destroy(fyzz_hip_context* ctx);
const char* fyzz_hip_last_error(const fyzz_hip_context* ctx);
/* hip_stream: a hipStream_t of the context's device, or NULL for the
 * default stream. */
int fyzz_hip_set_stream(fyzz_hip_context* ctx, void* hip_stream);
int fyzz_hip_synchronize(fyzz_hip_context* ctx);
/* Launch mode of the Dy grid-barrier kernels (no counterpart in the
 * reference): cooperative != 0 (default) uses hipLaunchCooperativeKernel,
 * which checks that every workgroup is resident; 0 launches the same
 * resident-sized grid as a plain kernel, as inside captured graphs (for
 * tools that do not intercept cooperative launches, such as counter
 * profilers). Results are identical; a non-resident grid sets the
 * FYZZ_DY_LOGIC status bit instead of waiting forever. */
int fyzz_hip_set_cooperative(fyzz_hip_context* ctx, int cooperative);
/* Stage timing: when enabled, named launches are bracketed by HIP events
 * on the context's stream (stands in for Context::profile_batch_gpu_nanoseconds,
 * include/fyzzxyzzy/compute.h:299). profile_read synchronizes, returns the
 * summed milliseconds and span count of one name, and resets it. Names:
 * "sap.bounds", "sap.sort", "sap.sort.scatter", "sap.pairs", "sap.lists",
 * "pcm.box_box", "pcm.plane_box", "dy.bodies", "dy.partition", "dy.prep",
 * "dy.pgs", "dy.integrate", "dy.adjoint", "world.islands", "world.managers",
 * "world.events", "world.scene_pairs". */
int fyzz_hip_profile_enable(fyzz_hip_context* ctx, int enable);
int fyzz_hip_profile_read(fyzz_hip_context* ctx, const char* name, double* total_ms,
                          uint64_t* count);
/* Library build identification (static string). */
const char* fyzz_hip_version(void);

/* ---- SAP stage entries (device pointers) ----------------------------- */
/* Scene::update_shape_bounds + encode_min/encode_max
 * (scene.cpp:26-35, 1233-1316) for n boxes: pos[3n], quat[4n] (x,y,z,w),
 * half[3n], contact_offset[n] -> bounds[6][n] (min x,y,z keys then max
 * x,y,z keys). *error_flag |= 1 when a bound is not finite (the
 * reference's std::overflow_error, scene.cpp:1298-1300). */
int fyzz_hip_sap_box_bounds(fyzz_hip_context* ctx, uint32_t n, const float* pos,
                            const float* quat, const float* half,
                            const float* contact_offset, uint32_t* bounds,
                            uint32_t* error_flag);

/* The sorted endpoint arrays of Scene::update_bounds (scene.cpp:1336-1400,
 * Endpoint scene.h:1295-1304) for all three axes: values[3][2n] and
 * payloads[3][2n], payload = slot << 1 | is_max, ordered by (value,
 * min-before-max, slot). scratch: fyzz_hip_sap_sort_scratch_words(n) u32. */
int fyzz_hip_sap_sort_endpoints(fyzz_hip_context* ctx, uint32_t n, const uint32_t* bounds,
                                uint32_t* values, uint32_t* payloads, uint32_t* scratch,
                                uint64_t scratch_words);
uint64_t fyzz_hip_sap_sort_scratch_words(uint32_t n);

/* No counterpart in the reference: for callers that cannot allocate
 * device memory (code under stream capture, tight serving loops,
 * harnesses that forbid hipMalloc). Same inputs, outputs and order as
 * fyzz_hip_sap_sort_endpoints, with the workspace held by the library:
 * one fixed array per device (6.5 MB, kept until the process ends),
 * looked up when a context is created on that device. Allocates nothing.
 * n above fyzz_hip_sap_sort_fixed_max_boxes() (131,072 boxes) returns
 * FYZZ_HIP_E_INVALID_ARGUMENT. Every call on one device shares that
 * workspace: calls that overlap on one device (other streams, other
 * threads) race. */
int fyzz_hip_sap_sort_endpoints_fixed(fyzz_hip_context* ctx, uint32_t n, const uint32_t* bounds,
                                      uint32_t* values, uint32_t* payloads);
uint32_t fyzz_hip_sap_sort_fixed_max_boxes(void);

/* ---- incremental broadphase (GpuBroadphaseProducer::advance) -------- */
/* n boxes in slot order (increasing shape id); body_ids[n] and dynamic[n]
 * are HOST arrays (Bp filter: pairs need different bodies and at least one
 * dynamic shape, scene.cpp:2606-2608). pair_capacity bounds the live pair
 * list, delta_capacity the per-frame candidate changes. */
int fyzz_hip_sap_create(fyzz_hip_context* ctx, uint32_t n, const uint32_t* body_ids,
                        const uint8_t* dynamic, uint32_t pair_capacity,
                        uint32_t delta_capacity, fyzz_hip_sap** out);
/* Destroy every broadphase object before its context. */
int fyzz_hip_sap_destroy(fyzz_hip_sap* sap);
/* GpuBroadphaseProducer::advance (scene.h:394-396). Device pose arrays as
 * in fyzz_hip_sap_box_bounds. dirty_slots (device, n_dirty entries) lists
 * the boxes whose pose may have changed (NULL: all). The first call, and
 * any call with force_full != 0, rebuilds from scratch
 * (GpuBroadphaseReports::full_rebuild). Asynchronous. */
int fyzz_hip_sap_advance(fyzz_hip_sap* sap, const float* pos, const float* quat,
                         const float* half, const float* contact_offset, uint32_t n_dirty,
                         const uint32_t* dirty_slots, int force_full);
/* Synchronizes. status_flags: bit0 non-finite bounds, bit1 capacity
 * overflow (GpuBroadphaseReports::overflow). */
int fyzz_hip_sap_counts(fyzz_hip_sap* sap, uint64_t* n_pairs, uint64_t* n_found,
                        uint64_t* n_lost, uint32_t* status_flags);
/* Synchronizes. which: 0 live pairs (Scene::pairs), 1 found, 2 lost
 * (GpuBroadphaseReports::found/lost), as sorted (lo, hi) SLOT pairs into a
 * HOST array of 2*capacity u32; *written = total count. Overflow or
 * non-finite bounds return FYZZ_HIP_E_OVERFLOW. */
int fyzz_hip_sap_fetch(fyzz_hip_sap* sap, int which, uint32_t* out_pairs, uint64_t capacity,
                       uint64_t* written);
/* Copies the current sorted endpoints (as fyzz_hip_sap_sort_endpoints
 * lays them out) into DEVICE arrays values[3][2n], payloads[3][2n]. */
int fyzz_hip_sap_endpoints(fyzz_hip_sap* sap, uint32_t* values, uint32_t* payloads);

/* ---- PCM narrowphase (persistent manifolds) ------------------------- */
/* Bytes of one persistent manifold cache record (Scene::PcmManifoldCache,
 * scene.h:1448-1459, without the write-only GJK warm start): 224. A
 * zero-filled record is an empty, uninitialised cache. */
uint32_t fyzz_hip_pcm_cache_bytes(void);

/* Scene::generate_pcm_box_box (scene.cpp:3466-3654) for n_pairs pairs in
 * Gu order (first, second are box slots; contact normals are expressed in
 * second's frame). Poses: pos[3n], quat[4n], half[3n] per slot (device).
 * cache_index (device, may be NULL = pair index) selects each pair's cache
 * record in `cache`. Outputs (device): contacts[n_pairs][4][7] =
 * (position, normal, separation) per published contact, counts[n_pairs],
 * flags[n_pairs] (bit0 regenerated, bit1 GJK/EPA continuation, bit2 the
 * reference would raise overflow_error). scratch: n_pairs + 1 u32. */
int fyzz_hip_pcm_box_box(fyzz_hip_context* ctx, uint32_t n_pairs, const uint32_t* first,
                         const uint32_t* second, const uint32_t* cache_index, const float* pos,
                         const float* quat, const float* half, float contact_distance, void* cache,
                         float* contacts, uint32_t* counts, uint32_t* flags, uint32_t* scratch);

/* Scene::generate_pcm_plane_box (scene.cpp:3774-3955) for n dynamic boxes
 * against the +y plane at HOST plane_position[3], with the scene's
 * published (+y) normal (scene.cpp:4652-4657). Arguments as above. */
int fyzz_hip_pcm_plane_box(fyzz_hip_context* ctx, uint32_t n, const uint32_t* box_slot,
                           const uint32_t* cache_index, const float* pos, const float* quat,
                           const float* half, const float* plane_position, float contact_distance,
                           void* cache, float* contacts, uint32_t* counts, uint32_t* flags);

/* Stateless parity entries over n pose pairs a[n][10], b[n][10] (half,
 * position, quat): source_pcm_box_box_candidates + batch selection
 * (GuPCMContactBoxBox.cpp:185-436) -> out[n][16][10] (local_a, local_b,
 * normal_b, separation), counts[n], overlapping[n], selection[n][4] (255
 * when <= 4 candidates); and source_pcm_box_box_gjk_epa
 * (GuPCMGjkEpa.cpp:527-566) -> out[n][10], ok[n]. */
int fyzz_hip_pcm_candidates(fyzz_hip_context* ctx, uint32_t n, const float* a, const float* b,
                            float contact_distance, float* out, uint32_t* counts, uint32_t* overlapping,
                            uint32_t* selection);
int fyzz_hip_pcm_gjk_epa(fyzz_hip_context* ctx, uint32_t n, const float* a, const float* b,
                         float contact_distance, float* out, uint32_t* ok);

/* ---- Dy stage (device records, one solver task) ---------------------- */
/* Status bits raised by a Dy step on the device (fyzz_hip_dy_status). */
#define FYZZ_DY_NONFINITE 1u /* overflow_error: a non-finite solver value */
#define FYZZ_DY_LOGIC 2u     /* logic_error: manager outside the admitted class */
#define FYZZ_DY_OVERFLOW 4u  /* overflow_error: partition capacity exceeded */
#define FYZZ_DY_BAD_INPUT 8u /* invalid_argument: malformed manager record */

/* Workspace for steps of up to body_capacity bodies and manager_capacity
 * managers (the producer's Dy buffers, GpuBroadphaseProducer::
 * solve_gpu_dy_distance's spans, scene.h:424-448). */
int fyzz_hip_dy_create(fyzz_hip_context* ctx, uint32_t body_capacity, uint32_t manager_capacity,
                       fyzz_hip_dy** out);
int fyzz_hip_dy_destroy(fyzz_hip_dy* dy);

/* Scene::partition_constraints (scene.cpp:7113-7489) for one task spanning
 * the managers in the given (descriptor) order: first-fit over 32-bit
 * masks for dynamic/dynamic managers, overflow waves of 32, and
 * dynamic/static managers after their body's highest dynamic partition.
 * bodies, managers (device) as for fyzz_hip_dy_step; partitions[n] (device). */
int fyzz_hip_dy_partition(fyzz_hip_dy* dy, uint32_t n_bodies, fyzz_dy_body* bodies, uint32_t n_managers,
                          const fyzz_dy_manager* managers, uint32_t* partitions);

/* One Dy step over device records, in three stages that fyzz_hip_dy_step
 * runs back to back:
 *   prepare   prepare_unconstrained_velocities (scene.cpp:10383-10507),
 *             prepare_solver_bodies (6940-7112), partition_constraints
 *             (7113-7489) and the contact part of prepare_constraints
 *             (7490-7515, 7763-8511): material combine, friction-patch
 *             correlation with history[n_history] (a manager's `history`
 *             indexes it), contact and friction rows; writes patches[nm],
 *             the patch each manager carries into the next step;
 *   solve     solve_constraint_task (8853-9148): position then velocity
 *             iterations over the partitions (counts: the maximum over the
 *             constrained bodies' `iterations`);
 *   integrate publish_solver_bodies (9509-9549) and integrate_solver_task
 *             (10508-10707): bodies[] are updated in place (pose,
 *             velocities, wake counter, sleep accumulators, ready flag);
 *             impulses[nm][4] receive each manifold contact's normal
 *             impulse and partitions[nm] each manager's partition (either
 *             may be NULL).
 * All pointers are device pointers; statics[k] is the static body that a
 * manager names as -1-k. Records stay borrowed until integrate returns. */
int fyzz_hip_dy_prepare(fyzz_hip_dy* dy, float dt, uint32_t n_bodies, fyzz_dy_body* bodies,
                        uint32_t n_statics, const fyzz_dy_static* statics, uint32_t n_managers,
                        const fyzz_dy_manager* managers, const fyzz_dy_patch* history,
                        uint32_t n_history, fyzz_dy_patch* patches);
int fyzz_hip_dy_solve(fyzz_hip_dy* dy);
int fyzz_hip_dy_integrate(fyzz_hip_dy* dy, float* impulses, uint32_t* partitions);
int fyzz_hip_dy_step(fyzz_hip_dy* dy, float dt, uint32_t n_bodies, fyzz_dy_body* bodies,
                     uint32_t n_statics, const fyzz_dy_static* statics, uint32_t n_managers,
                     const fyzz_dy_manager* managers, const fyzz_dy_patch* history, uint32_t n_history,
                     fyzz_dy_patch* patches, float* impulses, uint32_t* partitions);

/* Adjoint of a Dy step (no counterpart in the reference: the
 * differentiable form). With the tape on, a step records each row's
 * branch per iteration and the rows' unscaled lever arms.
 * fyzz_hip_dy_adjoint then runs reverse mode through that step's PGS and
 * integration with its contact rows frozen (normals, lever arms, effective
 * masses, biases and friction frames constant): inputs are the adjoints of
 * the step's output linear velocity, angular velocity and position
 * (device [nb][3] each); outputs the adjoints of its input linear and
 * angular velocity ([nb][3]) and of each body's dynamic friction ([nb]).
 * The position adjoint passes through unchanged. Orientation is not
 * differentiated. Call it after the step's integrate. */
int fyzz_hip_dy_set_tape(fyzz_hip_dy* dy, int enabled);
int fyzz_hip_dy_adjoint(fyzz_hip_dy* dy, const float* adj_v_out, const float* adj_w_out, const float* adj_p_out,
                        float* adj_v_in, float* adj_w_in, float* adj_friction);

/* Synchronizes the context's stream and reports the last Dy step:
 * flags (FYZZ_DY_* bits), the partition count and the dependency depth
 * of the partitioner (the longest chain of managers where each shares a
 * dynamic body with the previous, earlier in list order; a sequential
 * round-by-round colouring needs this many rounds). Returns the status matching the
 * first raised flag (the exception the reference would throw). */
int fyzz_hip_dy_status(fyzz_hip_dy* dy, uint32_t* flags, uint32_t* n_partitions, uint32_t* rounds);
/* The friction write-back of the last step's managers (Scene::
 * write_back_contacts, scene.cpp:9339-9388) into DEVICE out[n_managers],
 * computed on demand from the step's solved rows and the patches array
 * the step wrote (which must still hold them). Asynchronous. */
int fyzz_hip_dy_friction_anchors(fyzz_hip_dy* dy, fyzz_friction_anchors* out);

/* ---- Step pipeline (fyzz::World over the native GPU backend) --------- */
/* A world of one static floor plane and dynamic boxes whose whole step
 * (pose gather, broadphase advance, plane x box pairs, box/box and
 * plane/box PCM with persistent manifolds, touching-manager list with
 * friction history, Dy) is enqueued on the context's stream with no host
 * readback between stages. Stands in for World + the native GPU backend
 * (include/fyzzxyzzy/world.h:400-470; create_native_gpu_world,
 * src/native/native_gpu_world.h:84-138) for this scene class. Body ids
 * follow creation order as in the reference (the floor takes an id).
 * Authoring calls precede the first step. Manager (solver) order: touching
 * box/box managers in broadphase pair order, then touching plane managers
 * in creation order (the reference's island order is not rebuilt). */
int fyzz_hip_world_create(fyzz_hip_context* ctx, uint32_t max_bodies, fyzz_hip_world** out);
int fyzz_hip_world_destroy(fyzz_hip_world* world);
/* World::add_static_floor (world.h:424) */
int fyzz_hip_world_add_static_floor(fyzz_hip_world* world, float height, uint32_t* body_id);
/* World::add_static_box / add_static_sphere (world.h:269, 272;
 * Scene::add_static_box / add_static_sphere, scene.cpp:441-488): a static
 * body at HOST position[3] (identity orientation; fyzz_hip_world_set_body_pose
 * may turn it before the first step), the default material
 * (fyzz_hip_world_set_body_material changes it at any time). Statics pair
 * with dynamic bodies only; the Gu order puts them second, and the solver
 * treats them as it treats the floor. Single-scene worlds only. */
int fyzz_hip_world_add_static_box(fyzz_hip_world* world, const float* half_extents, const float* position,
                                  uint32_t* body_id);
int fyzz_hip_world_add_static_sphere(fyzz_hip_world* world, float radius, const float* position, uint32_t* body_id);
/* World::add_dynamic_sphere (world.h:427; Scene::add_dynamic_sphere,
 * scene.cpp:491-525): radius, density. Sphere pairs take the stateless
 * sphere contacts of GuPCMContactSphere.cpp (sphere/sphere, sphere/box,
 * sphere/plane; no persistent manifold), in single worlds and in
 * batched scenes alike. */
int fyzz_hip_world_add_dynamic_sphere(fyzz_hip_world* world, float radius, float density, uint32_t* body_id);
/* World::add_dynamic_box (world.h:425): HOST half_extents[3], density */
int fyzz_hip_world_add_dynamic_box(fyzz_hip_world* world, const float* half_extents, float density,
                                   uint32_t* body_id);
/* n add_dynamic_box calls followed by set_body_pose, set_body_velocity,
 * set_body_angular_velocity and set_body_material on each (HOST arrays
 * half_extents[3n], positions[3n], orientations[4n], linear[3n] and
 * angular[3n], the last two may be NULL); ids are consecutive from
 * *first_id. */
int fyzz_hip_world_add_dynamic_boxes(fyzz_hip_world* world, uint32_t n, const float* half_extents, float density,
                                     const float* positions, const float* orientations, const float* linear,
                                     const float* angular, float static_friction, float dynamic_friction,
                                     float restitution, uint32_t* first_id);
/* World::set_body_pose / set_body_velocity + set_body_angular_velocity
 * (world.h:475-478; Scene::set_pose, set_velocity, set_angular_velocity,
 * scene.cpp:589-781); HOST arrays, NULL leaves a part. Allowed before and
 * after the first step. After it, a pose edit wakes the body, resets the
 * PCM cache of every current broadphase pair holding it and, when the
 * pose changed, erases those pairs' friction history and wakes touching
 * peers that sleep or are ready to sleep; a velocity edit wakes the body.
 * Poses must be finite with |q| within 1e-2 of 1 (invalid_argument);
 * the floor cannot move (logic_error). Synchronizes. */
int fyzz_hip_world_set_body_pose(fyzz_hip_world* world, uint32_t body_id, const float* position,
                                 const float* orientation);
int fyzz_hip_world_set_body_velocity(fyzz_hip_world* world, uint32_t body_id, const float* linear,
                                     const float* angular);
/* World::set_body_material (world.h:491; scene.cpp:1022, 1099-1118):
 * frictions finite and >= 0, restitution in [0, 1]; any time. */
int fyzz_hip_world_set_body_material(fyzz_hip_world* world, uint32_t body_id, float static_friction,
                                     float dynamic_friction, float restitution);
/* fyzz::RigidBodyDynamics (world.h:115-140), same fields, order and
 * layout (40 bytes). */
typedef struct fyzz_body_dynamics {
    float linear_damping, angular_damping, max_linear_velocity, max_angular_velocity;
    float max_depenetration_velocity, max_contact_impulse, contact_slop_coefficient;
    uint8_t gravity_enabled;
    uint8_t pad0[3];
    uint32_t lock_flags;
    uint8_t gyroscopic_forces_enabled, retain_accelerations, position_iterations, velocity_iterations;
} fyzz_body_dynamics;
/* World::set_body_dynamics (world.h:506; Scene::set_dynamics,
 * scene.cpp:1119-1147): the reference's validation (invalid_argument),
 * then damping, speed limits, gravity and iteration counts, any time.
 * Non-default depenetration or impulse caps, slop, locks, gyroscopic
 * forces or retained accelerations are outside the admitted class
 * (logic_error). */
int fyzz_hip_world_set_body_dynamics(fyzz_hip_world* world, uint32_t body_id, const fyzz_body_dynamics* dynamics);
/* Independent scenes in one world (the batched form): scene k owns the
 * next box_counts[k] dynamic boxes in creation order (at most 1024 each);
 * boxes of different scenes never pair. The scenes share the floor plane
 * and its material. The broadphase then runs per scene (one workgroup per
 * scene, every pair of the scene tested with the broadphase's filter, so
 * each scene's pair set equals the SAP's), and every later stage runs once
 * over all scenes. Each scene steps exactly as a world holding only that
 * scene would. Before the first step; no contact notifications. */
int fyzz_hip_world_set_scenes(fyzz_hip_world* world, uint32_t n_scenes, const uint32_t* box_counts);
/* The solver order of the world's steps (before the first step):
 * 0 (default) the device order, touching box/box managers in broadphase
 *   pair order then plane managers, built on the device with no host
 *   round trip (graph capture, batched scenes);
 * 1 the native scene's order (World::step, world.h:325, of the reference's
 *   NativeCpu backend; descriptor order scene.cpp:7174-7260): each
 *   step hands its broadphase and touch changes to the host island
 *   authority (fyzz_hip_islands_*), whose order the solver list takes, and
 *   checks that the device island pass sleeps the same bodies. The step
 *   then synchronizes twice, so fyzz_hip_world_step_graph and
 *   fyzz_hip_world_autotune refuse it (logic_error), as do batched scenes
 *   and body edits after the first step. */
int fyzz_hip_world_set_solver_order(fyzz_hip_world* world, int native);
/* Per-body materials for n consecutive body ids from first_id (HOST
 * arrays): set_body_material on each. */
int fyzz_hip_world_set_materials(fyzz_hip_world* world, uint32_t first_id, uint32_t n, const float* static_friction,
                                 const float* dynamic_friction, const float* restitution);
/* World::step (world.h:459); asynchronous. */
int fyzz_hip_world_step(fyzz_hip_world* world, float dt);
/* One World::step through a captured HIP graph: the first two calls
 * capture one graph each (the step's double buffers alternate), later
 * calls replay them; a new dt recaptures. A world without scenes runs its
 * first step (the broadphase's full build) directly. Results equal
 * fyzz_hip_world_step's. Asynchronous. */
int fyzz_hip_world_step_graph(fyzz_hip_world* world, float dt);
/* Autotuner of the launch shapes (no counterpart in the reference): times
 * trial_steps real steps per candidate workgroup count (16-256) of the
 * partition and PGS launches, each candidate twice (increasing, then
 * decreasing order, so a world that changes under the trials biases no
 * candidate), and keeps the fastest; the 20 trials of 2 + trial_steps steps
 * advance the world. Launch shapes do not change results. Writes the chosen
 * counts. */
int fyzz_hip_world_autotune(fyzz_hip_world* world, float dt, uint32_t trial_steps, int* partition_blocks,
                            int* pgs_blocks);
/* Launch shapes set by hand (no counterpart in the reference): workgroup
 * counts of the partition colouring and PGS launches, each clamped to what
 * stays co-resident on the device; 0 keeps a count. Drops captured graphs.
 * Launch shapes do not change results. Writes the counts in effect. */
int fyzz_hip_world_set_workgroups(fyzz_hip_world* world, int partition_blocks, int pgs_blocks,
                                  int* partition_in_effect, int* pgs_in_effect);
/* Synchronizes; copies the dynamic bodies' records (creation order) to
 * HOST out[capacity]; *count = number of dynamic bodies. */
int fyzz_hip_world_body_records(fyzz_hip_world* world, fyzz_dy_body* out, uint32_t capacity, uint32_t* count);
/* Synchronizes; counts of the last step and its status bits (FYZZ_DY_*,
 * 0x100 non-finite bound, 0x200 pair capacity exceeded). */
int fyzz_hip_world_stats(fyzz_hip_world* world, uint32_t* n_pairs, uint32_t* n_managers,
                         uint32_t* n_partitions, uint32_t* status_flags);
/* Synchronizes; copies the last step's manager list (solver order), its
 * contact impulses [.][4] and partitions to HOST arrays of `capacity`
 * entries (any may be NULL); *count = number of managers. */
int fyzz_hip_world_fetch_managers(fyzz_hip_world* world, fyzz_dy_manager* managers, float* impulses,
                                  uint32_t* partitions, uint32_t capacity, uint32_t* count);
/* Scene::solved_friction_patches (scene.h:846-848): the last step's
 * friction write-back, one record per manager in the manager list's
 * order, into a HOST array of `capacity`; *count = number of managers.
 * Computed on demand (nothing is written while stepping); synchronizes. */
int fyzz_hip_world_friction_anchors(fyzz_hip_world* world, fyzz_friction_anchors* out, uint32_t capacity,
                                    uint32_t* count);
/* Synchronizes; running totals since creation or the last reset: [0]
 * steps, [1] broadphase pairs, [2] touching managers, [3] contact rows,
 * [4] friction rows, [5] partitions, [6] solver iterations, [7] partition
 * dependency depth, [8] broadphase found + lost pairs (out[9], HOST).
 * reset != 0 zeroes them afterwards. */
int fyzz_hip_world_totals(fyzz_hip_world* world, uint64_t* out, int reset);
/* Scene::set_contact_notifications (scene.h:654): when enabled, every
 * step also builds Scene::publish_contact_events' output on the device
 * (scene.cpp:9149-9259): touching managers in Gu pair order as found or
 * persists, merged with lost touches, each owning a window of points. */
int fyzz_hip_world_set_contact_notifications(fyzz_hip_world* world, int enabled);
/* Synchronizes; copies the last step's events and points to HOST arrays
 * (either may be NULL to query the counts). World::contact_events /
 * contact_event_points (world.h:467-468). */
int fyzz_hip_world_contact_events(fyzz_hip_world* world, fyzz_contact_event* events, uint32_t capacity,
                                  uint32_t* n_events, fyzz_contact_event_point* points,
                                  uint32_t point_capacity, uint32_t* n_points);
/* The last step's Dy inputs for the differentiable form (no counterpart
 * in the reference): its managers in solver order and the friction patch
 * each one read, with the manager's history index rewritten to its own
 * index (DEVICE arrays of `capacity`; NULL queries the count).
 * Synchronizes to read the count. */
int fyzz_hip_world_record_dy_inputs(fyzz_hip_world* world, fyzz_dy_manager* managers, fyzz_dy_patch* history,
                                    uint32_t capacity, uint32_t* n_managers);
/* The same records appended to DEVICE arenas of `capacity` records at
 * *cursor (DEVICE u64), with no host round trip (no counterpart in the
 * reference): *count (DEVICE u32) receives the step's manager count and
 * the cursor advances by it; when the records do not fit nothing is
 * copied and the cursor becomes all ones, so later appends copy nothing
 * either (a caller reads the cursor once, after its rollout). Each
 * record's history index is its index within the step's records. */
int fyzz_hip_world_append_dy_inputs(fyzz_hip_world* world, fyzz_dy_manager* managers, fyzz_dy_patch* history,
                                    uint64_t capacity, uint64_t* cursor, uint32_t* count);
/* Synchronizes; the last step's broadphase reports as sorted (lo, hi) box
 * SLOT pairs into a HOST array of 2*capacity u32 (*written = the total
 * count): which 0 the live pairs (Scene::pairs), 1 found, 2 lost
 * (GpuBroadphaseReports::found/lost, scene.h:78-91; worlds with scenes
 * have no found/lost reports and reject 1 and 2). */
int fyzz_hip_world_broadphase_reports(fyzz_hip_world* world, int which, uint32_t* out_pairs, uint64_t capacity,
                                      uint64_t* written);
/* Synchronizes; copies the narrowphase and friction state that the next
 * step reads (the state Scene::StepCheckpoint holds for these stages,
 * scene.h:1802-1900, here for checkpoints and parity checks) to HOST
 * arrays, any of which may be NULL: caches_bb[pair_capacity] the box/box
 * manifold caches aligned with the live pair list
 * (fyzz_hip_pcm_cache_bytes() each), caches_pl[n] the plane manifold cache
 * of each box, history[n_pairs + n] and stamps[n_pairs + n] the friction
 * patch history by manager slot (box/box managers aligned with the pair
 * list, then the plane manager of each box) with the step number that
 * wrote each entry; *step = the number of steps taken. A manager of the
 * next step reads its slot's patch when the slot's stamp equals *step. */
int fyzz_hip_world_fetch_persistent(fyzz_hip_world* world, void* caches_bb, uint32_t pair_capacity,
                                    uint32_t* n_pairs, void* caches_pl, fyzz_dy_patch* history, uint32_t* stamps,
                                    uint32_t* step);
/* Synchronizes; the touching flags the next step's island pass and event
 * transitions read (non-zero: the pair touched on the last step):
 * touch_bb[pair_capacity] aligned with the live pair list, touch_pl[n]
 * per box against the floor (HOST arrays, either may be NULL). */
int fyzz_hip_world_fetch_touch(fyzz_hip_world* world, uint32_t* touch_bb, uint32_t pair_capacity, uint32_t* touch_pl);
/* Device-to-device copy of the body records (creation order). */
int fyzz_hip_world_copy_bodies(fyzz_hip_world* world, fyzz_dy_body* dst);
/* Device pointers of the last step's manager list, contact impulses
 * [n][4], partitions and body records; valid until the next step. */
int fyzz_hip_world_device_views(fyzz_hip_world* world, const fyzz_dy_manager** managers,
                                const uint32_t** d_n_managers, const float** impulses,
                                const uint32_t** partitions, fyzz_dy_body** bodies);

/* ---- Solver-order authority (host) --------------------------------------
 * The native scene's accurate and speculative island graphs, kept on the
 * host from each step's broadphase and touch changes, and the solver order
 * they give: the active accurate islands' contact-manager arrays in
 * active-list order, batched into solver tasks that close once they hold
 * at least 128 bodies and one manager. It replaces the reference's host
 * IslandSim bookkeeping (scene.cpp:5119-5952 update_island_connectivity,
 * 4759-5118 lost-edge route repair and splits, 6148-6262
 * activate_published_islands, 6465-6536 deactivate_ready_accurate_islands,
 * 6746-6884 the active lists, 6885-6918 and 7174-7260 the descriptor and
 * task order), which its GPU producer keeps on the host as well. Bodies
 * are numbered as the scene numbers them (one shape per body), all
 * admitted before the first step; pairs are (low, high) body ids. */
typedef struct fyzz_hip_islands fyzz_hip_islands;

/* One step's changes, every pair list ascending (low, high) as the
 * broadphase reports them (scene.cpp:3233-3300 publish_pairs, 4640-4698
 * update_contacts): pairs_found / pairs_lost the broadphase pairs that
 * appeared / disappeared this step (static/dynamic pairs included),
 * touch_found / touch_lost the managers whose contact count became
 * non-zero / zero (a pair lost from the broadphase while touching is in
 * both pairs_lost and touch_lost). ready[n_bodies] (may be NULL: keep the
 * flags held) is the previous solve's sleep check per body. */
typedef struct fyzz_island_events {
  const uint32_t* pairs_found;
  uint32_t pairs_found_count;
  const uint32_t* pairs_lost;
  uint32_t pairs_lost_count;
  const uint32_t* touch_found;
  uint32_t touch_found_count;
  const uint32_t* touch_lost;
  uint32_t touch_lost_count;
  const uint8_t* ready;
} fyzz_island_events;

/* The step's solver order: managers[manager_capacity][2] the touching
 * pairs in descriptor order, task_ends[task_capacity] the manager end of
 * each solver task, sleeping[n_bodies] (may be NULL) each body's sleep
 * state after the step's island pass. The counts are written even when a
 * capacity is too small (FYZZ_HIP_E_OVERFLOW; the authority has still
 * advanced). */
typedef struct fyzz_island_order {
  uint32_t* managers;
  uint32_t manager_capacity;
  uint32_t manager_count;
  uint32_t* task_ends;
  uint32_t task_capacity;
  uint32_t task_count;
  uint8_t* sleeping;
} fyzz_island_order;

/* dynamic[body_count]: non-zero for a dynamic body, zero for a static one
 * (Scene::add_dynamic_box / add_static_floor, scene.cpp:584-730). Every
 * body starts awake. */
int fyzz_hip_islands_create(uint32_t body_count, const uint8_t* dynamic, fyzz_hip_islands** out);
void fyzz_hip_islands_destroy(fyzz_hip_islands* islands);
/* Applies one step's changes in the scene's own order (broadphase
 * publication, touches, connectivity, wakes, lost-edge repair, sleeps) and
 * writes the solver order. Pairs that are not ascending, a lost or touching
 * pair the authority does not hold, or a found pair it already holds are
 * FYZZ_HIP_E_INVALID_ARGUMENT; after any error the authority refuses
 * further steps (FYZZ_HIP_E_LOGIC), as part of the step may have applied. */
int fyzz_hip_islands_step(fyzz_hip_islands* islands, const fyzz_island_events* events, fyzz_island_order* order);
const char* fyzz_hip_islands_last_error(const fyzz_hip_islands* islands);

#ifdef __cplusplus
}
#endif
#endif /* FYZZ_HIP_H */
