// C-ABI driver over the compiled reference NativeCpu scene (fyzz::native::Scene).
//
// Test infrastructure only: it is linked against the reference's own
// sources (see oracle/Makefile, target `ref`) into oracle/_ref/, and is
// used in this container to write golden vectors and to pin the oracle
// restatement. Nothing in the product links or loads it.
//
// Reference interfaces used (paths relative to the reference root):
//   native::Scene public API             src/native/scene.h:633-1077
//   SourceSapEndpointFixture friend      src/native/scene.h:1085
//   sorted endpoint arrays               src/native/scene.h:1295-1304
//   PCM free functions                   src/native/scene.h:594-620
#include "native/scene.h"

#include <cstdint>
#include <cstring>
#include <exception>
#include <memory>
#include <string>

namespace fyzz::native {
// The reference declares this name a friend of Scene so that tests can
// read the sorted SAP endpoint arrays; the driver uses the same hook.
struct SourceSapEndpointFixture {
    static std::size_t endpoints(const Scene& s, unsigned axis,
                                 std::uint32_t* value, std::uint32_t* id,
                                 std::uint8_t* is_max, std::size_t cap) {
        const auto& e = s.sap_axis_endpoints_[axis];
        if (value) {
            const std::size_t n = e.size() < cap ? e.size() : cap;
            for (std::size_t i = 0; i < n; ++i) {
                value[i] = e[i].value;
                id[i] = e[i].id;
                is_max[i] = e[i].is_max ? 1 : 0;
            }
        }
        return e.size();
    }
    static void update_bounds(Scene& s) { s.update_bounds(); }
    static void update_pairs(Scene& s) { s.update_pairs(); }
    static std::size_t pending_pairs(const Scene& s, std::uint32_t* out,
                                     std::size_t cap) {
        const auto& p = s.pending_pairs_;
        if (out)
            for (std::size_t i = 0; i < p.size() && i < cap; ++i) {
                out[2 * i] = p[i].a;
                out[2 * i + 1] = p[i].b;
            }
        return p.size();
    }
    static std::size_t created_deleted(const Scene& s, bool created,
                                       std::uint32_t* out, std::size_t cap) {
        const auto& p = created ? s.sap_created_pairs_ : s.sap_deleted_pairs_;
        if (out)
            for (std::size_t i = 0; i < p.size() && i < cap; ++i) {
                out[2 * i] = p[i].a;
                out[2 * i + 1] = p[i].b;
            }
        return p.size();
    }
};
// The reference also names this fixture a friend (scene.h:1083); the
// driver reads the Dy descriptor stream, its conflict partitions and the
// solver task boundaries of the last step through it.
struct SolverIterationSourceFixture {
    static std::size_t descriptors(const Scene& s, std::uint32_t* index, std::uint32_t* body0,
                                   std::uint32_t* body1, std::uint64_t* partition,
                                   std::size_t cap) {
        const auto& d = s.constraint_descriptors_;
        if (index)
            for (std::size_t i = 0; i < d.size() && i < cap; ++i) {
                index[i] = d[i].index;
                body0[i] = d[i].body0;
                body1[i] = d[i].body1;
                partition[i] = i < s.constraint_partitions_.size()
                                   ? static_cast<std::uint64_t>(s.constraint_partitions_[i])
                                   : ~0ull;
            }
        return d.size();
    }
    static std::size_t tasks(const Scene& s, std::uint64_t* manager_end,
                             std::uint64_t* partition_end, std::size_t cap) {
        const auto& t = s.solver_tasks_;
        if (manager_end)
            for (std::size_t i = 0; i < t.size() && i < cap; ++i) {
                manager_end[i] = t[i].manager_end;
                partition_end[i] = t[i].partition_end;
            }
        return t.size();
    }
};
}  // namespace fyzz::native

namespace {
using fyzz::Quat;
using fyzz::Vec3;
using fyzz::native::Scene;
thread_local std::string g_error;

template <class F>
int guarded(F&& f) {
    try {
        f();
        return 0;
    } catch (const std::exception& e) {
        g_error = e.what();
        return -1;
    } catch (...) {
        g_error = "unknown exception";
        return -1;
    }
}
std::size_t copy_pairs(std::span<const fyzz::native::Pair> p,
                       std::uint32_t* out, std::size_t cap) {
    if (out)
        for (std::size_t i = 0; i < p.size() && i < cap; ++i) {
            out[2 * i] = p[i].a;
            out[2 * i + 1] = p[i].b;
        }
    return p.size();
}
}  // namespace

extern "C" {

const char* ref_last_error() { return g_error.c_str(); }

void* ref_scene_create() {
    try {
        return new Scene();
    } catch (...) {
        return nullptr;
    }
}
void ref_scene_destroy(void* s) { delete static_cast<Scene*>(s); }

int ref_add_floor(void* s, float height, std::uint32_t* id) {
    return guarded([&] { *id = static_cast<Scene*>(s)->add_static_floor(height); });
}
int ref_add_dynamic_box(void* s, const float* half, const float* pos,
                        float density, std::uint32_t* id) {
    return guarded([&] {
        *id = static_cast<Scene*>(s)->add_dynamic_box(
            {half[0], half[1], half[2]}, {pos[0], pos[1], pos[2]}, density);
    });
}
int ref_add_dynamic_sphere(void* s, float radius, const float* pos, float density, std::uint32_t* id) {
    return guarded([&] {
        *id = static_cast<Scene*>(s)->add_dynamic_sphere(radius, {pos[0], pos[1], pos[2]}, density);
    });
}
// n axis-aligned dynamic boxes (the broadphase workload's load); ids[n]
// receives the body ids.
int ref_add_aabbs(void* s, std::uint32_t n, const float* half, const float* pos, float density,
                  std::uint32_t* ids) {
    return guarded([&] {
        auto* sc = static_cast<Scene*>(s);
        for (std::uint32_t i = 0; i < n; ++i) {
            const float* h = half + 3ull * i;
            const float* p = pos + 3ull * i;
            ids[i] = sc->add_dynamic_box({h[0], h[1], h[2]}, {p[0], p[1], p[2]}, density);
        }
    });
}
// n dynamic boxes with pose, velocities and material (the bench's scene
// load); ids[n] receives the body ids.
int ref_add_boxes(void* s, std::uint32_t n, const float* half, const float* pos, const float* quat,
                  const float* vel, const float* ang, float density, float sf, float df, float rest,
                  std::uint32_t* ids) {
    return guarded([&] {
        auto* sc = static_cast<Scene*>(s);
        fyzz::Material m;
        m.static_friction = sf;
        m.dynamic_friction = df;
        m.restitution = rest;
        for (std::uint32_t i = 0; i < n; ++i) {
            const float* h = half + 3ull * i;
            const float* p = pos + 3ull * i;
            const float* q = quat + 4ull * i;
            const std::uint32_t id = sc->add_dynamic_box({h[0], h[1], h[2]}, {p[0], p[1], p[2]}, density);
            sc->set_pose(id, {p[0], p[1], p[2]}, {q[0], q[1], q[2], q[3]});
            if (vel) sc->set_velocity(id, {vel[3 * i], vel[3 * i + 1], vel[3 * i + 2]});
            if (ang) sc->set_angular_velocity(id, {ang[3 * i], ang[3 * i + 1], ang[3 * i + 2]});
            sc->set_material(id, m);
            ids[i] = id;
        }
    });
}
int ref_set_contact_notifications(void* s, int enabled) {
    return guarded([&] { static_cast<Scene*>(s)->set_contact_notifications(enabled != 0); });
}
// The last step's contact events: ev[n][6] u32 (body0, body1, kind, count,
// shape0, shape1), ev_offset[n] u64, ev_impulse[n] f64; points pt[m][7]
// f32 (position, normal, separation), pt_impulse[m] f64. Returns the event
// count; *n_points receives the point count. Null outputs query sizes.
std::size_t ref_contact_events(void* s, std::uint32_t* ev, std::uint64_t* ev_offset, double* ev_impulse,
                               std::size_t cap, float* pt, double* pt_impulse, std::size_t pt_cap,
                               std::size_t* n_points) {
    auto* sc = static_cast<Scene*>(s);
    const auto events = sc->contact_events();
    const auto points = sc->contact_event_points();
    if (n_points) *n_points = points.size();
    if (ev)
        for (std::size_t i = 0; i < events.size() && i < cap; ++i) {
            const auto& e = events[i];
            ev[6 * i] = e.body0;
            ev[6 * i + 1] = e.body1;
            ev[6 * i + 2] = static_cast<std::uint32_t>(e.kind);
            ev[6 * i + 3] = e.contact_count;
            ev[6 * i + 4] = e.shape0;
            ev[6 * i + 5] = e.shape1;
            ev_offset[i] = e.point_offset;
            ev_impulse[i] = e.normal_impulse;
        }
    if (pt)
        for (std::size_t i = 0; i < points.size() && i < pt_cap; ++i) {
            const auto& q = points[i];
            const float v[7] = {q.position.x, q.position.y, q.position.z, q.normal.x, q.normal.y, q.normal.z,
                                q.separation};
            for (int k = 0; k < 7; ++k) pt[7 * i + k] = v[k];
            pt_impulse[i] = q.normal_impulse;
        }
    return events.size();
}
int ref_add_static_box(void* s, const float* half, const float* pos,
                       std::uint32_t* id) {
    return guarded([&] {
        *id = static_cast<Scene*>(s)->add_static_box(
            {half[0], half[1], half[2]}, {pos[0], pos[1], pos[2]});
    });
}
int ref_add_static_sphere(void* s, float radius, const float* pos, std::uint32_t* id) {
    return guarded([&] { *id = static_cast<Scene*>(s)->add_static_sphere(radius, {pos[0], pos[1], pos[2]}); });
}
int ref_set_pose(void* s, std::uint32_t id, const float* pos, const float* q) {
    return guarded([&] {
        static_cast<Scene*>(s)->set_pose(id, {pos[0], pos[1], pos[2]},
                                         {q[0], q[1], q[2], q[3]});
    });
}
int ref_set_position(void* s, std::uint32_t id, const float* pos) {
    return guarded([&] {
        static_cast<Scene*>(s)->set_position(id, {pos[0], pos[1], pos[2]});
    });
}
int ref_set_velocity(void* s, std::uint32_t id, const float* v) {
    return guarded([&] {
        static_cast<Scene*>(s)->set_velocity(id, {v[0], v[1], v[2]});
    });
}
int ref_set_angular_velocity(void* s, std::uint32_t id, const float* v) {
    return guarded([&] {
        static_cast<Scene*>(s)->set_angular_velocity(id, {v[0], v[1], v[2]});
    });
}
// static friction, dynamic friction, restitution; combine modes default.
int ref_set_material(void* s, std::uint32_t id, float sf, float df, float rest) {
    return guarded([&] {
        fyzz::Material m;
        m.static_friction = sf;
        m.dynamic_friction = df;
        m.restitution = rest;
        static_cast<Scene*>(s)->set_material(id, m);
    });
}
int ref_set_gravity_enabled(void* s, std::uint32_t id, int enabled) {
    return guarded([&] {
        auto* sc = static_cast<Scene*>(s);
        fyzz::RigidBodyDynamics d;
        d.gravity_enabled = enabled != 0;
        sc->set_dynamics(id, d);
    });
}
// RigidBodyDynamics with damping, speed limits, gravity and iteration
// counts set (f = linear damping, angular damping, max linear velocity,
// max angular velocity); the other fields keep their defaults.
int ref_set_dynamics(void* s, std::uint32_t id, const float* f, int gravity, int position_iterations,
                     int velocity_iterations) {
    return guarded([&] {
        fyzz::RigidBodyDynamics d;
        d.linear_damping = f[0];
        d.angular_damping = f[1];
        d.max_linear_velocity = f[2];
        d.max_angular_velocity = f[3];
        d.gravity_enabled = gravity != 0;
        d.position_iterations = static_cast<std::uint8_t>(position_iterations);
        d.velocity_iterations = static_cast<std::uint8_t>(velocity_iterations);
        static_cast<Scene*>(s)->set_dynamics(id, d);
    });
}
int ref_advance_collision(void* s, int force) {
    return guarded([&] { static_cast<Scene*>(s)->advance_collision(force != 0); });
}
int ref_step(void* s, float dt) {
    return guarded([&] { static_cast<Scene*>(s)->step(dt); });
}
// SAP only: bounds + endpoints, then pairs; no narrowphase or islands.
int ref_sap_update(void* s) {
    return guarded([&] {
        auto& sc = *static_cast<Scene*>(s);
        fyzz::native::SourceSapEndpointFixture::update_bounds(sc);
        fyzz::native::SourceSapEndpointFixture::update_pairs(sc);
    });
}
std::size_t ref_sap_pending_pairs(void* s, std::uint32_t* out, std::size_t cap) {
    return fyzz::native::SourceSapEndpointFixture::pending_pairs(
        *static_cast<Scene*>(s), out, cap);
}
std::size_t ref_sap_created(void* s, std::uint32_t* out, std::size_t cap) {
    return fyzz::native::SourceSapEndpointFixture::created_deleted(
        *static_cast<Scene*>(s), true, out, cap);
}
std::size_t ref_sap_deleted(void* s, std::uint32_t* out, std::size_t cap) {
    return fyzz::native::SourceSapEndpointFixture::created_deleted(
        *static_cast<Scene*>(s), false, out, cap);
}
std::size_t ref_endpoints(void* s, unsigned axis, std::uint32_t* value,
                          std::uint32_t* id, std::uint8_t* is_max,
                          std::size_t cap) {
    return fyzz::native::SourceSapEndpointFixture::endpoints(
        *static_cast<Scene*>(s), axis, value, id, is_max, cap);
}
std::size_t ref_pairs(void* s, std::uint32_t* out, std::size_t cap) {
    return copy_pairs(static_cast<Scene*>(s)->pairs(), out, cap);
}
std::size_t ref_found(void* s, std::uint32_t* out, std::size_t cap) {
    return copy_pairs(static_cast<Scene*>(s)->found_pairs(), out, cap);
}
std::size_t ref_lost(void* s, std::uint32_t* out, std::size_t cap) {
    return copy_pairs(static_cast<Scene*>(s)->lost_pairs(), out, cap);
}
// Body state rows of 14 floats: p(3) v(3) q(4) w(3) sleeping(1).
int ref_body_state(void* s, std::uint32_t id, float* out) {
    return guarded([&] {
        const auto b = static_cast<Scene*>(s)->body_state(id);
        const float v[14] = {b.position.x, b.position.y, b.position.z,
                             b.velocity.x, b.velocity.y, b.velocity.z,
                             b.orientation.x, b.orientation.y, b.orientation.z,
                             b.orientation.w, b.angular_velocity.x,
                             b.angular_velocity.y, b.angular_velocity.z,
                             b.sleeping ? 1.0f : 0.0f};
        std::memcpy(out, v, sizeof v);
    });
}
// Manifold table: per manifold (pair.a, pair.b, count) and up to count
// points of 7 floats (position, normal, separation). Returns #manifolds.
std::size_t ref_manifolds(void* s, std::uint32_t* pairs_out, std::uint8_t* counts,
                          float* points, std::size_t cap) {
    const auto m = static_cast<Scene*>(s)->contacts();
    if (pairs_out)
        for (std::size_t i = 0; i < m.size() && i < cap; ++i) {
            pairs_out[2 * i] = m[i].pair.a;
            pairs_out[2 * i + 1] = m[i].pair.b;
            counts[i] = m[i].count;
            for (std::size_t k = 0; k < m[i].count && k < 4; ++k) {
                const auto& c = m[i].points[k];
                float* o = points + (4 * i + k) * 7;
                o[0] = c.position.x; o[1] = c.position.y; o[2] = c.position.z;
                o[3] = c.normal.x; o[4] = c.normal.y; o[5] = c.normal.z;
                o[6] = c.separation;
            }
        }
    return m.size();
}
// Scene::solved_friction_patches (scene.h:846-848): per patch the shape
// pair (2 u32), anchor count and broken flag (2 u32), anchor positions
// f32[2][3] and world impulses f64[2][3]. Returns the patch count.
std::size_t ref_friction_patches(void* s, std::uint32_t* pair_count_broken, float* positions, double* impulses,
                                 std::size_t cap) {
    const auto w = static_cast<Scene*>(s)->solved_friction_patches();
    if (pair_count_broken)
        for (std::size_t i = 0; i < w.size() && i < cap; ++i) {
            const auto& p = w[i];
            pair_count_broken[4 * i] = p.pair.a;
            pair_count_broken[4 * i + 1] = p.pair.b;
            pair_count_broken[4 * i + 2] = p.anchor_count;
            pair_count_broken[4 * i + 3] = p.broken ? 1u : 0u;
            for (int k = 0; k < 2; ++k) {
                positions[6 * i + 3 * k] = p.anchor_positions[k].x;
                positions[6 * i + 3 * k + 1] = p.anchor_positions[k].y;
                positions[6 * i + 3 * k + 2] = p.anchor_positions[k].z;
                impulses[6 * i + 3 * k] = p.anchor_impulses[k].x;
                impulses[6 * i + 3 * k + 1] = p.anchor_impulses[k].y;
                impulses[6 * i + 3 * k + 2] = p.anchor_impulses[k].z;
            }
        }
    return w.size();
}
std::size_t ref_indexed_order(void* s, std::uint64_t* out, std::size_t cap) {
    const auto o = static_cast<Scene*>(s)->indexed_contact_manager_order();
    if (out)
        for (std::size_t i = 0; i < o.size() && i < cap; ++i) out[i] = o[i];
    return o.size();
}
std::size_t ref_dy_descriptors(void* s, std::uint32_t* index, std::uint32_t* body0,
                               std::uint32_t* body1, std::uint64_t* partition, std::size_t cap) {
    return fyzz::native::SolverIterationSourceFixture::descriptors(
        *static_cast<Scene*>(s), index, body0, body1, partition, cap);
}
std::size_t ref_dy_tasks(void* s, std::uint64_t* manager_end, std::uint64_t* partition_end,
                         std::size_t cap) {
    return fyzz::native::SolverIterationSourceFixture::tasks(*static_cast<Scene*>(s), manager_end,
                                                             partition_end, cap);
}
// wake counter and ready-for-sleep of a body (scene.h:790-799)
int ref_sleep_state(void* s, std::uint32_t id, float* wake_counter, int* ready) {
    return guarded([&] {
        const auto* sc = static_cast<Scene*>(s);
        *wake_counter = sc->body_wake_counter(id);
        *ready = sc->body_ready_for_sleep(id) ? 1 : 0;
    });
}
std::size_t ref_partition_count(void* s) {
    return static_cast<Scene*>(s)->solver_partition_count();
}

// --- Gu PCM free functions (GuPCMContactBoxBox.cpp, GuPCMGjkEpa.cpp) ---
// Pose layout: half(3) position(3) quat(4) = 10 floats per box.
// Candidate layout: local_a(3) local_b(3) normal_b(3) separation(1).
int ref_pcm_box_box_candidates(const float* a, const float* b, float contact_distance,
                               float* cand_out, std::uint32_t cap,
                               std::uint32_t* count, int* ok) {
    return guarded([&] {
        std::vector<fyzz::native::PcmContactCandidate> c(cap);
        std::size_t n = 0;
        const bool r = fyzz::native::source_pcm_box_box_candidates(
            {a[0], a[1], a[2]}, {a[3], a[4], a[5]}, {a[6], a[7], a[8], a[9]},
            {b[0], b[1], b[2]}, {b[3], b[4], b[5]}, {b[6], b[7], b[8], b[9]},
            contact_distance, c, n);
        *ok = r ? 1 : 0;
        *count = static_cast<std::uint32_t>(n);
        for (std::size_t i = 0; i < n && i < cap; ++i) {
            float* o = cand_out + 10 * i;
            o[0] = c[i].local_a.x; o[1] = c[i].local_a.y; o[2] = c[i].local_a.z;
            o[3] = c[i].local_b.x; o[4] = c[i].local_b.y; o[5] = c[i].local_b.z;
            o[6] = c[i].normal_b.x; o[7] = c[i].normal_b.y; o[8] = c[i].normal_b.z;
            o[9] = c[i].separation;
        }
    });
}
int ref_pcm_batch_selection(const float* cand, std::uint32_t n, float tolerance,
                            std::uint8_t* sel4) {
    return guarded([&] {
        std::vector<fyzz::native::PcmContactCandidate> c(n);
        for (std::uint32_t i = 0; i < n; ++i) {
            const float* o = cand + 10 * i;
            c[i].local_a = {o[0], o[1], o[2]};
            c[i].local_b = {o[3], o[4], o[5]};
            c[i].normal_b = {o[6], o[7], o[8]};
            c[i].separation = o[9];
        }
        const auto s = fyzz::native::source_pcm_box_batch_selection(c, tolerance);
        for (int k = 0; k < 4; ++k) sel4[k] = s[k];
    });
}
int ref_pcm_gjk_epa(const float* a, const float* b, float contact_distance,
                    float* cand_out, int* ok) {
    return guarded([&] {
        fyzz::native::PcmContactCandidate c{};
        const bool r = fyzz::native::source_pcm_box_box_gjk_epa(
            {a[0], a[1], a[2]}, {a[3], a[4], a[5]}, {a[6], a[7], a[8], a[9]},
            {b[0], b[1], b[2]}, {b[3], b[4], b[5]}, {b[6], b[7], b[8], b[9]},
            contact_distance, c, nullptr);
        *ok = r ? 1 : 0;
        cand_out[0] = c.local_a.x; cand_out[1] = c.local_a.y; cand_out[2] = c.local_a.z;
        cand_out[3] = c.local_b.x; cand_out[4] = c.local_b.y; cand_out[5] = c.local_b.z;
        cand_out[6] = c.normal_b.x; cand_out[7] = c.normal_b.y; cand_out[8] = c.normal_b.z;
        cand_out[9] = c.separation;
    });
}

}  // extern "C"
