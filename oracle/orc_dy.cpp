// Oracle Dy restatement: unconstrained pre-integration, solver bodies,
// conflict partitions, contact preparation with correlated friction and
// anchor growth, PGS (4 position + 1 velocity iterations by default),
// integration and the wake-counter sleep check, for one step of the
// admitted scene class (dynamic boxes + static planes, PCM manifolds,
// non-compliant materials, no joints, locks, slop or gyroscopic terms).
// Templated on the scalar: T = float follows the reference's FP32
// operation order exactly; T = double is the FP64 twin used to state each
// FP32 result's error. Test infrastructure only.
//
// Reference (paths relative to the reference root, src/native/scene.cpp):
//   prepare_unconstrained_velocities        10383-10507
//   prepare_solver_bodies                   6940-7112
//   partition_constraints                   7113-7489
//   prepare_constraints (contacts)          7490-7515, 7763-8511
//   cached_friction_patch_valid             3981-4010
//   solve_constraint_task                   8853-9148
//   publish_solver_bodies                   9509-9549
//   integrate_solver_task, integrate_quaternion  10508-10707, 111-146
#include <algorithm>
#include <array>
#include <bit>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <vector>

#include "../include/fyzz_hip.h"
#include "orc.h"
#include "orc_math.h"

namespace {

template <class T>
struct Dy {
    using V = orc::V3<T>;
    using Q = orc::Q4<T>;
    static V v3(const float* p) { return {T(p[0]), T(p[1]), T(p[2])}; }
    static Q q4(const float* p) { return {T(p[0]), T(p[1]), T(p[2]), T(p[3])}; }
    static V sub(V a, V b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
    static T dot(V a, V b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
    static T len2(V v) { return v.x * v.x + v.y * v.y + v.z * v.z; }
    static bool finite(V v) { return std::isfinite(v.x) && std::isfinite(v.y) && std::isfinite(v.z); }
    static bool finite(Q q) {
        return std::isfinite(q.x) && std::isfinite(q.y) && std::isfinite(q.z) && std::isfinite(q.w);
    }
    // mass pose with the identity mass-local frame of a box (scene.cpp:359-365)
    static V mass_position(V p, Q q) {
        const V r = orc::rotate(q, V{0, 0, 0});
        return {p.x + r.x, p.y + r.y, p.z + r.z};
    }
    static Q mass_orientation(Q q) { return orc::multiply(q, Q{0, 0, 0, 1}); }

    struct Row {  // SolverBodyRow (scene.h:1188-1232)
        V lin_delta{}, ang_state{};
        std::uint32_t mask = 0;
        std::size_t max_dyn = 0, static_count = 0;
        V v{}, w{};
        T inv_mass = 0;
        V sii[3]{};  // sqrt inverse inertia columns
        V pos{};
        Q q{0, 0, 0, 1};
        V motion_lin{}, motion_ang{};
    };
    struct ContactRow {  // SolverContactPoint
        V n, ra, rb;
        T vm, biased, unbiased, applied, max_impulse;
        // velocity dependence, kept for the frozen-row rollout (not used by
        // the step): err = kv * vrel * vm - c, vrel = n.(va - vb) + raxn.wa - rbxn.wb
        V raxn{}, rbxn{};
        T kv = 0, cb = 0, cu = 0;
    };
    struct FrictionRow {  // SolverContactFriction
        V t, ra, rb;
        T vm, bias, target, applied;
        V ax0{}, ax1{};  // unscaled lever arms: target = -(t.(va - vb) + ax0.wa - ax1.wb)
    };
    struct Patch {  // FrictionPatch (fields the admitted class needs)
        V normal{}, n0{}, n1{};
        T sf = 0, df = 0, rest = 0, damping = 0;
        V a0[2]{}, a1[2]{};
        V bmin{}, bmax{};
        std::uint32_t contacts = 0, anchors = 0;
        bool broken = false, disable_strong = false;
        std::size_t row_offset = 0;
    };
    struct BodyView {  // actor pose for body-local friction frames
        V pos;
        Q q;
    };

    // scene.cpp:10383-10507, default dynamics (no velocity mods pending)
    static void pre_integrate(fyzz_dy_body& b, T dt, V& v, V& w) {
        const T gravity = T(-9.81f);
        V pre{T(b.linear_velocity[0]) + T(0) + T(0) * dt, T(b.linear_velocity[1]) + T(0) + T(0) * dt,
              T(b.linear_velocity[2]) + T(0) + T(0) * dt};
        T vx = pre.x;
        T vy = pre.y + ((b.flags & 4u) ? gravity * dt : T(0));
        T vz = pre.z;
        const T lm = std::max(T(1) - T(b.linear_damping) * dt, T(0));
        vx *= lm;
        vy *= lm;
        vz *= lm;
        V ang{T(b.angular_velocity[0]) + T(0) + T(0) * dt, T(b.angular_velocity[1]) + T(0) + T(0) * dt,
              T(b.angular_velocity[2]) + T(0) + T(0) * dt};
        const T am = std::max(T(1) - T(b.angular_damping) * dt, T(0));
        ang.x *= am;
        ang.y *= am;
        ang.z *= am;
        const T s2 = vx * vx + vy * vy + vz * vz;
        const T ml2 = T(b.max_linear_velocity) * T(b.max_linear_velocity);
        const T ma2 = T(b.max_angular_velocity) * T(b.max_angular_velocity);
        if (s2 > ml2) {
            const T f = std::sqrt(ml2 / s2);
            vx *= f;
            vy *= f;
            vz *= f;
        }
        const T a2 = ang.x * ang.x + ang.y * ang.y + ang.z * ang.z;
        if (a2 > ma2) {
            const T f = std::sqrt(ma2 / a2);
            ang.x *= f;
            ang.y *= f;
            ang.z *= f;
        }
        v = {vx, vy, vz};
        w = ang;
    }

    // scene.cpp:7092-7110: R diag(sqrt I^-1) R^T, identity fast path
    static void world_sqrt_inertia(const fyzz_dy_body& b, Q q, V out[3]) {
        const V local{std::sqrt(T(b.inverse_inertia[0])), std::sqrt(T(b.inverse_inertia[1])),
                      std::sqrt(T(b.inverse_inertia[2]))};
        if (q.x == T(0) && q.y == T(0) && q.z == T(0) && q.w == T(1)) {
            out[0] = {local.x, 0, 0};
            out[1] = {0, local.y, 0};
            out[2] = {0, 0, local.z};
            return;
        }
        const V rx = orc::rotate(q, V{1, 0, 0});
        const V ry = orc::rotate(q, V{0, 1, 0});
        const V rz = orc::rotate(q, V{0, 0, 1});
        const T xx = local.x * rx.x * rx.x + local.y * ry.x * ry.x + local.z * rz.x * rz.x;
        const T yy = local.x * rx.y * rx.y + local.y * ry.y * ry.y + local.z * rz.y * rz.y;
        const T zz = local.x * rx.z * rx.z + local.y * ry.z * ry.z + local.z * rz.z * rz.z;
        const T xy = local.x * rx.x * rx.y + local.y * ry.x * ry.y + local.z * rz.x * rz.y;
        const T xz = local.x * rx.x * rx.z + local.y * ry.x * ry.z + local.z * rz.x * rz.z;
        const T yz = local.x * rx.y * rx.z + local.y * ry.y * ry.z + local.z * rz.y * rz.z;
        out[0] = {xx, xy, xz};
        out[1] = {xy, yy, yz};
        out[2] = {xz, yz, zz};
    }
    static V transform(const Row& d, V v) {
        const V* c = d.sii;
        return {c[0].x * v.x + c[1].x * v.y + c[2].x * v.z, c[0].y * v.x + c[1].y * v.y + c[2].y * v.z,
                c[0].z * v.x + c[1].z * v.y + c[2].z * v.z};
    }
    static V local_point(const BodyView& a, V world) {
        const Q q = mass_orientation(a.q);
        const V p = mass_position(a.pos, a.q);
        return orc::rotate(Q{-q.x, -q.y, -q.z, q.w}, V{world.x - p.x, world.y - p.y, world.z - p.z});
    }
    static V local_normal(const BodyView& a, V n) {
        const Q q = mass_orientation(a.q);
        return orc::rotate(Q{-q.x, -q.y, -q.z, q.w}, n);
    }
    static V world_point(const BodyView& a, V local) {
        const V r = orc::rotate(mass_orientation(a.q), local);
        const V p = mass_position(a.pos, a.q);
        return {r.x + p.x, r.y + p.y, r.z + p.z};
    }
    // scene.cpp:3981-4010
    static bool history_valid(const Patch& old, const BodyView& a0, const BodyView& a1) {
        if (old.broken || old.disable_strong || old.anchors == 0u) return false;
        const T same_normal = T(0.999f);
        const T corr = T(0.025f);
        const Q q0 = mass_orientation(a0.q), q1 = mass_orientation(a1.q);
        const V n10 = orc::rotate_inverse(q0, orc::rotate(q1, old.n1));
        if (dot(old.n0, n10) <= same_normal) return false;
        const V p0 = mass_position(a0.pos, a0.q), p1 = mass_position(a1.pos, a1.q);
        for (std::uint32_t k = 0; k < old.anchors; ++k) {
            const V l = orc::rotate(q1, old.a1[k]);
            const V world{l.x + p1.x, l.y + p1.y, l.z + p1.z};
            const V in0 = orc::rotate_inverse(q0, V{world.x - p0.x, world.y - p0.y, world.z - p0.z});
            const V d{old.a0[k].x - in0.x, old.a0[k].y - in0.y, old.a0[k].z - in0.z};
            if (std::abs(dot(d, old.n0)) >= corr) return false;
        }
        return true;
    }
    // scene.cpp:111-146
    static bool integrate_quaternion(Q pose, V w, T dt, Q& out) {
        T ww = w.x * w.x + w.y * w.y + w.z * w.z;
        if (ww != T(0)) {
            ww = std::sqrt(ww);
            const T max_w = T(1.0e7f);
            if (ww > max_w) {
                const T s = max_w / ww;
                w = {w.x * s, w.y * s, w.z * s};
                ww = max_w;
            }
            const T v = dt * ww * T(0.5f);
            const T s = std::sin(v) / ww;
            const T c = std::cos(v);
            const Q dq = orc::multiply(Q{w.x * s, w.y * s, w.z * s, T(0)}, pose);
            const Q nx{dq.x + pose.x * c, dq.y + pose.y * c, dq.z + pose.z * c, dq.w + pose.w * c};
            const T len = std::sqrt(nx.x * nx.x + nx.y * nx.y + nx.z * nx.z + nx.w * nx.w);
            if (!(len > T(0)) || !std::isfinite(len)) return false;
            pose = {nx.x / len, nx.y / len, nx.z / len, nx.w / len};
        }
        out = pose;
        return true;
    }

    // Full step. bodies_out64 (optional) receives the FP64 twin's state.
    static int step(T dt, std::uint32_t nb, fyzz_dy_body* bodies, std::uint32_t ns, const fyzz_dy_static* statics,
                    std::uint32_t nm, const fyzz_dy_manager* mgr, const fyzz_dy_patch* history,
                    std::uint32_t nh, fyzz_dy_patch* patches_out, float* impulses_out, std::uint32_t* parts_out,
                    double* state_out) {
        const T inv_dt = T(1) / dt;
        // ---- pre-integration (awake bodies)
        std::vector<V> v(nb), w(nb);
        for (std::uint32_t i = 0; i < nb; ++i) {
            if (bodies[i].flags & 1u) continue;
            pre_integrate(bodies[i], dt, v[i], w[i]);
            if (!finite(v[i]) || !finite(w[i])) return -3;
        }
        auto view = [&](int32_t id) -> BodyView {
            if (id >= 0) return {v3(bodies[id].position), q4(bodies[id].orientation)};
            const fyzz_dy_static& s = statics[-1 - id];
            return {v3(s.position), q4(s.orientation)};
        };
        // ---- solver rows for bodies of touching managers (scene.cpp:6987-7013, 7051-7110)
        std::vector<Row> rows(nb);
        std::vector<char> has_row(nb, 0);
        for (std::uint32_t m = 0; m < nm; ++m)
            for (int32_t id : {mgr[m].body0, mgr[m].body1})
                if (id >= 0 && !(bodies[id].flags & 1u)) has_row[id] = 1;
        for (std::uint32_t i = 0; i < nb; ++i) {
            if (!has_row[i]) continue;
            Row& r = rows[i];
            const fyzz_dy_body& b = bodies[i];
            const V p = v3(b.position);
            const Q q = q4(b.orientation);
            r.pos = mass_position(p, q);
            r.q = mass_orientation(q);
            r.v = v[i];
            r.w = w[i];
            r.inv_mass = T(b.inverse_mass);
            world_sqrt_inertia(b, r.q, r.sii);
        }
        Row world{};  // the static-world solver body: zero response
        auto row_of = [&](int32_t id) -> Row& { return id >= 0 ? rows[id] : world; };

        // ---- partitions (scene.cpp:7352-7445), one task over the list
        std::vector<std::size_t> part(nm, ~std::size_t(0));
        std::vector<std::uint32_t> pending;
        for (std::uint32_t m = 0; m < nm; ++m)
            if (mgr[m].body0 >= 0 && mgr[m].body1 >= 0) pending.push_back(m);
        for (std::size_t wave = 0; !pending.empty(); wave += 32) {
            if (wave)
                for (std::uint32_t m : pending) {
                    rows[mgr[m].body0].mask = 0;
                    rows[mgr[m].body1].mask = 0;
                }
            std::size_t over = 0;
            for (std::size_t k = 0; k < pending.size(); ++k) {
                const std::uint32_t m = pending[k];
                Row& a = rows[mgr[m].body0];
                Row& b = rows[mgr[m].body1];
                const std::uint32_t avail = ~(a.mask | b.mask);
                if (!avail) {
                    pending[over++] = m;
                    continue;
                }
                const unsigned bit = std::countr_zero(avail);
                a.mask |= 1u << bit;
                b.mask |= 1u << bit;
                const std::size_t lp = wave + bit;
                a.max_dyn = std::max(a.max_dyn, lp + 1);
                b.max_dyn = std::max(b.max_dyn, lp + 1);
                part[m] = lp;
            }
            if (over == pending.size()) return -4;
            pending.resize(over);
        }
        for (std::uint32_t m = 0; m < nm; ++m) {
            if (part[m] != ~std::size_t(0)) continue;
            Row& act = rows[mgr[m].body0 >= 0 ? mgr[m].body0 : mgr[m].body1];
            part[m] = act.max_dyn + act.static_count++;
        }

        // ---- contact preparation (scene.cpp:7794-8505)
        const T rigid_contact_bias = T(0.8f);
        const T bounce_threshold = T(-2.0f);
        const T same_normal = T(0.999f);
        const T friction_offset_threshold = T(0.04f);
        std::vector<Patch> patches(nm);
        std::vector<std::vector<ContactRow>> crow(nm);
        std::vector<FrictionRow> frows;
        for (std::uint32_t m = 0; m < nm; ++m) {
            const fyzz_dy_manager& M = mgr[m];
            const BodyView a0 = view(M.body0), a1 = view(M.body1);
            const Row& a = row_of(M.body0);
            const Row& b = row_of(M.body1);
            auto mat = [&](int32_t id, T& sf, T& df, T& rs) {
                if (id >= 0) {
                    sf = T(bodies[id].static_friction);
                    df = T(bodies[id].dynamic_friction);
                    rs = T(bodies[id].restitution);
                } else {
                    sf = T(statics[-1 - id].static_friction);
                    df = T(statics[-1 - id].dynamic_friction);
                    rs = T(statics[-1 - id].restitution);
                }
            };
            T sf0, df0, r0, sf1, df1, r1;
            mat(M.body0, sf0, df0, r0);
            mat(M.body1, sf1, df1, r1);
            if (r0 < T(0) || r1 < T(0)) return -4;  // compliant contacts are not admitted
            const T restitution = T(0.5f) * (r0 + r1);
            const T damping = T(0.5f) * (T(0) + T(0));
            const T dynamic_friction = std::max(T(0), T(0.5f) * (df0 + df1));
            const T static_friction = std::max(dynamic_friction, T(0.5f) * (sf0 + sf1));
            const T max_pen_bias = -T(1.0e32f);
            const T rest_distance = T(0);
            // contiguous same-normal runs (scene.cpp:7964-7997)
            struct Run {
                std::uint32_t start, count;
                V normal, bmin, bmax;
                std::size_t next;
            };
            Run runs[4];
            std::uint32_t nr = 0;
            for (std::uint32_t j = 0; j < M.count; ++j) {
                const V pn = v3(M.contacts[j].normal);
                const V pp = v3(M.contacts[j].position);
                if (nr == 0 || dot(runs[nr - 1].normal, pn) < same_normal) {
                    runs[nr++] = {j, 1, pn, pp, pp, ~std::size_t(0)};
                } else {
                    Run& r = runs[nr - 1];
                    ++r.count;
                    r.bmin = {std::min(r.bmin.x, pp.x), std::min(r.bmin.y, pp.y), std::min(r.bmin.z, pp.z)};
                    r.bmax = {std::max(r.bmax.x, pp.x), std::max(r.bmax.y, pp.y), std::max(r.bmax.z, pp.z)};
                }
            }
            // correlation against the previous frame's patches (scene.cpp:8007-8099)
            std::vector<Patch> local;
            if (M.history != 0xffffffffu && M.history < nh) {
                const fyzz_dy_patch& h = history[M.history];
                Patch old;
                old.n0 = v3(h.body0_normal);
                old.n1 = v3(h.body1_normal);
                for (int k = 0; k < 2; ++k) {
                    old.a0[k] = v3(h.body0_anchors[k]);
                    old.a1[k] = v3(h.body1_anchors[k]);
                }
                old.sf = T(h.static_friction);
                old.df = T(h.dynamic_friction);
                old.rest = T(h.restitution);
                old.damping = T(h.damping);
                old.anchors = h.anchor_count;
                old.broken = h.flags & 1u;
                old.disable_strong = (h.flags >> 1) & 1u;
                if (history_valid(old, a0, a1)) {
                    Patch c = old;
                    c.normal = orc::rotate(mass_orientation(a0.q), old.n0);
                    c.contacts = 0;
                    const T big = std::numeric_limits<T>::max();
                    c.bmin = {big, big, big};
                    c.bmax = {-big, -big, -big};
                    local.push_back(c);
                }
            }
            std::size_t heads[32];
            for (auto& hh : heads) hh = ~std::size_t(0);
            std::uint32_t patch_of_run[4];
            for (std::uint32_t ri = 0; ri < nr; ++ri) {
                Run& cp = runs[ri];
                std::size_t lp = 0;
                for (; lp < local.size(); ++lp) {
                    const Patch& p = local[lp];
                    if (dot(p.normal, cp.normal) >= same_normal && p.rest == restitution && p.damping == damping &&
                        p.sf == static_friction && p.df == dynamic_friction)
                        break;
                }
                if (lp == local.size()) {
                    Patch p{};
                    p.normal = cp.normal;
                    p.sf = static_friction;
                    p.df = dynamic_friction;
                    p.rest = restitution;
                    p.damping = damping;
                    p.n0 = local_normal(a0, cp.normal);
                    p.n1 = local_normal(a1, cp.normal);
                    local.push_back(p);
                }
                Patch& p = local[lp];
                if (p.contacts == 0) {
                    p.bmin = cp.bmin;
                    p.bmax = cp.bmax;
                } else {
                    p.bmin = {std::min(p.bmin.x, cp.bmin.x), std::min(p.bmin.y, cp.bmin.y),
                              std::min(p.bmin.z, cp.bmin.z)};
                    p.bmax = {std::max(p.bmax.x, cp.bmax.x), std::max(p.bmax.y, cp.bmax.y),
                              std::max(p.bmax.z, cp.bmax.z)};
                }
                p.contacts += cp.count;
                cp.next = heads[lp];
                heads[lp] = ri;
                patch_of_run[ri] = static_cast<std::uint32_t>(lp);
            }
            // compaction: keep patches with contacts
            std::vector<Patch> kept;
            std::vector<std::size_t> kept_head;
            std::size_t remap[32];
            for (std::size_t j = 0; j < local.size(); ++j) {
                if (local[j].contacts == 0) continue;
                remap[j] = kept.size();
                kept_head.push_back(heads[j]);
                kept.push_back(local[j]);
            }
            if (kept.size() != 1) return -4;  // PCM managers publish one normal: one patch
            Patch& P = kept[0];
            const V header = v3(M.contacts[runs[kept_head[0]].start].normal);
            // contact rows in the correlation-list order (scene.cpp:8204-8225, 8237-8330)
            std::uint32_t order[4];
            std::uint32_t no = 0;
            for (std::size_t cp = kept_head[0]; cp != ~std::size_t(0); cp = runs[cp].next)
                for (std::uint32_t c = runs[cp].start; c < runs[cp].start + runs[cp].count; ++c) order[no++] = c;
            (void)remap;
            (void)patch_of_run;
            crow[m].resize(M.count);
            for (std::uint32_t j = 0; j < M.count; ++j) {
                const fyzz_dy_contact& C = M.contacts[j];
                const V n = header;
                const V cpos = v3(C.position);
                const V ra_x_n = orc::cross(sub(cpos, a.pos), n);
                const V rb_x_n = orc::cross(sub(cpos, b.pos), n);
                const T lin_vrel = dot(n, sub(a.v, b.v));
                const V ra = transform(a, ra_x_n);
                const V rb = transform(b, rb_x_n);
                const T ang_vrel = dot(ra_x_n, a.w) - dot(rb_x_n, b.w);
                const T vrel = lin_vrel + ang_vrel;
                const T response = (a.inv_mass + dot(ra, ra)) + (b.inv_mass + dot(rb, rb));
                const T vm = response > T(0) ? T(1) / response : T(0);
                const T pen = T(C.separation) - rest_distance;
                const T pen_inv_dt = pen * inv_dt;
                const bool separated = pen >= T(0);
                const bool closing = -vrel > pen_inv_dt;
                const bool bouncing = restitution > T(0) && bounce_threshold > vrel && closing;
                const T target_velocity = bouncing ? -vrel * restitution : T(0);
                ContactRow r{};
                r.n = n;
                r.ra = ra;
                r.rb = rb;
                r.max_impulse = T(std::numeric_limits<float>::max());
                const T target = target_velocity - vrel;
                const T pen_scaled = separated ? pen_inv_dt : pen * (inv_dt * rigid_contact_bias);
                const T scaled_bias = bouncing ? T(0) : vm * std::max(max_pen_bias, pen_scaled);
                r.vm = vm;
                r.biased = target * vm - scaled_bias;
                r.unbiased = target * vm - (bouncing ? T(0) : std::max(scaled_bias, T(0)));
                r.applied = T(0);
                r.raxn = ra_x_n;
                r.rbxn = rb_x_n;
                r.kv = bouncing ? -(T(1) + restitution) : T(-1);
                r.cb = scaled_bias;
                r.cu = bouncing ? T(0) : std::max(scaled_bias, T(0));
                if (!std::isfinite(response) || !std::isfinite(r.biased) || !std::isfinite(r.unbiased)) return -3;
                crow[m][j] = r;
            }
            // anchor growth (scene.cpp:8359-8433)
            {
                const V extent = sub(P.bmax, P.bmin);
                const T diag2 = len2(extent);
                bool keep_pair = false;
                if (P.anchors == 2) {
                    const V d = sub(P.a0[0], P.a0[1]);
                    keep_pair = len2(d) * T(4) >= diag2;
                    if (!keep_pair) P.anchors = 0;
                }
                V anchors[2]{};
                bool changed[2] = {false, false};
                std::uint32_t count = P.anchors;
                if (!keep_pair) {
                    if (count == 1) anchors[0] = world_point(a0, P.a0[0]);
                    T pd2 = T(0);
                    for (std::uint32_t q = 0; q < no; ++q) {
                        const std::uint32_t j = order[q];
                        if (T(M.contacts[j].separation) >= friction_offset_threshold + rest_distance) continue;
                        const V wp = v3(M.contacts[j].position);
                        if (count == 0) {
                            anchors[0] = wp;
                            changed[0] = true;
                            ++count;
                        } else if (count == 1) {
                            const T d = len2(sub(wp, anchors[0]));
                            if (d > T(1.0e-8f)) {
                                anchors[1] = wp;
                                changed[1] = true;
                                pd2 = d;
                                ++count;
                            }
                        } else {
                            const T d0 = len2(sub(wp, anchors[0]));
                            const T d1 = len2(sub(wp, anchors[1]));
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
                    for (std::uint32_t k = 0; k < count; ++k) {
                        if (!changed[k]) continue;
                        P.a0[k] = local_point(a0, anchors[k]);
                        P.a1[k] = local_point(a1, anchors[k]);
                    }
                    if (count == 0) {
                        P.a0[0] = V{0, 0, 0};
                        P.a1[0] = V{0, 0, 0};
                    }
                }
                P.anchors = count;
                P.row_offset = frows.size();
                if (count) {
                    const V n = header;
                    const V rel = sub(a.v, b.v);
                    V t0 = sub(rel, V{n.x * dot(rel, n), n.y * dot(rel, n), n.z * dot(rel, n)});
                    if (len2(t0) <= T(0.0001f))
                        t0 = std::abs(n.x) < T(0.70710678f) ? V{T(0), -n.z, n.y} : V{-n.y, n.x, T(0)};
                    const T il = T(1) / std::sqrt(len2(t0));
                    t0 = {t0.x * il, t0.y * il, t0.z * il};
                    const V t1 = orc::cross(n, t0);
                    for (std::uint32_t k = 0; k < count; ++k) {
                        const V ra = orc::rotate(mass_orientation(a0.q), P.a0[k]);
                        const V rb = orc::rotate(mass_orientation(a1.q), P.a1[k]);
                        const V c0 = mass_position(a0.pos, a0.q), c1 = mass_position(a1.pos, a1.q);
                        const V err = sub(V{ra.x + c0.x, ra.y + c0.y, ra.z + c0.z},
                                          V{rb.x + c1.x, rb.y + c1.y, rb.z + c1.z});
                        for (const V tangent : {t0, t1}) {
                            const V ax0 = orc::cross(ra, tangent);
                            const V ax1 = orc::cross(rb, tangent);
                            const V j0 = transform(a, ax0);
                            const V j1 = transform(b, ax1);
                            const T response = (a.inv_mass + dot(j0, j0)) + (b.inv_mass + dot(j1, j1));
                            FrictionRow fr{};
                            fr.t = tangent;
                            fr.ra = j0;
                            fr.rb = j1;
                            fr.vm = response > T(0) ? rigid_contact_bias / response : T(0);
                            fr.bias = dot(tangent, err) * inv_dt;
                            fr.target = -(dot(tangent, rel) + dot(ax0, a.w) - dot(ax1, b.w));
                            fr.applied = T(0);
                            fr.ax0 = ax0;
                            fr.ax1 = ax1;
                            frows.push_back(fr);
                        }
                    }
                }
            }
            patches[m] = P;
            // the solve walks the patch's correlation-list order; q-th row
            // belongs to manifold contact order[q] (the Gu force slot)
            std::vector<ContactRow> ordered(no);
            for (std::uint32_t q = 0; q < no; ++q) {
                ordered[q] = crow[m][order[q]];
                patch_order_[m][q] = order[q];
            }
            crow[m] = ordered;
        }

        // ---- PGS in partition order (scene.cpp:8853-9148)
        std::size_t nparts = 0;
        for (std::uint32_t m = 0; m < nm; ++m) nparts = std::max(nparts, part[m] + 1);
        std::vector<std::uint32_t> sorted;
        sorted.reserve(nm);
        {
            std::vector<std::size_t> cnt(nparts + 1, 0);
            for (std::uint32_t m = 0; m < nm; ++m) ++cnt[part[m] + 1];
            for (std::size_t p = 1; p <= nparts; ++p) cnt[p] += cnt[p - 1];
            sorted.resize(nm);
            for (std::uint32_t m = 0; m < nm; ++m) sorted[cnt[part[m]]++] = m;
        }
        unsigned pos_iters = 0, vel_iters = 0;
        for (std::uint32_t i = 0; i < nb; ++i)
            if (has_row[i]) {
                pos_iters = std::max(pos_iters, bodies[i].iterations & 0xffu);
                vel_iters = std::max(vel_iters, (bodies[i].iterations >> 8) & 0xffu);
            }
        vel_iters = std::max(vel_iters, 1u);
        if (capture_) {
            Frozen& F = *capture_;
            F.dt = dt;
            F.pos_iters = pos_iters;
            F.vel_iters = vel_iters;
            F.bodies.assign(nb, FrozenBody{});
            for (std::uint32_t i = 0; i < nb; ++i) {
                FrozenBody& fb = F.bodies[i];
                const fyzz_dy_body& b = bodies[i];
                fb.awake = !(b.flags & 1u);
                fb.has_row = has_row[i] != 0;
                fb.inv_mass = has_row[i] ? rows[i].inv_mass : T(0);
                for (int k = 0; k < 3; ++k) fb.c[k] = has_row[i] ? rows[i].sii[k] : V{0, 0, 0};
                fb.lm = std::max(T(1) - T(b.linear_damping) * dt, T(0));
                fb.am = std::max(T(1) - T(b.angular_damping) * dt, T(0));
                fb.gravity = (b.flags & 4u) != 0;
            }
            F.mgrs.clear();
            for (std::uint32_t m : sorted) {
                FrozenMgr fm;
                fm.b0 = mgr[m].body0;
                fm.b1 = mgr[m].body1;
                fm.c = crow[m];
                const Patch& P = patches[m];
                fm.anchors = P.anchors;
                fm.f.assign(frows.begin() + static_cast<std::ptrdiff_t>(P.row_offset),
                            frows.begin() + static_cast<std::ptrdiff_t>(P.row_offset + 2 * P.anchors));
                for (int k = 0; k < 2; ++k) {
                    const int32_t id = k ? mgr[m].body1 : mgr[m].body0;
                    fm.static_sf[k] = id < 0 ? T(statics[-1 - id].static_friction) : T(0);
                    fm.static_df[k] = id < 0 ? T(statics[-1 - id].dynamic_friction) : T(0);
                }
                F.mgrs.push_back(std::move(fm));
            }
        }
        // the task-local copy of the static-world body (scene.cpp:8866)
        Row wcopy = world;
        for (unsigned iter = 0; iter < pos_iters + vel_iters; ++iter) {
            for (std::uint32_t m : sorted) {
                Row& aa = mgr[m].body0 >= 0 ? rows[mgr[m].body0] : wcopy;
                Row& bb = mgr[m].body1 >= 0 ? rows[mgr[m].body1] : wcopy;
                Patch& P = patches[m];
                T normal_impulse = T(0);
                for (ContactRow& r : crow[m]) {
                    const V n = r.n;
                    const V va = aa.lin_delta, vb = bb.lin_delta;
                    const T nv = dot(n, va) - dot(n, vb) + dot(aa.ang_state, r.ra) - dot(bb.ang_state, r.rb);
                    const T err = iter < pos_iters ? r.biased : r.unbiased;
                    const T dc = std::max(-nv * r.vm + err, -r.applied);
                    const T next = std::min(r.max_impulse, T(1) * r.applied + dc);
                    const T delta = next - r.applied;
                    const T da = delta * aa.inv_mass;
                    const T db = delta * bb.inv_mass;
                    aa.lin_delta = {va.x + n.x * da, va.y + n.y * da, va.z + n.z * da};
                    bb.lin_delta = {vb.x - n.x * db, vb.y - n.y * db, vb.z - n.z * db};
                    aa.ang_state = {aa.ang_state.x + r.ra.x * delta, aa.ang_state.y + r.ra.y * delta,
                                    aa.ang_state.z + r.ra.z * delta};
                    bb.ang_state = {bb.ang_state.x - r.rb.x * delta, bb.ang_state.y - r.rb.y * delta,
                                    bb.ang_state.z - r.rb.z * delta};
                    if (!finite(aa.lin_delta) || !finite(bb.lin_delta) || !finite(aa.ang_state) ||
                        !finite(bb.ang_state))
                        return -3;
                    r.applied = next;
                    normal_impulse += next;
                }
                if ((iter >= pos_iters || pos_iters - iter <= 3) && P.anchors) {
                    const T coef = P.anchors == 2 ? T(0.5f) : T(1.0f);
                    const T max_static = P.sf * coef * normal_impulse;
                    const T max_dynamic = P.df * coef * normal_impulse;
                    bool broken = false;
                    for (std::size_t k = P.row_offset; k < P.row_offset + P.anchors * 2; ++k) {
                        FrictionRow& fr = frows[k];
                        const V n = fr.t;
                        const V va = aa.lin_delta, vb = bb.lin_delta;
                        const T td = dot(n, va) - dot(n, vb) + dot(aa.ang_state, fr.ra) - dot(bb.ang_state, fr.rb);
                        const T bias = iter < pos_iters ? fr.bias : T(0);
                        const T cand = fr.applied + (fr.target - bias - td) * fr.vm;
                        const bool clamped = std::abs(cand) > max_static;
                        const T next = clamped ? std::min(max_dynamic, std::max(-max_dynamic, cand)) : cand;
                        const T delta = next - fr.applied;
                        const T da = delta * aa.inv_mass;
                        const T db = delta * bb.inv_mass;
                        aa.lin_delta = {va.x + n.x * da, va.y + n.y * da, va.z + n.z * da};
                        bb.lin_delta = {vb.x - n.x * db, vb.y - n.y * db, vb.z - n.z * db};
                        aa.ang_state = {aa.ang_state.x + fr.ra.x * delta, aa.ang_state.y + fr.ra.y * delta,
                                        aa.ang_state.z + fr.ra.z * delta};
                        bb.ang_state = {bb.ang_state.x - fr.rb.x * delta, bb.ang_state.y - fr.rb.y * delta,
                                        bb.ang_state.z - fr.rb.z * delta};
                        if (!finite(aa.lin_delta) || !finite(bb.lin_delta) || !finite(aa.ang_state) ||
                            !finite(bb.ang_state))
                            return -3;
                        fr.applied = next;
                        broken |= clamped;
                    }
                    if (iter + 1 == pos_iters + vel_iters) P.broken = broken;
                }
            }
            if (iter + 1 == pos_iters)
                for (std::uint32_t i = 0; i < nb; ++i)
                    if (has_row[i]) {
                        rows[i].motion_lin = rows[i].lin_delta;
                        rows[i].motion_ang = rows[i].ang_state;
                    }
        }

        // ---- outputs: patches, impulses, partitions
        for (std::uint32_t m = 0; m < nm; ++m) {
            const Patch& P = patches[m];
            fyzz_dy_patch& o = patches_out[m];
            std::memset(&o, 0, sizeof o);
            auto put = [](float* d, V s) {
                d[0] = float(s.x);
                d[1] = float(s.y);
                d[2] = float(s.z);
            };
            put(o.body0_normal, P.n0);
            put(o.body1_normal, P.n1);
            for (int k = 0; k < 2; ++k) {
                put(o.body0_anchors[k], P.a0[k]);
                put(o.body1_anchors[k], P.a1[k]);
            }
            o.static_friction = float(P.sf);
            o.dynamic_friction = float(P.df);
            o.restitution = float(P.rest);
            o.damping = float(P.damping);
            o.anchor_count = P.anchors;
            o.flags = (P.broken ? 1u : 0u) | (P.disable_strong ? 2u : 0u);
            for (std::uint32_t q = 0; q < 4; ++q) impulses_out[4 * m + q] = 0.0f;
            for (std::uint32_t q = 0; q < mgr[m].count; ++q)
                impulses_out[4 * m + patch_order_[m][q]] = float(crow[m][q].applied);
            if (parts_out) parts_out[m] = static_cast<std::uint32_t>(part[m]);
            // friction write-back (Scene::write_back_contacts, scene.cpp:
            // 9339-9388): anchors on body0's mass pose at the step's start,
            // impulses as double sums of tangent x applied force
            if (anchors_out_) {
                fyzz_friction_anchors& A = anchors_out_[m];
                std::memset(&A, 0, sizeof A);
                A.anchor_count = P.anchors;
                A.broken = P.broken ? 1u : 0u;
                const BodyView a0 = view(mgr[m].body0);
                const Q mq = mass_orientation(a0.q);
                const V c0 = mass_position(a0.pos, a0.q);
                for (std::uint32_t k = 0; k < P.anchors && k < 2; ++k) {
                    const V r = orc::rotate(mq, P.a0[k]);
                    A.position[k][0] = float(c0.x + r.x);
                    A.position[k][1] = float(c0.y + r.y);
                    A.position[k][2] = float(c0.z + r.z);
                    const FrictionRow& x = frows[P.row_offset + 2 * k];
                    const FrictionRow& y = frows[P.row_offset + 2 * k + 1];
                    A.impulse[k][0] = double(x.t.x) * double(x.applied) + double(y.t.x) * double(y.applied);
                    A.impulse[k][1] = double(x.t.y) * double(x.applied) + double(y.t.y) * double(y.applied);
                    A.impulse[k][2] = double(x.t.z) * double(x.applied) + double(y.t.z) * double(y.applied);
                }
            }
        }

        // ---- publish + integrate + sleep check (scene.cpp:9509-9549, 10508-10707)
        for (std::uint32_t i = 0; i < nb; ++i) {
            fyzz_dy_body& B = bodies[i];
            if (B.flags & 1u) continue;
            V vel = v[i], ang = w[i];
            V motion_lin = vel, motion_ang = ang;
            if (has_row[i]) {
                const Row& r = rows[i];
                vel = {r.v.x + r.lin_delta.x, r.v.y + r.lin_delta.y, r.v.z + r.lin_delta.z};
                const V* c = r.sii;
                ang = {r.w.x + c[0].x * r.ang_state.x + c[1].x * r.ang_state.y + c[2].x * r.ang_state.z,
                       r.w.y + c[0].y * r.ang_state.x + c[1].y * r.ang_state.y + c[2].y * r.ang_state.z,
                       r.w.z + c[0].z * r.ang_state.x + c[1].z * r.ang_state.y + c[2].z * r.ang_state.z};
                motion_lin = {r.v.x + r.motion_lin.x, r.v.y + r.motion_lin.y, r.v.z + r.motion_lin.z};
                motion_ang = {
                    r.w.x + c[0].x * r.motion_ang.x + c[1].x * r.motion_ang.y + c[2].x * r.motion_ang.z,
                    r.w.y + c[0].y * r.motion_ang.x + c[1].y * r.motion_ang.y + c[2].y * r.motion_ang.z,
                    r.w.z + c[0].z * r.motion_ang.x + c[1].z * r.motion_ang.y + c[2].z * r.motion_ang.z};
                if (!finite(vel) || !finite(ang)) return -3;
            } else {
                motion_lin = vel;
                motion_ang = ang;
            }
            const V p0 = v3(B.position);
            const Q q0 = q4(B.orientation);
            const V center = mass_position(p0, q0);
            V p{center.x + motion_lin.x * dt, center.y + motion_lin.y * dt, center.z + motion_lin.z * dt};
            Q q;
            if (!integrate_quaternion(mass_orientation(q0), motion_ang, dt, q)) return -3;
            q = orc::multiply(q, Q{-T(0), -T(0), -T(0), T(1)});
            const V off = orc::rotate(q, V{0, 0, 0});
            p = {p.x - off.x, p.y - off.y, p.z - off.z};
            if (!finite(p) || !finite(q)) return -3;
            // sleep check (scene.cpp:10640-10705)
            const T wake_reset = T(20.0f * 0.02f);
            const T energy_threshold = T(5.0e-5f * 10.0f * 10.0f);
            T wc = T(B.wake_counter);
            V lin_acc = v3(B.sleep_linear_acc), ang_acc = v3(B.sleep_angular_acc);
            if (wc < wake_reset * T(0.5f) || wc < dt) {
                const V la = orc::rotate_inverse(q, motion_ang);
                const V l2{lin_acc.x + motion_lin.x, lin_acc.y + motion_lin.y, lin_acc.z + motion_lin.z};
                const V a2{ang_acc.x + la.x, ang_acc.y + la.y, ang_acc.z + la.z};
                const V ii = v3(B.inverse_inertia);
                const V inertia{ii.x > T(0) ? T(1) / ii.x : T(1), ii.y > T(0) ? T(1) / ii.y : T(1),
                                ii.z > T(0) ? T(1) / ii.z : T(1)};
                const T ang_e = (a2.x * a2.x * inertia.x + a2.y * a2.y * inertia.y + a2.z * a2.z * inertia.z) *
                                T(B.inverse_mass);
                const T lin_e = l2.x * l2.x + l2.y * l2.y + l2.z * l2.z;
                const T ne = T(0.5f) * (ang_e + lin_e);
                const T cluster = T(1) + T(float(B.counted_interactions));
                const T th = cluster * energy_threshold;
                if (ne >= th) {
                    lin_acc = {0, 0, 0};
                    ang_acc = {0, 0, 0};
                    const T factor = th == T(0) ? T(2) : std::min(ne / th, T(2));
                    wc = factor * T(0.5f) * wake_reset + dt * (cluster - T(1));
                } else {
                    lin_acc = l2;
                    ang_acc = a2;
                    wc = std::max(wc - dt, T(0));
                }
            } else {
                wc = std::max(wc - dt, T(0));
            }
            const bool ready = wc == T(0);
            if (ready) {
                lin_acc = {0, 0, 0};
                ang_acc = {0, 0, 0};
            }
            if (state_out) {
                double* s = state_out + 13ull * i;
                s[0] = double(p.x); s[1] = double(p.y); s[2] = double(p.z);
                s[3] = double(vel.x); s[4] = double(vel.y); s[5] = double(vel.z);
                s[6] = double(q.x); s[7] = double(q.y); s[8] = double(q.z); s[9] = double(q.w);
                s[10] = double(ang.x); s[11] = double(ang.y); s[12] = double(ang.z);
            }
            auto put = [](float* d, V s) {
                d[0] = float(s.x);
                d[1] = float(s.y);
                d[2] = float(s.z);
            };
            put(B.position, p);
            B.orientation[0] = float(q.x);
            B.orientation[1] = float(q.y);
            B.orientation[2] = float(q.z);
            B.orientation[3] = float(q.w);
            put(B.linear_velocity, vel);
            put(B.angular_velocity, ang);
            B.wake_counter = float(wc);
            put(B.sleep_linear_acc, lin_acc);
            put(B.sleep_angular_acc, ang_acc);
            B.flags = (B.flags & ~2u) | (ready ? 2u : 0u);
        }
        return 0;
    }
    static inline thread_local std::vector<std::array<std::uint32_t, 4>> patch_order_;
    // optional friction write-back output (one record per manager)
    static inline thread_local fyzz_friction_anchors* anchors_out_ = nullptr;

    // ---- frozen-row rollout (the differentiable form's reference)
    struct FrozenBody {
        bool awake = true, has_row = false, gravity = true;
        T inv_mass = 0, lm = 1, am = 1;
        V c[3]{};
    };
    struct FrozenMgr {
        int32_t b0 = 0, b1 = 0;
        std::vector<ContactRow> c;
        std::vector<FrictionRow> f;
        std::uint32_t anchors = 0;
        T static_sf[2]{}, static_df[2]{};
    };
    struct Frozen {
        T dt = 0;
        unsigned pos_iters = 0, vel_iters = 1;
        std::vector<FrozenBody> bodies;
        std::vector<FrozenMgr> mgrs;  // solve order
    };
    static inline thread_local Frozen* capture_ = nullptr;
};

// One step of the frozen-row function: rows (geometry, effective masses,
// biases, friction frames, branches of the bounce test) fixed at the
// nominal step; velocities, the PGS, and positions follow the inputs.
// Records each row update's branch for the reverse pass.
struct FrozenTape {
    std::vector<std::vector<unsigned char>> c, f;  // per iteration, per solve-ordered row
    std::vector<std::vector<double>> ni;            // per iteration, per manager
    std::vector<double> vpre, wpre;                 // [nb][3]
};
using D = Dy<double>;

void frozen_friction(const D::Frozen& F, const D::FrozenMgr& fm, const double* mu, double& sf, double& df) {
    double sfb[2], dfb[2];
    for (int k = 0; k < 2; ++k) {
        const int32_t id = k ? fm.b1 : fm.b0;
        sfb[k] = id >= 0 ? mu[id] : fm.static_sf[k];
        dfb[k] = id >= 0 ? mu[id] : fm.static_df[k];
    }
    df = std::max(0.0, 0.5 * (dfb[0] + dfb[1]));
    sf = std::max(df, 0.5 * (sfb[0] + sfb[1]));
    (void)F;
}

void frozen_step(const D::Frozen& F, std::vector<double>& v, std::vector<double>& w, std::vector<double>& p,
                 const double* mu, FrozenTape* tape) {
    using V = D::V;
    const std::size_t nb = F.bodies.size();
    const double dt = F.dt;
    std::vector<double> vpre(3 * nb), wpre(3 * nb), lin(3 * nb, 0.0), ang(3 * nb, 0.0), mlin(3 * nb, 0.0);
    for (std::size_t i = 0; i < nb; ++i) {
        const auto& fb = F.bodies[i];
        for (int k = 0; k < 3; ++k) {
            vpre[3 * i + k] = v[3 * i + k];
            wpre[3 * i + k] = w[3 * i + k];
        }
        if (!fb.awake) continue;
        if (fb.gravity) vpre[3 * i + 1] += -9.81 * dt;
        for (int k = 0; k < 3; ++k) {
            vpre[3 * i + k] *= fb.lm;
            wpre[3 * i + k] *= fb.am;
        }
    }
    auto get = [&](const std::vector<double>& a, int32_t id) -> V {
        if (id < 0) return {0, 0, 0};
        return {a[3 * id], a[3 * id + 1], a[3 * id + 2]};
    };
    auto add = [&](std::vector<double>& a, int32_t id, V d, double s) {
        if (id < 0) return;
        a[3 * id] += d.x * s;
        a[3 * id + 1] += d.y * s;
        a[3 * id + 2] += d.z * s;
    };
    auto dot = [](V a, V b) { return a.x * b.x + a.y * b.y + a.z * b.z; };
    // velocity-dependent row terms
    std::vector<std::vector<double>> biased(F.mgrs.size()), unbiased(F.mgrs.size()), target(F.mgrs.size());
    std::vector<std::vector<double>> capp(F.mgrs.size()), fapp(F.mgrs.size());
    for (std::size_t m = 0; m < F.mgrs.size(); ++m) {
        const auto& fm = F.mgrs[m];
        const V va = get(vpre, fm.b0), vb = get(vpre, fm.b1), wa = get(wpre, fm.b0), wb = get(wpre, fm.b1);
        for (const auto& r : fm.c) {
            const double vrel = dot(r.n, {va.x - vb.x, va.y - vb.y, va.z - vb.z}) + dot(r.raxn, wa) - dot(r.rbxn, wb);
            biased[m].push_back(r.kv * vrel * r.vm - r.cb);
            unbiased[m].push_back(r.kv * vrel * r.vm - r.cu);
        }
        for (const auto& fr : fm.f)
            target[m].push_back(-(dot(fr.t, {va.x - vb.x, va.y - vb.y, va.z - vb.z}) + dot(fr.ax0, wa) - dot(fr.ax1, wb)));
        capp[m].assign(fm.c.size(), 0.0);
        fapp[m].assign(fm.f.size(), 0.0);
    }
    const unsigned total = F.pos_iters + F.vel_iters;
    if (tape) {
        tape->c.assign(total, {});
        tape->f.assign(total, {});
        tape->ni.assign(total, std::vector<double>(F.mgrs.size(), 0.0));
        tape->vpre = vpre;
        tape->wpre = wpre;
    }
    for (unsigned it = 0; it < total; ++it) {
        for (std::size_t m = 0; m < F.mgrs.size(); ++m) {
            const auto& fm = F.mgrs[m];
            const double ima = fm.b0 >= 0 ? F.bodies[fm.b0].inv_mass : 0.0;
            const double imb = fm.b1 >= 0 ? F.bodies[fm.b1].inv_mass : 0.0;
            double ni = 0.0;
            for (std::size_t q = 0; q < fm.c.size(); ++q) {
                const auto& r = fm.c[q];
                const V va = get(lin, fm.b0), vb = get(lin, fm.b1), wa = get(ang, fm.b0), wb = get(ang, fm.b1);
                const double nv = dot(r.n, va) - dot(r.n, vb) + dot(wa, r.ra) - dot(wb, r.rb);
                const double err = it < F.pos_iters ? biased[m][q] : unbiased[m][q];
                const double x = -nv * r.vm + err;
                const bool free = !(x < -capp[m][q]);
                const double next = free ? capp[m][q] + x : 0.0;
                const double delta = next - capp[m][q];
                add(lin, fm.b0, r.n, delta * ima);
                add(lin, fm.b1, r.n, -delta * imb);
                add(ang, fm.b0, r.ra, delta);
                add(ang, fm.b1, r.rb, -delta);
                capp[m][q] = next;
                ni += next;
                if (tape) tape->c[it].push_back(free ? 1 : 0);
            }
            if (tape) tape->ni[it][m] = ni;
            if ((it >= F.pos_iters || F.pos_iters - it <= 3) && fm.anchors) {
                double sf, df;
                frozen_friction(F, fm, mu, sf, df);
                const double coef = fm.anchors == 2 ? 0.5 : 1.0;
                const double max_static = sf * coef * ni, max_dynamic = df * coef * ni;
                for (std::size_t k = 0; k < fm.f.size(); ++k) {
                    const auto& fr = fm.f[k];
                    const V va = get(lin, fm.b0), vb = get(lin, fm.b1), wa = get(ang, fm.b0), wb = get(ang, fm.b1);
                    const double td = dot(fr.t, va) - dot(fr.t, vb) + dot(wa, fr.ra) - dot(wb, fr.rb);
                    const double bias = it < F.pos_iters ? fr.bias : 0.0;
                    const double cand = fapp[m][k] + (target[m][k] - bias - td) * fr.vm;
                    const bool clamped = std::abs(cand) > max_static;
                    unsigned char br = 0;
                    double next = cand;
                    if (clamped) {
                        br = 1;
                        if (cand > max_dynamic) next = max_dynamic, br = 2;
                        else if (cand < -max_dynamic) next = -max_dynamic, br = 3;
                    }
                    const double delta = next - fapp[m][k];
                    add(lin, fm.b0, fr.t, delta * ima);
                    add(lin, fm.b1, fr.t, -delta * imb);
                    add(ang, fm.b0, fr.ra, delta);
                    add(ang, fm.b1, fr.rb, -delta);
                    fapp[m][k] = next;
                    if (tape) tape->f[it].push_back(br);
                }
            }
        }
        if (it + 1 == F.pos_iters) mlin = lin;
    }
    for (std::size_t i = 0; i < nb; ++i) {
        const auto& fb = F.bodies[i];
        if (!fb.awake) continue;
        for (int k = 0; k < 3; ++k) {
            const double motion = vpre[3 * i + k] + (fb.has_row ? mlin[3 * i + k] : 0.0);
            p[3 * i + k] += motion * dt;
            v[3 * i + k] = vpre[3 * i + k] + lin[3 * i + k];
            double s = 0.0;
            for (int j = 0; j < 3; ++j) s += (&fb.c[j].x)[k] * ang[3 * i + j];
            w[3 * i + k] = wpre[3 * i + k] + s;
        }
    }
}

// Reverse of frozen_step: adjoints of (v, w, p) after the step in, of
// (v, w, p) before it out; per-body dynamic friction adjoints accumulate.
void frozen_step_adjoint(const D::Frozen& F, const FrozenTape& tape, const double* mu, std::vector<double>& av,
                         std::vector<double>& aw, const std::vector<double>& ap, std::vector<double>& amu) {
    using V = D::V;
    const std::size_t nb = F.bodies.size();
    const double dt = F.dt;
    auto dot = [](V a, V b) { return a.x * b.x + a.y * b.y + a.z * b.z; };
    std::vector<double> alin(3 * nb, 0.0), aang(3 * nb, 0.0), avpre(3 * nb, 0.0), awpre(3 * nb, 0.0);
    for (std::size_t i = 0; i < nb; ++i) {
        const auto& fb = F.bodies[i];
        if (!fb.awake) continue;
        for (int k = 0; k < 3; ++k) {
            alin[3 * i + k] = av[3 * i + k];
            avpre[3 * i + k] = av[3 * i + k] + dt * ap[3 * i + k];
            awpre[3 * i + k] = aw[3 * i + k];
        }
        for (int j = 0; j < 3; ++j)
            aang[3 * i + j] = fb.c[j].x * aw[3 * i] + fb.c[j].y * aw[3 * i + 1] + fb.c[j].z * aw[3 * i + 2];
    }
    auto get = [&](const std::vector<double>& a, int32_t id) -> V {
        if (id < 0) return {0, 0, 0};
        return {a[3 * id], a[3 * id + 1], a[3 * id + 2]};
    };
    auto put = [&](std::vector<double>& a, int32_t id, V x) {
        if (id < 0) return;
        a[3 * id] = x.x;
        a[3 * id + 1] = x.y;
        a[3 * id + 2] = x.z;
    };
    auto axpy = [](V a, V d, double s) { return V{a.x + d.x * s, a.y + d.y * s, a.z + d.z * s}; };
    std::vector<std::vector<double>> acapp(F.mgrs.size()), afapp(F.mgrs.size()), aerr(F.mgrs.size()),
        atgt(F.mgrs.size());
    std::vector<double> adf(F.mgrs.size(), 0.0);
    for (std::size_t m = 0; m < F.mgrs.size(); ++m) {
        acapp[m].assign(F.mgrs[m].c.size(), 0.0);
        aerr[m].assign(F.mgrs[m].c.size(), 0.0);
        afapp[m].assign(F.mgrs[m].f.size(), 0.0);
        atgt[m].assign(F.mgrs[m].f.size(), 0.0);
    }
    // row offsets into the tape, per iteration
    const unsigned total = F.pos_iters + F.vel_iters;
    for (unsigned it = total; it-- > 0;) {
        if (it + 1 == F.pos_iters)
            for (std::size_t i = 0; i < nb; ++i)
                if (F.bodies[i].awake && F.bodies[i].has_row)
                    for (int k = 0; k < 3; ++k) alin[3 * i + k] += dt * ap[3 * i + k];
        std::vector<std::size_t> coff(F.mgrs.size() + 1, 0), foff(F.mgrs.size() + 1, 0);
        for (std::size_t m = 0; m < F.mgrs.size(); ++m) {
            coff[m + 1] = coff[m] + F.mgrs[m].c.size();
            const bool fr_on = (it >= F.pos_iters || F.pos_iters - it <= 3) && F.mgrs[m].anchors;
            foff[m + 1] = foff[m] + (fr_on ? F.mgrs[m].f.size() : 0);
        }
        for (std::size_t m = F.mgrs.size(); m-- > 0;) {
            const auto& fm = F.mgrs[m];
            const double ima = fm.b0 >= 0 ? F.bodies[fm.b0].inv_mass : 0.0;
            const double imb = fm.b1 >= 0 ? F.bodies[fm.b1].inv_mass : 0.0;
            V ava = get(alin, fm.b0), avb = get(alin, fm.b1), awa = get(aang, fm.b0), awb = get(aang, fm.b1);
            double a_ni = 0.0;
            if (foff[m + 1] > foff[m]) {
                double sf, df;
                frozen_friction(F, fm, mu, sf, df);
                const double coef = fm.anchors == 2 ? 0.5 : 1.0;
                const double ni = tape.ni[it][m];
                double a_maxd = 0.0;
                for (std::size_t k = fm.f.size(); k-- > 0;) {
                    const auto& fr = fm.f[k];
                    const double a_delta = dot(fr.t, ava) * ima - dot(fr.t, avb) * imb + dot(fr.ra, awa) - dot(fr.rb, awb);
                    const double a_next = afapp[m][k] + a_delta;
                    double a_applied = -a_delta, a_cand = 0.0;
                    const unsigned char br = tape.f[it][foff[m] + k];
                    if (br <= 1) a_cand = a_next;
                    else if (br == 2) a_maxd += a_next;
                    else a_maxd -= a_next;
                    a_applied += a_cand;
                    atgt[m][k] += fr.vm * a_cand;
                    const double a_td = -fr.vm * a_cand;
                    ava = axpy(ava, fr.t, a_td);
                    avb = axpy(avb, fr.t, -a_td);
                    awa = axpy(awa, fr.ra, a_td);
                    awb = axpy(awb, fr.rb, -a_td);
                    afapp[m][k] = a_applied;
                }
                adf[m] += coef * ni * a_maxd;
                a_ni = df * coef * a_maxd;
            }
            for (std::size_t q = fm.c.size(); q-- > 0;) {
                const auto& r = fm.c[q];
                const double a_delta = dot(r.n, ava) * ima - dot(r.n, avb) * imb + dot(r.ra, awa) - dot(r.rb, awb);
                const double a_next = acapp[m][q] + a_delta + a_ni;
                double a_applied = -a_delta;
                if (tape.c[it][coff[m] + q]) {
                    a_applied += a_next;
                    const double a_nv = -r.vm * a_next;
                    aerr[m][q] += a_next;
                    ava = axpy(ava, r.n, a_nv);
                    avb = axpy(avb, r.n, -a_nv);
                    awa = axpy(awa, r.ra, a_nv);
                    awb = axpy(awb, r.rb, -a_nv);
                }
                acapp[m][q] = a_applied;
            }
            put(alin, fm.b0, ava);
            put(alin, fm.b1, avb);
            put(aang, fm.b0, awa);
            put(aang, fm.b1, awb);
        }
    }
    for (std::size_t m = 0; m < F.mgrs.size(); ++m) {
        const auto& fm = F.mgrs[m];
        V ava{0, 0, 0}, avb{0, 0, 0}, awa{0, 0, 0}, awb{0, 0, 0};
        for (std::size_t q = 0; q < fm.c.size(); ++q) {
            const auto& r = fm.c[q];
            const double a_vrel = r.kv * r.vm * aerr[m][q];
            ava = axpy(ava, r.n, a_vrel);
            avb = axpy(avb, r.n, -a_vrel);
            awa = axpy(awa, r.raxn, a_vrel);
            awb = axpy(awb, r.rbxn, -a_vrel);
        }
        for (std::size_t k = 0; k < fm.f.size(); ++k) {
            const auto& fr = fm.f[k];
            ava = axpy(ava, fr.t, -atgt[m][k]);
            avb = axpy(avb, fr.t, atgt[m][k]);
            awa = axpy(awa, fr.ax0, -atgt[m][k]);
            awb = axpy(awb, fr.ax1, atgt[m][k]);
        }
        double sf, df;
        frozen_friction(F, fm, mu, sf, df);
        const double a_df = df > 0.0 ? 0.5 * adf[m] : 0.0;
        for (int k = 0; k < 2; ++k) {
            const int32_t id = k ? fm.b1 : fm.b0;
            if (id < 0 || !F.bodies[id].has_row) continue;
            const V dv = k ? avb : ava, dw = k ? awb : awa;
            avpre[3 * id] += dv.x;
            avpre[3 * id + 1] += dv.y;
            avpre[3 * id + 2] += dv.z;
            awpre[3 * id] += dw.x;
            awpre[3 * id + 1] += dw.y;
            awpre[3 * id + 2] += dw.z;
            amu[id] += a_df;
        }
    }
    for (std::size_t i = 0; i < nb; ++i) {
        const auto& fb = F.bodies[i];
        if (!fb.awake) continue;  // unchanged by the step: adjoints pass through
        for (int k = 0; k < 3; ++k) {
            av[3 * i + k] = fb.lm * avpre[3 * i + k];
            aw[3 * i + k] = fb.am * awpre[3 * i + k];
        }
    }
}

struct DiffRollout {
    std::vector<D::Frozen> steps;
    std::vector<double> p0;
};

// Widens rows frozen along an FP32 nominal trajectory to the FP64 form.
D::V wide(const Dy<float>::V& v) { return {double(v.x), double(v.y), double(v.z)}; }
D::Frozen wide(const Dy<float>::Frozen& f) {
    D::Frozen o;
    o.dt = f.dt;
    o.pos_iters = f.pos_iters;
    o.vel_iters = f.vel_iters;
    o.bodies.resize(f.bodies.size());
    for (std::size_t i = 0; i < f.bodies.size(); ++i) {
        const auto& a = f.bodies[i];
        auto& b = o.bodies[i];
        b.awake = a.awake;
        b.has_row = a.has_row;
        b.gravity = a.gravity;
        b.inv_mass = a.inv_mass;
        b.lm = a.lm;
        b.am = a.am;
        for (int k = 0; k < 3; ++k) b.c[k] = wide(a.c[k]);
    }
    o.mgrs.resize(f.mgrs.size());
    for (std::size_t m = 0; m < f.mgrs.size(); ++m) {
        const auto& a = f.mgrs[m];
        auto& b = o.mgrs[m];
        b.b0 = a.b0;
        b.b1 = a.b1;
        b.anchors = a.anchors;
        for (int k = 0; k < 2; ++k) {
            b.static_sf[k] = a.static_sf[k];
            b.static_df[k] = a.static_df[k];
        }
        for (const auto& r : a.c) {
            D::ContactRow c;
            c.n = wide(r.n);
            c.ra = wide(r.ra);
            c.rb = wide(r.rb);
            c.vm = r.vm;
            c.biased = r.biased;
            c.unbiased = r.unbiased;
            c.applied = r.applied;
            c.max_impulse = r.max_impulse;
            c.raxn = wide(r.raxn);
            c.rbxn = wide(r.rbxn);
            c.kv = r.kv;
            c.cb = r.cb;
            c.cu = r.cu;
            b.c.push_back(c);
        }
        for (const auto& r : a.f) {
            D::FrictionRow c;
            c.t = wide(r.t);
            c.ra = wide(r.ra);
            c.rb = wide(r.rb);
            c.vm = r.vm;
            c.bias = r.bias;
            c.target = r.target;
            c.applied = r.applied;
            c.ax0 = wide(r.ax0);
            c.ax1 = wide(r.ax1);
            b.f.push_back(c);
        }
    }
    return o;
}

}  // namespace

extern "C" {

// Frozen-row rollout of the differentiable form (FP64). create: runs the
// nominal Dy steps from bodies0 with each step's recorded managers and
// history (managers concatenated, nm[s] per step; history likewise, manager
// history indices local to the step) and keeps every step's rows. With
// nominal_fp32 the nominal trajectory is the FP32 restatement's (rows
// widened to FP64), else the FP64 twin's. The Dy step has discontinuous
// decisions on the poses (friction anchors kept or reset), so two nominal
// trajectories of different precision can freeze different rows; the FP32
// nominal freezes the rows an FP32 implementation freezes.
void* orc_diff_create(std::uint32_t nb, const fyzz_dy_body* bodies0, std::uint32_t ns, const fyzz_dy_static* statics,
                      std::uint32_t steps, const std::uint32_t* nm, const fyzz_dy_manager* managers,
                      const std::uint32_t* nh, const fyzz_dy_patch* history, float dt, int nominal_fp32) {
    auto* R = new DiffRollout();
    std::vector<fyzz_dy_body> b(bodies0, bodies0 + nb);
    R->p0.resize(3 * nb);
    for (std::uint32_t i = 0; i < nb; ++i)
        for (int k = 0; k < 3; ++k) R->p0[3 * i + k] = b[i].position[k];
    std::size_t mo = 0, ho = 0;
    R->steps.resize(steps);
    for (std::uint32_t s = 0; s < steps; ++s) {
        std::vector<fyzz_dy_patch> patches(std::max<std::uint32_t>(nm[s], 1));
        std::vector<float> imp(4 * std::max<std::uint32_t>(nm[s], 1));
        std::vector<double> st(13ull * nb);
        int rc;
        if (nominal_fp32) {
            Dy<float>::Frozen f;
            Dy<float>::patch_order_.assign(nm[s], {});
            Dy<float>::capture_ = &f;
            rc = Dy<float>::step(dt, nb, b.data(), ns, statics, nm[s], managers + mo, history + ho, nh[s],
                                 patches.data(), imp.data(), nullptr, nullptr);
            Dy<float>::capture_ = nullptr;
            R->steps[s] = wide(f);
        } else {
            D::patch_order_.assign(nm[s], {});
            D::capture_ = &R->steps[s];
            rc = D::step(double(dt), nb, b.data(), ns, statics, nm[s], managers + mo, history + ho, nh[s],
                         patches.data(), imp.data(), nullptr, st.data());
            D::capture_ = nullptr;
        }
        if (rc != 0) {
            delete R;
            return nullptr;
        }
        mo += nm[s];
        ho += nh[s];
    }
    return R;
}

void orc_diff_destroy(void* h) { delete static_cast<DiffRollout*>(h); }

// Final positions [nb][3] of the frozen-row rollout for initial linear and
// angular velocities v0, w0 [nb][3] and per-body friction mu [nb].
int orc_diff_eval(void* h, const double* v0, const double* w0, const double* mu, double* p_final) {
    auto* R = static_cast<DiffRollout*>(h);
    const std::size_t nb = R->p0.size() / 3;
    std::vector<double> v(v0, v0 + 3 * nb), w(w0, w0 + 3 * nb), p = R->p0;
    for (const auto& F : R->steps) frozen_step(F, v, w, p, mu, nullptr);
    std::copy(p.begin(), p.end(), p_final);
    return 0;
}

// Reverse mode: adjoint of the final positions in, gradients of v0, w0 and
// mu out (and the final positions).
int orc_diff_grad(void* h, const double* v0, const double* w0, const double* mu, const double* adj_p_final,
                  double* g_v0, double* g_w0, double* g_mu, double* p_final) {
    auto* R = static_cast<DiffRollout*>(h);
    const std::size_t nb = R->p0.size() / 3;
    std::vector<double> v(v0, v0 + 3 * nb), w(w0, w0 + 3 * nb), p = R->p0;
    std::vector<FrozenTape> tapes(R->steps.size());
    std::vector<std::vector<double>> vs, ws;
    for (std::size_t s = 0; s < R->steps.size(); ++s) frozen_step(R->steps[s], v, w, p, mu, &tapes[s]);
    std::copy(p.begin(), p.end(), p_final);
    std::vector<double> av(3 * nb, 0.0), aw(3 * nb, 0.0), ap(adj_p_final, adj_p_final + 3 * nb), amu(nb, 0.0);
    for (std::size_t s = R->steps.size(); s-- > 0;) frozen_step_adjoint(R->steps[s], tapes[s], mu, av, aw, ap, amu);
    std::copy(av.begin(), av.end(), g_v0);
    std::copy(aw.begin(), aw.end(), g_w0);
    std::copy(amu.begin(), amu.end(), g_mu);
    return 0;
}

// One Dy step of the FP32 restatement. bodies are updated in place;
// patches_out[nm], impulses_out[nm][4] (per manifold contact), parts_out[nm]
// (conflict partition of each manager within the step's single task).
int orc_dy_step(float dt, std::uint32_t nb, fyzz_dy_body* bodies, std::uint32_t ns, const fyzz_dy_static* statics,
                std::uint32_t nm, const fyzz_dy_manager* managers, const fyzz_dy_patch* history,
                std::uint32_t nh, fyzz_dy_patch* patches_out, float* impulses_out, std::uint32_t* parts_out) {
    Dy<float>::patch_order_.assign(nm, {});
    return Dy<float>::step(dt, nb, bodies, ns, statics, nm, managers, history, nh, patches_out, impulses_out,
                           parts_out, nullptr);
}

// orc_dy_step plus the friction write-back of the step's managers into
// anchors_out[nm] (Scene::write_back_contacts, scene.cpp:9339-9388).
int orc_dy_step_anchors(float dt, std::uint32_t nb, fyzz_dy_body* bodies, std::uint32_t ns,
                        const fyzz_dy_static* statics, std::uint32_t nm, const fyzz_dy_manager* managers,
                        const fyzz_dy_patch* history, std::uint32_t nh, fyzz_dy_patch* patches_out,
                        float* impulses_out, std::uint32_t* parts_out, fyzz_friction_anchors* anchors_out) {
    Dy<float>::patch_order_.assign(nm, {});
    Dy<float>::anchors_out_ = anchors_out;
    const int rc = Dy<float>::step(dt, nb, bodies, ns, statics, nm, managers, history, nh, patches_out, impulses_out,
                                   parts_out, nullptr);
    Dy<float>::anchors_out_ = nullptr;
    return rc;
}

// The FP64 twin: same inputs; state64[nb][13] receives position, linear
// velocity, orientation, angular velocity in double. bodies are updated
// with the rounded values.
int orc_dy_step64(float dt, std::uint32_t nb, fyzz_dy_body* bodies, std::uint32_t ns,
                  const fyzz_dy_static* statics, std::uint32_t nm, const fyzz_dy_manager* managers,
                  const fyzz_dy_patch* history, std::uint32_t nh, fyzz_dy_patch* patches_out, float* impulses_out,
                  double* state64) {
    Dy<double>::patch_order_.assign(nm, {});
    return Dy<double>::step(double(dt), nb, bodies, ns, statics, nm, managers, history, nh, patches_out,
                            impulses_out, nullptr, state64);
}

}  // extern "C"
