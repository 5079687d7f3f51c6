// PCM narrowphase on gfx950: persistent box/box and plane/box contact
// manifolds with the sparse GJK/EPA continuation.
//
// Semantics follow the reference NativeCpu Gu/Sc code (paths relative to
// the reference root):
//   manifold refresh                  src/native/scene.cpp:3432-3465
//   box/box manifold + thresholds     src/native/scene.cpp:3466-3654
//   five-point reduction              src/native/scene.cpp:3317-3431
//   plane/box manifold                src/native/scene.cpp:3774-3955
//   SAT, reference face, clipping     src/native/GuPCMContactBoxBox.cpp:20-436
//   GJK/EPA continuation              src/native/GuPCMGjkEpa.cpp:18-566
// Contacts must agree with the reference to 1e-5 and the manifold cache
// makes discrete regeneration choices, so every expression keeps the
// reference's floating-point operation order (contraction is off for the
// whole library). The reference uses two quaternion rotations with
// different association: scene.cpp computes (v + w t) + u x t (`srot`),
// the Gu files v + (t w + u x t) (`grot`).
//
// Work split: k_pcm_box_box runs one thread per pair (refresh, regeneration
// test, SAT + clipping into a 16-slot candidate array, batch reduction,
// emission). Pairs whose SAT overlaps but clip yields no witness are
// appended to a device list and finished by k_pcm_gjk (GJK + EPA with
// 64-vertex/64-facet working sets), the sparse continuation.
// k_pcm_plane_box runs one thread per dynamic box against the floor.
#include <algorithm>
#include <memory>
#include <string>

#include "../../include/fyzz_hip.h"
#include "context.hpp"
#include "stages.hpp"
#include "vecmath.hpp"

namespace fyzz_hip::pcm {

using namespace fyzz_hip::vm;

struct Pt {  // persistent witness: local_a, local_b, normal_b, separation
    V a, b, n;
    float s;
};

// Cache record, 224 B: Scene::PcmManifoldCache (scene.h:1448-1459)
// without the write-only GJK warm start.
struct alignas(16) Cache {
    Pt pts[4];
    V rel_pos;
    Q rel_rot, orient_a, orient_b;
    u32 flags;  // count | initialized << 8
};
static_assert(sizeof(Cache) == 224, "cache layout");

struct Box3 {
    V half, pos;
    Q rot;
};
__device__ inline Box3 load_box(const float* pos, const float* quat, const float* half, u32 s) {
    return {{half[3 * s], half[3 * s + 1], half[3 * s + 2]},
            {pos[3 * s], pos[3 * s + 1], pos[3 * s + 2]},
            {quat[4 * s], quat[4 * s + 1], quat[4 * s + 2], quat[4 * s + 3]}};
}

// ---------------- GuPCMContactSphere.cpp ----------------
// Spheres carry no persistent manifold: one stateless contact per step
// (position, normal, separation), in the reference's operation order.
struct SPt {
    V p, n;
    float s;
};
// source_pcm_sphere_sphere (GuPCMContactSphere.cpp:49-67): the normal
// points from the second center to the first, the witness on the second
// sphere's surface.
__device__ inline u32 sphere_sphere(V c0, float r0, V c1, float r1, float cd, SPt& o) {
    const V d = vsub(c0, c1);
    const float d2 = vdot(d, d);
    const float rs = r0 + r1;
    const float inflated = rs + cd;
    if (!(inflated * inflated > d2)) return 0u;
    const float dist = sqrtf(d2);
    const V n = dist <= 0.00001f ? V{1.f, 0.f, 0.f} : vmul(d, 1.f / dist);
    o = {vadd(c1, vmul(n, r1)), n, dist - rs};
    return 1u;
}
// source_pcm_sphere_plane (GuPCMContactSphere.cpp:69-79): the horizontal
// plane at height py, the +y normal the solver receives.
__device__ inline u32 sphere_plane(V c, float r, float py, float cd, SPt& o) {
    const float sep = c.y - py - r;
    if (!(cd >= sep)) return 0u;
    o = {{c.x, c.y - r, c.z}, {0.f, 1.f, 0.f}, sep};
    return 1u;
}
// source_pcm_sphere_box (GuPCMContactSphere.cpp:81-129): the closest point
// of the oriented box, or the nearest exit face when the center is inside;
// the normal points from the box to the sphere.
__device__ inline u32 sphere_box(V c, float r, V bc, Q bq, V h, float cd, SPt& o) {
    const V local = grot_inv(bq, vsub(c, bc));
    const V cl{sclamp(local.x, -h.x, h.x), sclamp(local.y, -h.y, h.y), sclamp(local.z, -h.z, h.z)};
    const V gap = vsub(local, cl);
    const float l2 = vdot(gap, gap);
    const float inflated = r + cd;
    if (!(inflated * inflated > l2)) return 0u;
    const bool inside = fabsf(local.x) <= h.x && fabsf(local.y) <= h.y && fabsf(local.z) <= h.z;
    if (inside) {
        const V dist{h.x - fabsf(cl.x), h.y - fabsf(cl.y), h.z - fabsf(cl.z)};
        V nl;
        float depth;
        if (dist.z <= dist.x && dist.z <= dist.y) {
            nl = {0.f, 0.f, cl.z >= 0.f ? 1.f : -1.f};
            depth = dist.z;
        } else if (dist.x <= dist.y && dist.x <= dist.z) {
            nl = {cl.x >= 0.f ? 1.f : -1.f, 0.f, 0.f};
            depth = dist.x;
        } else {
            nl = {0.f, cl.y >= 0.f ? 1.f : -1.f, 0.f};
            depth = dist.y;
        }
        const V n = grot(bq, nl);
        o = {vadd(c, vmul(n, depth)), n, -depth - r};
        return 1u;
    }
    const float len = sqrtf(l2);
    const V n = grot(bq, vmul(gap, 1.f / len));
    o = {vadd(bc, grot(bq, cl)), n, len - r};
    return 1u;
}
// ---------------- GuPCMContactBoxBox.cpp ----------------
struct GBox {
    V half, pos;
    Q rot;
    V axis[3];
};
__device__ inline GBox make_gbox(const Box3& b) {
    GBox g{b.half, b.pos, b.rot, {}};
    g.axis[0] = grot(b.rot, {1, 0, 0});
    g.axis[1] = grot(b.rot, {0, 1, 0});
    g.axis[2] = grot(b.rot, {0, 0, 1});
    return g;
}
struct Frame {
    V x, y, z, p;
};
__device__ inline V f_project(const Frame& f, V w) {
    const V d = vsub(w, f.p);
    return {vdot(d, f.x), vdot(d, f.y), vdot(d, f.z)};
}
__device__ inline V f_unproject(const Frame& f, V v) {
    return vadd(f.p, vadd(vadd(vmul(f.x, v.x), vmul(f.y, v.y)), vmul(f.z, v.z)));
}
// GuPCMContactBoxBox.cpp:63-83
__device__ bool poly_contains(const V* poly, V p, V low, V high) {
    if (p.x < low.x || p.y < low.y || p.x > high.x || p.y > high.y) return false;
    unsigned hits = 0;
    for (unsigned i = 0, j = 3; i < 4; j = i++) {
        const V a = poly[j], b = poly[i];
        if ((p.x == a.x && p.y == a.y) || (p.x == b.x && p.y == b.y)) return true;
        if ((a.y > p.y) != (b.y > p.y)) {
            const float dx = b.x - a.x, dy = b.y - a.y;
            const float part = (p.y - a.y) * dx + (a.x + kFltEps) * dy;
            const float target = p.x * dy;
            const bool hit = dy > 0.0f ? part >= target : target >= part;
            if (hit && ++hits == 2) return false;
        }
    }
    return hits > 0;
}
// GuPCMContactBoxBox.cpp:87-104
__device__ bool segment_slab(V p, V d, V high, V low, float& enter, float& exit) {
    enter = 0.0f;
    exit = 1.0f;
    for (unsigned a = 0; a < 3; ++a) {
        const float x = comp(p, a), dd = comp(d, a);
        const float mn = comp(low, a), mx = comp(high, a);
        if (fabsf(dd) < 1.0e-6f) {
            if (x < mn || x > mx) return false;
            continue;
        }
        const float one = (mn - x) / dd, two = (mx - x) / dd;
        enter = smax(enter, smin(one, two));
        exit = smin(exit, smax(one, two));
    }
    return enter <= exit && enter <= 1.0f;
}

constexpr int kMaxCand = 16;  // 4 inside corners + 4 projected + 8 edge cuts
constexpr u32 kWitnessThreads = 128;  // largest workgroup that keeps witnesses in LDS

// The clip witnesses of one pair. Every witness the clipper emits has its
// reference-face point at the incident point's (x, y, 0) in the clip frame
// and separation -z, so the incident point alone is stored, in three LDS
// planes indexed [plane][witness][thread] (conflict-free); the witness
// (point on A, point on B, normal in B, separation) is rebuilt from it with
// the arithmetic that generated it.
struct Witnesses {
    float* lds;  // this thread's entry of plane 0, witness 0
    u32 stride;  // threads per workgroup
    int count;
    Frame f;
    Q arot, brot;
    V apos, bpos;
    V nb;  // contact normal in B's frame (the same for every witness)
    bool flipped;
    __device__ float& at(int plane, int k) const { return lds[(plane * kMaxCand + k) * stride]; }
    __device__ void emit(V inc) {
        if (count >= kMaxCand) return;  // unreachable: at most 16 witnesses
        at(0, count) = inc.x;
        at(1, count) = inc.y;
        at(2, count) = inc.z;
        ++count;
    }
    __device__ V inc(int k) const { return {at(0, k), at(1, k), at(2, k)}; }
    __device__ float s(int k) const { return -at(2, k); }
    __device__ V n(int) const { return nb; }
    __device__ V b(int k) const {
        const V in = inc(k);
        const V w1 = f_unproject(f, flipped ? V{in.x, in.y, 0} : in);
        return grot_inv(brot, vsub(w1, bpos));
    }
    __device__ Pt pt(int k) const {
        const V in = inc(k);
        const V ref{in.x, in.y, 0};
        const V w0 = f_unproject(f, flipped ? in : ref);
        const V w1 = f_unproject(f, flipped ? ref : in);
        return {grot_inv(arot, vsub(w0, apos)), grot_inv(brot, vsub(w1, bpos)), nb, -in.z};
    }
};

// GuPCMContactBoxBox.cpp:119-178
__device__ void clip_contacts(Witnesses& g, const V* poly, V inc_n, float fx, float fy, float cd) {
    fx *= 1.0001f;
    fy *= 1.0001f;
    bool pen[4], in[4];
    V mn{kFltMax, kFltMax, kFltMax};
    V mx = vneg(mn);
    const V high{fx, fy, kFltMax};
    for (unsigned i = 0; i < 4; ++i) {
        const V p = poly[i];
        mn.x = smin(mn.x, p.x);
        mn.y = smin(mn.y, p.y);
        mx.x = smax(mx.x, p.x);
        mx.y = smax(mx.y, p.y);
        pen[i] = false;
        in[i] = false;
        if (-p.z < cd) {
            pen[i] = true;
            in[i] = fabsf(p.x) <= fx && fabsf(p.y) <= fy;
            if (in[i]) g.emit(p);
        }
    }
    if (g.count == 4) return;
    if (inc_n.z != 0.0f) {
        const V corners[4] = {{fx, fy, 0}, {fx, -fy, 0}, {-fx, fy, 0}, {-fx, -fy, 0}};
        for (int k = 0; k < 4; ++k) {
            const V c = corners[k];
            if (!poly_contains(poly, c, mn, mx)) continue;
            const float nom = vdot(inc_n, vsub(poly[0], c));
            const float t = nom / inc_n.z;
            const float sep = -t;
            if (cd > sep) g.emit({c.x, c.y, t});
        }
    }
    const V low{-fx, -fy, -cd - kFltEps};
    for (unsigned s = 0, e = 3; s < 4; e = s++) {
        if (!pen[s] && !pen[e]) continue;
        const bool c0 = pen[s] && in[s];
        const bool c1 = pen[e] && in[e];
        if (c0 && c1) continue;
        const V edge = vsub(poly[e], poly[s]);
        float t0, t1;
        if (!segment_slab(poly[s], edge, high, low, t0, t1)) continue;
        if (!c0) {
            const V p = vadd(poly[s], vmul(edge, t0));
            g.emit(p);
        }
        if (!c1) {
            const V p = vadd(poly[s], vmul(edge, t1));
            g.emit(p);
        }
    }
}

// GuPCMContactBoxBox.cpp:301-436. Returns overlap; gen.count = #witnesses
// (gen.lds and gen.stride set by the caller).
__device__ bool box_box_candidates(const Box3& ba, const Box3& bb, float cd, Witnesses& gen) {
    const GBox a = make_gbox(ba), b = make_gbox(bb);
    gen.count = 0;
    const V delta = vsub(b.pos, a.pos);
    float sgn[6], ov[6];
    constexpr float se = 1.0e-6f;
    for (unsigned i = 0; i < 6; ++i) {
        const bool first = i < 3;
        const GBox& ref = first ? a : b;
        const GBox& inc = first ? b : a;
        const unsigned ax = i % 3;
        const V n = ref.axis[ax];
        sgn[i] = vdot(delta, n);
        const float r = (fabsf(vdot(n, inc.axis[0])) + se) * inc.half.x +
                        (fabsf(vdot(n, inc.axis[1])) + se) * inc.half.y +
                        (fabsf(vdot(n, inc.axis[2])) + se) * inc.half.z;
        ov[i] = comp(ref.half, ax) + r - fabsf(sgn[i]) + cd;
        if (ov[i] < 0.0f) return false;
    }
    for (unsigned i = 0; i < 3; ++i)
        for (unsigned j = 0; j < 3; ++j) {
            const V ca = vcross(a.axis[i], b.axis[j]);
            float ra = 0.0f, rb = 0.0f;
            for (unsigned k = 0; k < 3; ++k) {
                if (k != i) ra += (fabsf(vdot(ca, a.axis[k])) + se) * comp(a.half, k);
                if (k != j) rb += (fabsf(vdot(ca, b.axis[k])) + se) * comp(b.half, k);
            }
            if (fabsf(vdot(delta, ca)) > ra + rb + cd) return false;
        }
    unsigned feature = 0;
    float mo = ov[0], msgn = sgn[0];
#pragma unroll
    for (unsigned i = 1; i < 6; ++i)
        if (mo > ov[i]) {
            mo = ov[i];
            msgn = sgn[i];
            feature = i;
        }
    const bool flipped = feature >= 3;
    const GBox inc = flipped ? a : b;
    const bool np = msgn <= 0.0f;
    Frame fr;
    V normal;
    float fx, fy;
    switch (feature) {
        case 0:
            normal = np ? a.axis[0] : vneg(a.axis[0]);
            fr = {np ? vneg(a.axis[2]) : a.axis[2], a.axis[1], normal,
                  vadd(a.pos, vmul(a.axis[0], np ? -a.half.x : a.half.x))};
            fx = a.half.z;
            fy = a.half.y;
            break;
        case 1:
            normal = np ? a.axis[1] : vneg(a.axis[1]);
            fr = {a.axis[0], np ? vneg(a.axis[2]) : a.axis[2], normal,
                  vadd(a.pos, vmul(a.axis[1], np ? -a.half.y : a.half.y))};
            fx = a.half.x;
            fy = a.half.z;
            break;
        case 2:
            normal = np ? a.axis[2] : vneg(a.axis[2]);
            fr = {a.axis[0], np ? a.axis[1] : vneg(a.axis[1]), normal,
                  vadd(a.pos, vmul(a.axis[2], np ? -a.half.z : a.half.z))};
            fx = a.half.x;
            fy = a.half.y;
            break;
        case 3:
            normal = np ? b.axis[0] : vneg(b.axis[0]);
            fr = {np ? b.axis[2] : vneg(b.axis[2]), b.axis[1], vneg(normal),
                  vadd(b.pos, vmul(b.axis[0], np ? b.half.x : -b.half.x))};
            fx = b.half.z;
            fy = b.half.y;
            break;
        case 4:
            normal = np ? b.axis[1] : vneg(b.axis[1]);
            fr = {b.axis[0], np ? b.axis[2] : vneg(b.axis[2]), vneg(normal),
                  vadd(b.pos, vmul(b.axis[1], np ? b.half.y : -b.half.y))};
            fx = b.half.x;
            fy = b.half.z;
            break;
        default:
            normal = np ? b.axis[2] : vneg(b.axis[2]);
            fr = {b.axis[0], np ? vneg(b.axis[1]) : b.axis[1], vneg(normal),
                  vadd(b.pos, vmul(b.axis[2], np ? b.half.z : -b.half.z))};
            fx = b.half.x;
            fy = b.half.y;
            break;
    }
    const V search = vneg(fr.z);
    const float dx = fabsf(vdot(search, inc.axis[0]));
    const float dy = fabsf(vdot(search, inc.axis[1]));
    const float dz = fabsf(vdot(search, inc.axis[2]));
    const unsigned ia = dx >= dy && dx >= dz ? 0 : (dy >= dz ? 1 : 2);
    const unsigned ua = ia == 0 ? 1 : 0;
    const unsigned va = ia == 2 ? 1 : 2;
    const V axis_i = ia == 0 ? inc.axis[0] : (ia == 1 ? inc.axis[1] : inc.axis[2]);
    const V axis_u = ua == 0 ? inc.axis[0] : inc.axis[1];
    const V axis_v = va == 1 ? inc.axis[1] : inc.axis[2];
    const float facing = vdot(search, axis_i);
    const float sext = facing > 0.0f ? -comp(inc.half, ia) : comp(inc.half, ia);
    const V fc = vadd(inc.pos, vmul(axis_i, sext));
    const V u = vmul(axis_u, comp(inc.half, ua));
    const V v = vmul(axis_v, comp(inc.half, va));
    V poly[4];
    poly[0] = f_project(fr, vadd(fc, vadd(u, v)));
    poly[1] = f_project(fr, vadd(fc, vsub(u, v)));
    poly[2] = f_project(fr, vsub(fc, vadd(u, v)));
    poly[3] = f_project(fr, vsub(fc, vsub(u, v)));
    const V fnorm = f_project(fr, vadd(fr.p, facing > 0.0f ? vneg(axis_i) : axis_i));
    gen.f = fr;
    gen.arot = a.rot;
    gen.apos = a.pos;
    gen.brot = b.rot;
    gen.bpos = b.pos;
    gen.nb = grot_inv(b.rot, normal);
    gen.flipped = flipped;
    clip_contacts(gen, poly, fnorm, fx, fy, cd);
    return true;
}

// A list of at most 16 witness numbers, 4 bits each (registers only).
struct Avail {
    u64 v;
    __device__ int get(int i) const { return static_cast<int>((v >> (4 * i)) & 15u); }
    __device__ void set(int i, int x) { v = (v & ~(15ull << (4 * i))) | (static_cast<u64>(x) << (4 * i)); }
};

// GuPCMContactBoxBox.cpp:185-299
__device__ void batch_selection(const Witnesses& c, int n, float tol, int sel[4]) {
    Avail avail{0xfedcba9876543210ull};
    int rem = n;
    float deepest_d = c.s(0), shallow = deepest_d;
    int deepest = 0;
    for (int i = 1; i < rem; ++i) {
        const float d = c.s(i);
        shallow = smax(shallow, d);
        if (deepest_d > d) {
            deepest_d = d;
            deepest = i;
        }
    }
    sel[0] = deepest;
    avail.set(deepest, avail.get(--rem));
    const V b0 = c.b(sel[0]);
    int fp = 0;
    float far2;
    {
        const V d = vsub(c.b(avail.get(0)), b0);
        far2 = vdot(d, d);
    }
    for (int i = 1; i < rem; ++i) {
        const V d = vsub(c.b(avail.get(i)), b0);
        const float v = vdot(d, d);
        if (v > far2) {
            far2 = v;
            fp = i;
        }
    }
    sel[1] = avail.get(fp);
    avail.set(fp, avail.get(--rem));
    const V base = vsub(c.b(sel[1]), b0);
    V tr = vcross(base, c.n(sel[0]));
    const float tr2 = vdot(tr, tr);
    if (tr2 > 0.0f) {
        const float inv = 1.0f / sqrtf(tr2);
        tr = {tr.x * inv, tr.y * inv, tr.z * inv};
    } else {
        tr = c.n(sel[0]);
    }
    float greatest = -kFltMax, least = kFltMax;
    int gp = 0, lp = 0;
    for (int i = 0; i < rem; ++i) {
        const float side = vdot(vsub(c.b(avail.get(i)), b0), tr);
        if (side > greatest) {
            greatest = side;
            gp = i;
        }
        if (side < least) {
            least = side;
            lp = i;
        }
    }
    sel[2] = avail.get(gp);
    avail.set(gp, avail.get(--rem));
    if (lp == rem) lp = gp;
    if (least * greatest > 0.0f) {
        greatest = -kFltMax;
        for (int i = 0; i < rem; ++i) {
            const float side = vdot(vsub(c.b(avail.get(i)), b0), tr);
            if (side > greatest) {
                greatest = side;
                lp = i;
            }
        }
    }
    sel[3] = avail.get(lp);
    avail.set(lp, avail.get(--rem));
    const float th = tol * 0.02f;
    if (deepest_d < th && shallow > th) {
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            float depth = c.s(sel[i]);
            if (depth > th) {
                int rep = rem;
                for (int j = 0; j < rem; ++j) {
                    const float cdp = c.s(avail.get(j));
                    if (depth > cdp && th > cdp) {
                        depth = cdp;
                        rep = j;
                    }
                }
                if (rep < rem) {
                    const int t = sel[i];
                    sel[i] = avail.get(rep);
                    avail.set(rep, t);
                }
            }
        }
    }
}

// ---------------- GuPCMGjkEpa.cpp (box/box polyhedral cores) ----------------
__device__ inline float box_min_margin(V h) { return 0.05f * smin(smin(h.x, h.y), h.z); }
struct Sup {
    V a, b, q;
    u32 ai, bi;
};
__device__ inline V vsel3(bool c, V x, V y) { return {c ? x.x : y.x, c ? x.y : y.y, c ? x.z : y.z}; }
// (field-wise select: keeps a simplex slot written at a runtime position in
// registers)
__device__ inline Sup ssel(bool c, const Sup& x, const Sup& y) {
    return {vsel3(c, x.a, y.a), vsel3(c, x.b, y.b), vsel3(c, x.q, y.q), c ? x.ai : y.ai, c ? x.bi : y.bi};
}
struct GPair {
    V ha, hb, a_in_b;
    Q rel;
};
__device__ inline V box_vertex(V e, V d, u32& idx) {
    const bool x = d.x > 0.f, y = d.y > 0.f, z = d.z > 0.f;
    idx = (x ? 1u : 0u) | (y ? 2u : 0u) | (z ? 4u : 0u);
    return {x ? e.x : -e.x, y ? e.y : -e.y, z ? e.z : -e.z};
}
__device__ inline Sup support(const GPair& p, V dir) {
    const V da = grot(Q{-p.rel.x, -p.rel.y, -p.rel.z, p.rel.w}, dir);
    Sup s;
    const V pa = box_vertex(p.ha, da, s.ai);
    s.b = box_vertex(p.hb, vneg(dir), s.bi);
    s.a = vadd(p.a_in_b, grot(p.rel, pa));
    s.q = vsub(s.a, s.b);
    return s;
}
__device__ inline V pair_local_a(const GPair& p, V pb) {
    return grot(Q{-p.rel.x, -p.rel.y, -p.rel.z, p.rel.w}, vsub(pb, p.a_in_b));
}
// GuPCMGjkEpa.cpp:87-112 (explicit fused multiply-add, as written there),
// for an N x N system with its right-hand side in column N. Every index is
// a compile-time constant (the pivot row swap is a select over the rows),
// so the system stays in registers.
template <int N>
__device__ inline bool solve_n(float (&a)[3][4], float (&r)[3]) {
    float scale = 0.0f;
#pragma unroll
    for (int i = 0; i < N; ++i) scale = smax(scale, fabsf(a[i][i]));
    if (!(scale > 0.0f) || !isfinite(scale)) return false;
    const float tol = kFltEps * scale;
#pragma unroll
    for (int col = 0; col < N; ++col) {
        int piv = col;
        float best = fabsf(a[col][col]);
#pragma unroll
        for (int row = col + 1; row < N; ++row) {
            const float v = fabsf(a[row][col]);
            if (v > best) {
                piv = row;
                best = v;
            }
        }
        if (!(best > tol)) return false;
#pragma unroll
        for (int row = col + 1; row < N; ++row)
            if (piv == row) {
#pragma unroll
                for (int j = col; j <= N; ++j) {
                    const float t = a[col][j];
                    a[col][j] = a[row][j];
                    a[row][j] = t;
                }
            }
        const float div = a[col][col];
#pragma unroll
        for (int j = col; j <= N; ++j) a[col][j] /= div;
#pragma unroll
        for (int row = 0; row < N; ++row)
            if (row != col) {
                const float f = a[row][col];
#pragma unroll
                for (int j = col; j <= N; ++j) a[row][j] = fmaf(-f, a[col][j], a[row][j]);
            }
    }
#pragma unroll
    for (int i = 0; i < N; ++i) r[i] = a[i][N];
    return true;
}
struct Closest {
    V p, a, b;
    float w[4];
    unsigned mask;
    float d2;
};
// GuPCMGjkEpa.cpp:119-161 for a simplex of COUNT vertices: every subset in
// increasing mask order, barycentric weights from the subset's normal
// equations, the feasible point nearest the origin (first one on ties).
// Fully unrolled, so the subsets' index lists are constants.
template <int COUNT>
__device__ Closest closest_simplex_n(const Sup (&sx)[4]) {
    Closest best;
    best.p = best.a = best.b = V{0, 0, 0};
    best.w[0] = best.w[1] = best.w[2] = best.w[3] = 0.0f;
    best.mask = 0;
    best.d2 = kFltMax;
#pragma unroll
    for (unsigned mask = 1; mask < (1u << COUNT); ++mask) {
        int ids[4] = {0, 0, 0, 0}, n = 0;
#pragma unroll
        for (int i = 0; i < COUNT; ++i)
            if (mask & (1u << i)) ids[n++] = i;
        float w[4] = {0, 0, 0, 0};
        w[ids[0]] = 1.0f;
        bool solved = true;
        if (n > 1) {
            const V base = sx[ids[0]].q;
            V dirs[3];
            float m[3][4];
#pragma unroll
            for (int i = 1; i < n; ++i) dirs[i - 1] = vsub(sx[ids[i]].q, base);
#pragma unroll
            for (int i = 1; i < n; ++i) {
#pragma unroll
                for (int j = 1; j < n; ++j) m[i - 1][j - 1] = vdot(dirs[i - 1], dirs[j - 1]);
                m[i - 1][n - 1] = -vdot(dirs[i - 1], base);
            }
            float co[3];
            solved = n == 2 ? solve_n<1>(m, co) : n == 3 ? solve_n<2>(m, co) : solve_n<3>(m, co);
            if (solved) {
#pragma unroll
                for (int i = 1; i < n; ++i) {
                    w[ids[i]] = co[i - 1];
                    w[ids[0]] -= co[i - 1];
                }
            }
        }
        if (!solved) continue;
        bool feasible = true;
#pragma unroll
        for (int i = 0; i < n; ++i)
            if (w[ids[i]] < -1.0e-5f) feasible = false;
        if (!feasible) continue;
        V q{0, 0, 0}, pa{0, 0, 0}, pb{0, 0, 0};
#pragma unroll
        for (int i = 0; i < COUNT; ++i) {
            q = vadd(q, vmul(sx[i].q, w[i]));
            pa = vadd(pa, vmul(sx[i].a, w[i]));
            pb = vadd(pb, vmul(sx[i].b, w[i]));
        }
        const float d = vdot(q, q);
        if (d < best.d2) {
            best.p = q;
            best.a = pa;
            best.b = pb;
#pragma unroll
            for (int k = 0; k < 4; ++k) best.w[k] = w[k];
            best.mask = mask;
            best.d2 = d;
        }
    }
    return best;
}
__device__ inline Closest closest_simplex(const Sup (&sx)[4], int count) {
    switch (count) {
    case 1:
        return closest_simplex_n<1>(sx);
    case 2:
        return closest_simplex_n<2>(sx);
    case 3:
        return closest_simplex_n<3>(sx);
    default:
        return closest_simplex_n<4>(sx);
    }
}
__device__ bool contact_from_gjk(const GPair& pr, const Closest& c, Pt& out) {
    const float dist = sqrtf(c.d2);
    if (!(dist > 0.0f) || !isfinite(dist)) return false;
    const V n = vmul(c.p, 1.0f / dist);
    out = {pair_local_a(pr, c.a), c.b, n, dist};
    return true;
}
struct Facet {
    int i, j, k;
    V n;
    float d;
    bool valid, in_heap;
};
__device__ Facet make_facet_q(V a, V b, V c, int i, int j, int k) {
    Facet f;
    f.valid = false;
    f.in_heap = true;
    const V e0 = vsub(b, a), e1 = vsub(c, a);
    V n = vcross(e0, e1);
    const float n2 = vdot(n, n);
    if (!(n2 > kFltEps)) return f;
    n = vmul(n, 1.0f / sqrtf(n2));
    f.i = i;
    f.j = j;
    f.k = k;
    f.n = n;
    f.d = vdot(n, a);
    f.valid = true;
    return f;
}
__device__ inline Facet make_facet(const Sup* v, int i, int j, int k) {
    return make_facet_q(v[i].q, v[j].q, v[k].q, i, j, k);
}
// (all slots indexed with constants: the polytope stays in registers)
__device__ bool init_epa(const GPair& pr, Sup (&v)[4], int& count, Facet (&fa)[4], int& fc) {
    if (count <= 0 || count > 4) return false;
    if (count == 1) {
        const Sup nx = support(pr, {1, 0, 0});
        const V d = vsub(nx.q, v[0].q);
        if (vdot(d, d) == 0.0f) return false;
        v[1] = nx;
        count = 2;
    }
    if (count == 2) {
        const V e = vsub(v[1].q, v[0].q);
        const V ab{fabsf(e.x), fabsf(e.y), fabsf(e.z)};
        V ax{1, 0, 0};
        if (ab.x > ab.y && ab.z > ab.y)
            ax = {0, 1, 0};
        else if (ab.x > ab.z)
            ax = {0, 0, 1};
        V perp = vcross(ax, e);
        const float sq = vdot(perp, perp);
        if (!(sq > 0.0f)) return false;
        perp = vmul(perp, 1.0f / sqrtf(sq));
        v[2] = support(pr, perp);
        count = 3;
    }
    if (count == 3) {
        const Facet f0 = make_facet(v, 0, 1, 2);
        const Facet f1 = make_facet(v, 1, 0, 2);
        if (!f0.valid || !f1.valid) return false;
        fa[0] = f0;
        fa[1] = f1;
        fc = 2;
        return true;
    }
    const V p0 = v[0].q;
    const V v1 = vsub(v[1].q, p0), v2 = vsub(v[2].q, p0);
    if (vdot(vcross(v1, v2), vsub(v[3].q, p0)) > 0.0f) {
        const Sup t = v[1];
        v[1] = v[2];
        v[2] = t;
    }
    const int faces[4][3] = {{0, 1, 2}, {0, 3, 1}, {0, 2, 3}, {1, 3, 2}};
    bool valid = true;
#pragma unroll
    for (int f = 0; f < 4; ++f) {
        fa[f] = make_facet(v, faces[f][0], faces[f][1], faces[f][2]);
        valid = valid && fa[f].valid;
    }
    fc = 4;
    return valid;
}
__device__ bool contact_from_facet(const GPair& pr, const Sup& p, const Sup& q, const Sup& r, const Facet& f,
                                   Pt& out) {
    const V e0 = vsub(q.q, p.q), e1 = vsub(r.q, p.q);
    const float a = vdot(e0, e0), b = vdot(e0, e1), c = vdot(e1, e1);
    const float den = a * c - b * b;
    if (!(den > kFltEps * a * c)) return false;
    const float x = -vdot(p.q, e0), y = -vdot(p.q, e1);
    const float u = (x * c - y * b) / den;
    const float vv = (y * a - x * b) / den;
    const float w = 1.0f - u - vv;
    const V pa = vadd(vadd(vmul(p.a, w), vmul(q.a, u)), vmul(r.a, vv));
    const V pb = vadd(vadd(vmul(p.b, w), vmul(q.b, u)), vmul(r.b, vv));
    if (!isfinite(pa.x) || !isfinite(pb.x) || !isfinite(f.d)) return false;
    const float depth = fabsf(f.d);
    out = {pair_local_a(pr, pa), pb, vneg(f.n), -depth};
    const V diff = vsub(pa, pb);
    return sqrtf(vdot(diff, diff)) <= depth + 1.0e-3f;
}
// ---- EPA, one wave per pair (GuPCMGjkEpa.cpp:302-395) ----
// Lane t holds vertex t and facet t of the working sets (at most 64 of
// each) and horizon edge t (at most 32). Every step keeps the sequential
// order of the reference: the closest facet is the lowest (distance,
// index) among live facets; the horizon is built over the visible facets
// in index order, three edges each, a repeated edge cancelling its
// reverse (the last entry moves into its place); new facets fill the
// freed slots in increasing order, then append. All lanes run the same
// control flow and hold the same uniform values.
__device__ inline V shfl_v(V x, int src) {
    return {__shfl(x.x, src), __shfl(x.y, src), __shfl(x.z, src)};
}
__device__ inline Sup shfl_sup(const Sup& x, int src) {
    Sup r;
    r.a = shfl_v(x.a, src);
    r.b = shfl_v(x.b, src);
    r.q = shfl_v(x.q, src);
    r.ai = static_cast<u32>(__shfl(static_cast<int>(x.ai), src));
    r.bi = static_cast<u32>(__shfl(static_cast<int>(x.bi), src));
    return r;
}
__device__ inline Facet shfl_facet(const Facet& x, int src) {
    Facet r;
    r.i = __shfl(x.i, src);
    r.j = __shfl(x.j, src);
    r.k = __shfl(x.k, src);
    r.n = shfl_v(x.n, src);
    r.d = __shfl(x.d, src);
    r.valid = __shfl(x.valid ? 1 : 0, src) != 0;
    r.in_heap = __shfl(x.in_heap ? 1 : 0, src) != 0;
    return r;
}
// sx: the simplex (sc vertices) held by every lane.
__device__ bool epa_wave(const GPair& pr, const Sup (&sx)[4], int sc, Pt& out) {
    const int lane = static_cast<int>(lane_id());
    // the initial polytope, built by lane 0 and spread one entry per lane
    Sup v4[4];
    Facet f4[4];
    int vc = sc, fc = 0, ok = 0;
    if (lane == 0) {
#pragma unroll
        for (int i = 0; i < 4; ++i) v4[i] = sx[i];
        ok = init_epa(pr, v4, vc, f4, fc) ? 1 : 0;
    }
    if (!__shfl(ok, 0)) return false;
    vc = __shfl(vc, 0);
    fc = __shfl(fc, 0);
    Sup myv{};
    Facet myf{};
    myf.valid = false;
    myf.in_heap = false;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const Sup vq = shfl_sup(v4[q], 0);
        const Facet fq = shfl_facet(f4[q], 0);
        if (lane == q) {
            if (q < vc) myv = vq;
            if (q < fc) myf = fq;
        }
    }
    int bfrom = 0, bto = 0;
    const float eps = 0.1f * smin(box_min_margin(pr.ha), box_min_margin(pr.hb));
    float upper = kFltMax;
    const u64 below = (1ull << lane) - 1ull;
    for (int it = 0; it < 60; ++it) {
        // closest live facet: lowest distance below FLT_MAX, lowest index on ties
        float key = (lane < fc && myf.valid && myf.in_heap && myf.d < kFltMax) ? myf.d : __builtin_inff();
        int idx = lane;
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
            const float ok = __shfl_xor(key, off);
            const int oi = __shfl_xor(idx, off);
            if (ok < key || (ok == key && oi < idx)) {
                key = ok;
                idx = oi;
            }
        }
        if (!(key < __builtin_inff())) return false;
        const Facet sel = shfl_facet(myf, idx);
        const float dist = key;
        const Sup np = support(pr, sel.n);
        const float proj = vdot(np.q, sel.n);
        if (fabsf(proj - dist) <= eps)
            return contact_from_facet(pr, shfl_sup(myv, sel.i & 63), shfl_sup(myv, sel.j & 63),
                                      shfl_sup(myv, sel.k & 63), sel, out);
        if (vc == 64) return false;
        upper = smin(upper, proj);
        if (__ballot(lane < vc && myv.ai == np.ai && myv.bi == np.bi))
            return contact_from_facet(pr, shfl_sup(myv, sel.i & 63), shfl_sup(myv, sel.j & 63),
                                      shfl_sup(myv, sel.k & 63), sel, out);
        const int ni = vc++;
        if (lane == ni) myv = np;
        // visible facets (each lane its own), then the horizon in index order
        const V fiq = shfl_v(myv.q, myf.i & 63);
        const bool vis = lane < fc && myf.valid && !(vdot(myf.n, vsub(np.q, fiq)) < 0.0f);
        u64 vmask = __ballot(vis);
        if (vis) myf.valid = false;
        int bc = 0;
        while (vmask) {
            const int f = __builtin_ctzll(vmask);
            vmask &= vmask - 1ull;
            const int fi = __shfl(myf.i, f), fj = __shfl(myf.j, f), fk = __shfl(myf.k, f);
            const int ef[3] = {fi, fj, fk}, et[3] = {fj, fk, fi};
#pragma unroll
            for (int q = 0; q < 3; ++q) {
                const u64 m = __ballot(lane < bc && bfrom == et[q] && bto == ef[q]);
                if (m) {
                    const int rev = __builtin_ctzll(m);
                    --bc;
                    const int lf = __shfl(bfrom, bc), lt = __shfl(bto, bc);
                    if (lane == rev) {
                        bfrom = lf;
                        bto = lt;
                    }
                } else {
                    if (bc == 32) return false;
                    if (lane == bc) {
                        bfrom = ef[q];
                        bto = et[q];
                    }
                    ++bc;
                }
            }
        }
        if (bc == 0) return false;
        const u64 inval = __ballot(lane < fc && !myf.valid);
        const int ninv = __popcll(inval);
        if (bc > 64 - fc + ninv) return false;
        // lane t < bc builds the facet of horizon edge t
        Facet nf = make_facet_q(shfl_v(myv.q, bfrom & 63), shfl_v(myv.q, bto & 63), shfl_v(myv.q, ni), bfrom, bto,
                                ni);
        if (__ballot(lane < bc && !nf.valid)) return false;
        nf.in_heap = nf.d <= upper;
        // edge t goes to the t-th freed slot, or after the last facet
        int t = -1;
        if (lane < fc && !myf.valid) {
            const int k = __popcll(inval & below);
            if (k < bc) t = k;
        } else if (lane >= fc) {
            const int k = ninv + (lane - fc);
            if (k < bc) t = k;
        }
        const Facet got = shfl_facet(nf, t >= 0 ? t : 0);
        if (t >= 0) myf = got;
        fc += bc > ninv ? bc - ninv : 0;
    }
    return false;
}
// GuPCMGjkEpa.cpp:396-515 with the polyhedral (no sphere) branches, on one
// lane. Returns 0 (no contact), 1 (contact in out) or 2 (EPA from the
// simplex sx[0, count)).
__device__ int gjk_lane(const GPair& pr, float cd, Pt& out, Sup (&sx)[4], int& count) {
    const float eps = 0.1f * smin(box_min_margin(pr.ha), box_min_margin(pr.hb));
    constexpr float progress = 1.0f - 0.000225f;
    count = 0;
    V closest = vdot(pr.a_in_b, pr.a_in_b) > 0.0f ? pr.a_in_b : V{1, 0, 0};
    float previous = kFltMax;
    for (int it = 0; it < 32; ++it) {
        const float dist = sqrtf(vdot(closest, closest));
        if (count > 0 && dist <= eps) return 2;
        if (dist <= 0.0f) return 0;
        const V n = vmul(closest, 1.0f / dist);
        const Sup s = support(pr, vneg(closest));
        const float proj = vdot(n, s.q);
        if (proj > cd) return 0;
        if (count > 0 && proj > dist * progress) {
            const Closest cur = closest_simplex(sx, count);
            return cur.mask && contact_from_gjk(pr, cur, out) ? 1 : 0;
        }
        // (the simplex array is indexed with constants only: it stays in
        // registers)
        bool dup = false;
#pragma unroll
        for (int i = 0; i < 4; ++i) dup |= i < count && s.ai == sx[i].ai && s.bi == sx[i].bi;
        if (dup) {
            if (proj > 0.f) {
                const Closest cur = closest_simplex(sx, count);
                return cur.mask && contact_from_gjk(pr, cur, out) && out.s <= cd ? 1 : 0;
            }
            return 0;
        }
        if (count == 4) return 2;
#pragma unroll
        for (int k = 0; k < 4; ++k) sx[k] = ssel(k == count, s, sx[k]);
        ++count;
        const Closest nx = closest_simplex(sx, count);
        if (!nx.mask) return 0;
        closest = nx.p;
        const float nd = sqrtf(nx.d2);
        // keep the vertices with weight, in order
        Sup kept[4] = {sx[0], sx[1], sx[2], sx[3]};
        int wr = 0;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const bool keep = r < count && (nx.mask & (1u << r)) && nx.w[r] > 1.0e-6f;
#pragma unroll
            for (int k = 0; k <= r; ++k) kept[k] = ssel(keep && k == wr, sx[r], kept[k]);
            wr += keep ? 1 : 0;
        }
        if (wr == 0 && count > 0) {
            wr = 1;
            kept[0] = sx[0];
        }
#pragma unroll
        for (int k = 0; k < 4; ++k) sx[k] = ssel(k < wr, kept[k], sx[k]);
        count = wr;
        if (nd >= previous) return 0;
        previous = nd;
    }
    return 0;
}
__device__ inline Pt shfl_pt(const Pt& x, int src) { return {shfl_v(x.a, src), shfl_v(x.b, src), shfl_v(x.n, src), __shfl(x.s, src)}; }
// GJK on lane 0; its result, or the EPA it hands over, shared by the wave.
__device__ bool gjk_epa(const GPair& pr, float cd, Pt& out) {
    Sup sx[4];
    int count = 0, r = 0;
    if (lane_id() == 0) r = gjk_lane(pr, cd, out, sx, count);
    r = __shfl(r, 0);
    if (r == 2) {
        count = __shfl(count, 0);
        Sup all[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) all[i] = shfl_sup(sx[i], 0);
        return epa_wave(pr, all, count, out);
    }
    out = shfl_pt(out, 0);
    return r == 1;
}
// GuPCMGjkEpa.cpp:527-566, box/box entry.
__device__ bool box_box_gjk_epa(const Box3& b0, const Box3& b1, float cd, Pt& out) {
    const V rel = grot_inv(b1.rot, vsub(b0.pos, b1.pos));
    const Q ib{-b1.rot.x, -b1.rot.y, -b1.rot.z, b1.rot.w};
    const Q r0 = b0.rot;
    const Q qa{ib.w * r0.x + r0.w * ib.x + ib.y * r0.z - ib.z * r0.y,
               ib.w * r0.y + r0.w * ib.y + ib.z * r0.x - ib.x * r0.z,
               ib.w * r0.z + r0.w * ib.z + ib.x * r0.y - ib.y * r0.x,
               ib.w * r0.w - ib.x * r0.x - ib.y * r0.y - ib.z * r0.z};
    const GPair pr{b0.half, b1.half, rel, qa};
    if (!gjk_epa(pr, cd, out)) return false;
    return out.s <= cd;
}

// ---------------- scene.cpp persistent manifolds ----------------
__device__ inline u32 c_count(const Cache& c) { return c.flags & 0xffu; }
__device__ inline bool c_init(const Cache& c) { return (c.flags >> 8) & 1u; }
__device__ inline void c_set(Cache& c, u32 count, bool init) { c.flags = (count & 0xffu) | (init ? 0x100u : 0u); }

// The manifold's four points are indexed with constants only (loops
// unrolled over the four slots, runtime slot numbers resolved by selects),
// so a cache copy stays in registers.
__device__ inline V vsel(bool c, V x, V y) { return {c ? x.x : y.x, c ? x.y : y.y, c ? x.z : y.z}; }
__device__ inline Pt psel(bool c, const Pt& x, const Pt& y) {
    return {vsel(c, x.a, y.a), vsel(c, x.b, y.b), vsel(c, x.n, y.n), c ? x.s : y.s};
}
// (selects on the scalar fields: a whole-struct copy from a selected slot
// would become a pointer select, which keeps the slots in memory)
__device__ inline Pt pick4(const Pt (&p)[4], u32 i) {
    return psel(i == 3u, p[3], psel(i == 2u, p[2], psel(i == 1u, p[1], p[0])));
}
// (every slot written through a select: conditional stores to the slots
// would be merged into one store through a selected pointer)
__device__ inline void put4(Pt (&p)[4], u32 i, const Pt& v) {
#pragma unroll
    for (u32 k = 0; k < 4; ++k) p[k] = psel(i == k, v, p[k]);
}

// scene.cpp:3432-3465 (newest slot first; a dropped point takes the last
// slot's point, which has been refreshed already)
__device__ void refresh(Cache& c, V pos_ab, Q rot_ab, float brk) {
    const float b2 = brk * brk;
    u32 count = c_count(c);
#pragma unroll
    for (int i = 3; i >= 0; --i) {
        if (static_cast<u32>(i) >= count) continue;
        Pt& p = c.pts[i];
        const V r = srot(rot_ab, p.a);
        const V ainb{pos_ab.x + r.x, pos_ab.y + r.y, pos_ab.z + r.z};
        const V v{ainb.x - p.b.x, ainb.y - p.b.y, ainb.z - p.b.z};
        const float d = v.x * p.n.x + v.y * p.n.y + v.z * p.n.z;
        const V pd{v.x - p.n.x * d, v.y - p.n.y * d, v.z - p.n.z * d};
        const float t2 = pd.x * pd.x + pd.y * pd.y + pd.z * pd.z;
        if (t2 > b2)
            c.pts[i] = pick4(c.pts, --count);
        else
            p.s = d;
    }
    c_set(c, count, c_init(c));
}

__device__ float seg_sq(V p, V a, V b) {
    const V ab = vsub(b, a), ap = vsub(p, a);
    const float den = vdot(ab, ab);
    const float t = den == 0.0f ? 0.0f : sclamp(vdot(ap, ab) / den, 0.0f, 1.0f);
    const V d{ap.x - t * ab.x, ap.y - t * ab.y, ap.z - t * ab.z};
    return vdot(d, d);
}
__device__ float tri_sq(V p, V a, V b, V c) {
    const V ab = vsub(b, a), ac = vsub(c, a);
    const V ap = vsub(p, a), bp = vsub(p, b), cp = vsub(p, c);
    const float d1 = vdot(ab, ap), d2 = vdot(ac, ap);
    const float d3 = vdot(ab, bp), d4 = vdot(ac, bp);
    const float d5 = vdot(ab, cp), d6 = vdot(ac, cp);
    if (d1 < 0.0f && d2 < 0.0f) return vdot(ap, ap);
    if (d3 >= 0.0f && d3 >= d4) return vdot(bp, bp);
    if (d6 >= 0.0f && d6 >= d5) return vdot(cp, cp);
    const float vc = d1 * d4 - d3 * d2;
    if (vc < 0.0f && d1 >= 0.0f && d3 < 0.0f) return seg_sq(p, a, b);
    const float va = d3 * d6 - d5 * d4;
    if (va < 0.0f && d4 >= d3 && d5 >= d6) return seg_sq(p, b, c);
    const float vb = d5 * d2 - d1 * d6;
    if (vb < 0.0f && d2 >= 0.0f && d6 < 0.0f) return seg_sq(p, a, c);
    const V n{ab.y * ac.z - ab.z * ac.y, ab.z * ac.x - ab.x * ac.z, ab.x * ac.y - ab.y * ac.x};
    const float den = vdot(n, n);
    if (den == 0.0f) return 0.0f;
    const float pr = vdot(n, ap);
    return pr * pr / den;
}
// scene.cpp:3317-3431 (the five points indexed with constants; runtime
// picks resolved by selects)
__device__ inline Pt pick5(const Pt (&p)[5], int i) {
    return psel(i == 4, p[4], psel(i == 3, p[3], psel(i == 2, p[2], psel(i == 1, p[1], p[0]))));
}
__device__ void reduce_five(Cache& cache, const Pt& nc) {
    const Pt cd[5] = {cache.pts[0], cache.pts[1], cache.pts[2], cache.pts[3], nc};
    unsigned chosen = 0;
    int deepest = 4;
    float ds = cd[4].s;
#pragma unroll
    for (int i = 0; i < 4; ++i)
        if (cd[i].s < ds) {
            deepest = i;
            ds = cd[i].s;
        }
    chosen |= 1u << deepest;
    const V bd = pick5(cd, deepest).b;
    int far = 0;
    V d0 = vsub(cd[0].b, bd);
    float largest = vdot(d0, d0);
#pragma unroll
    for (int i = 1; i < 5; ++i)
        if (!(chosen & (1u << i))) {
            const V d = vsub(cd[i].b, bd);
            const float v = vdot(d, d);
            if (v > largest) {
                largest = v;
                far = i;
            }
        }
    chosen |= 1u << far;
    const V bf = pick5(cd, far).b;
    largest = -kFltMax;
    int third = 0;
#pragma unroll
    for (int i = 0; i < 5; ++i)
        if (!(chosen & (1u << i))) {
            const float v = seg_sq(cd[i].b, bd, bf);
            if (v > largest) {
                largest = v;
                third = i;
            }
        }
    chosen |= 1u << third;
    const V bt = pick5(cd, third).b;
    largest = -kFltMax;
    int fourth = third;
#pragma unroll
    for (int i = 0; i < 5; ++i)
        if (!(chosen & (1u << i))) {
            const float v = tri_sq(cd[i].b, bd, bf, bt);
            if (v > largest) {
                largest = v;
                fourth = i;
            }
        }
    if (chosen & (1u << fourth)) {
        cache.pts[0] = pick5(cd, deepest);
        cache.pts[1] = pick5(cd, far);
        cache.pts[2] = pick5(cd, third);
        c_set(cache, 3, c_init(cache));
        return;
    }
    chosen |= 1u << fourth;
    cache.pts[0] = pick5(cd, deepest);
    cache.pts[1] = pick5(cd, far);
    cache.pts[2] = pick5(cd, third);
    cache.pts[3] = pick5(cd, fourth);
    c_set(cache, 4, c_init(cache));
    const unsigned rest = ~chosen & 0x1fu;
    if (!rest) return;
    const int uns = __ffs(rest) - 1;
    const Pt pu = pick5(cd, uns);
    float closest = kFltMax;
    int rep = 0;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const V d = vsub(cache.pts[i].a, pu.a);
        const float dd = vdot(d, d);
        if (dd < closest) {
            closest = dd;
            rep = i;
        }
    }
    if (pu.s < pick4(cache.pts, static_cast<u32>(rep)).s) put4(cache.pts, static_cast<u32>(rep), pu);
}

__device__ inline float pcm_margin(V h) { return smin(0.15f * smin(smin(h.x, h.y), h.z), 0.15f); }

// Per-pair flags word written for every pair:
enum : u32 {
    kRegenerated = 1u,
    kGjk = 2u,          // SAT overlapped with no clip witness: GJK/EPA continuation ran
    kInvalid = 4u,      // reference would throw (invalid normal / non-finite contact)
};

// One sphere pair of the box/box list (scene.cpp:4514-4537): first / second
// in Gu order; a box first takes the reversed normal, so the normal always
// points toward the first shape. A non-finite contact is flagged invalid as
// in the box path (the reference raises overflow_error there).
__device__ inline u32 sphere_pair(const Box3& f, bool fs, const Box3& s, bool ss, float cd, float* out,
                                  u32& flags) {
    SPt o{};
    u32 e;
    if (fs && ss) {
        e = sphere_sphere(f.pos, f.half.x, s.pos, s.half.x, cd, o);
    } else {
        const Box3& sph = fs ? f : s;
        const Box3& box = fs ? s : f;
        e = sphere_box(sph.pos, sph.half.x, box.pos, box.rot, box.half, cd, o);
        if (e && !fs) o.n = vneg(o.n);
    }
    if (e) {
        const float v[7] = {o.p.x, o.p.y, o.p.z, o.n.x, o.n.y, o.n.z, o.s};
        for (int k = 0; k < 7; ++k) {
            out[k] = v[k];
            if (!isfinite(v[k])) flags |= kInvalid;
        }
    }
    return e;
}


// Manifold emission (scene.cpp:3598-3653). direct: contacts came from SAT
// clipping this frame; normal source as in the reference.
__device__ u32 emit_box_box(const Cache& cache, const Box3& second, float cd, bool direct, bool gjk, V gjk_n,
                            float* out, u32& flags) {
    const u32 count = c_count(cache);
    if (count == 0) return 0;
    V nb{0, 0, 0};
    if (gjk)
        nb = gjk_n;
    else if (direct)
        nb = cache.pts[0].n;
    else
#pragma unroll
        for (u32 i = 0; i < 4; ++i)
            if (i < count) {
                nb.x += cache.pts[i].n.x;
                nb.y += cache.pts[i].n.y;
                nb.z += cache.pts[i].n.z;
            }
    const float n2 = nb.x * nb.x + nb.y * nb.y + nb.z * nb.z;
    if (!(n2 > 0.0f) || !isfinite(n2)) {
        flags |= kInvalid;
        return 0;
    }
    const float il = 1.0f / sqrtf(n2);
    const V nn{nb.x * il, nb.y * il, nb.z * il};
    const V normal = srot(second.rot, nn);
    u32 e = 0;
#pragma unroll
    for (u32 i = 0; i < 4; ++i) {
        if (i >= count) continue;
        const Pt& p = cache.pts[i];
        if (!direct && p.s > cd) continue;
        const V r = srot(second.rot, p.b);
        float* o = out + 7 * e;
        o[0] = r.x + second.pos.x;
        o[1] = r.y + second.pos.y;
        o[2] = r.z + second.pos.z;
        o[3] = normal.x;
        o[4] = normal.y;
        o[5] = normal.z;
        o[6] = p.s;
        if (!isfinite(o[0]) || !isfinite(o[1]) || !isfinite(o[2]) || !isfinite(o[3]) || !isfinite(o[6]))
            flags |= kInvalid;
        ++e;
    }
    return e;
}

struct PairArgs {
    u32 n_pairs;
    const u32* first;   // slot of the Gu first shape
    const u32* second;  // slot of the Gu second shape (normal frame)
    const u32* cache_index;  // per pair, or null (= pair index)
    const float* pos;
    const float* quat;
    const float* half;
    float cd;
    Cache* cache;
    float* contacts;  // [n_pairs][4][7]
    u32* counts;
    u32* flags;
    u32* gjk_list;    // pairs deferred to the continuation
    u32* gjk_count;
    const u32* d_n = nullptr;          // device pair count (at most n_pairs), or null
    const Cache* cache_src = nullptr;  // read caches from here (double-buffered), or null
    const u32* src_index = nullptr;    // per pair: cache_src record, kNoSource = empty
    const uint8_t* sphere = nullptr;   // per slot: non-zero for a sphere (radius in half[3 s]), or null
};

// scene.cpp:3466-3654 up to the zero-witness continuation.
__global__ __launch_bounds__(kWitnessThreads) void k_pcm_box_box(PairArgs A) {
    __shared__ float wit_lds[3 * kMaxCand * kWitnessThreads];
    const u32 t = blockIdx.x * blockDim.x + threadIdx.x;
    const u32 n = A.d_n ? min(*A.d_n, A.n_pairs) : A.n_pairs;
    if (t >= n) return;
    const Box3 f = load_box(A.pos, A.quat, A.half, A.first[t]);
    const Box3 s = load_box(A.pos, A.quat, A.half, A.second[t]);
    Cache& cref = A.cache[A.cache_index ? A.cache_index[t] : t];
    if (A.sphere && (A.sphere[A.first[t]] || A.sphere[A.second[t]])) {
        u32 sflags = 0;
        A.counts[t] = sphere_pair(f, A.sphere[A.first[t]] != 0, s, A.sphere[A.second[t]] != 0, A.cd,
                                  A.contacts + 28ull * t, sflags);
        A.flags[t] = sflags;
        cref = Cache{};
        return;
    }
    Cache cache{};
    if (!A.cache_src)
        cache = cref;
    else if (A.src_index[t] != kNoSource)
        cache = A.cache_src[A.src_index[t]];
    u32 flags = 0;
    const V delta{f.pos.x - s.pos.x, f.pos.y - s.pos.y, f.pos.z - s.pos.z};
    const V rel = srot_inv(s.rot, delta);
    const Q rot = qmul(Q{-s.rot.x, -s.rot.y, -s.rot.z, s.rot.w}, f.rot);
    const float mm = smin(pcm_margin(f.half), pcm_margin(s.half));
    const u32 initial = c_count(cache);
    refresh(cache, rel, rot, mm * 0.8f);
    const u32 cnt = c_count(cache);
    const bool lost = cnt != initial;
    // position / orientation thresholds by point count (0..4)
    const float thc = cnt == 0u ? 0.5f : cnt == 1u ? 0.125f : cnt == 2u ? 0.25f : 0.375f;
    const float qthc = cnt == 0u ? 0.9998f : 0.9999f;
    const bool init = c_init(cache);
    const float td = init ? smax(smax(fabsf(rel.x - cache.rel_pos.x), fabsf(rel.y - cache.rel_pos.y)),
                                 fabsf(rel.z - cache.rel_pos.z))
                          : __builtin_inff();
    const float qa = f.rot.x * cache.orient_a.x + f.rot.y * cache.orient_a.y + f.rot.z * cache.orient_a.z +
                     f.rot.w * cache.orient_a.w;
    const float qb = s.rot.x * cache.orient_b.x + s.rot.y * cache.orient_b.y + s.rot.z * cache.orient_b.z +
                     s.rot.w * cache.orient_b.w;
    const float tp = mm * thc;
    bool regen = !init || lost || td > tp || qa < qthc || qb < qthc;
    if (!regen) {
        auto travel = [](float d, V e) {
            const float ang = d < 1.0f ? acosf(smax(-1.0f, d)) : 0.0f;
            return ang * sqrtf(e.x * e.x + e.y * e.y + e.z * e.z);
        };
        regen = travel(qa, f.half) > tp || travel(qb, s.half) > tp;
    }
    bool direct = false;
    if (regen) {
        flags |= kRegenerated;
        cache.rel_pos = rel;
        cache.rel_rot = rot;
        cache.orient_a = f.rot;
        cache.orient_b = s.rot;
        c_set(cache, c_count(cache), true);
        Witnesses cands;
        cands.lds = wit_lds + threadIdx.x;
        cands.stride = blockDim.x;
        const bool overlapping = box_box_candidates(f, s, A.cd, cands);
        const int count = overlapping ? cands.count : 0;
        if (!overlapping) {
            c_set(cache, 0, true);
        } else if (count == 0) {
            // finished by k_pcm_gjk, which owns this pair's outputs
            flags |= kGjk;
            cref = cache;
            A.flags[t] = flags;
            A.counts[t] = 0;
            const u32 slot = atomicAdd(A.gjk_count, 1u);
            A.gjk_list[slot] = t;
            return;
        } else if (count <= 4) {
#pragma unroll
            for (int i = 0; i < 4; ++i)
                if (i < count) cache.pts[i] = cands.pt(i);
            c_set(cache, static_cast<u32>(count), true);
        } else {
            int sel[4];
            batch_selection(cands, count, 1.0f, sel);
#pragma unroll
            for (int i = 0; i < 4; ++i) cache.pts[i] = cands.pt(sel[i]);
            c_set(cache, 4, true);
        }
        direct = count > 0;
    }
    const u32 e = emit_box_box(cache, s, A.cd, direct, false, V{0, 0, 0}, A.contacts + 28ull * t, flags);
    cref = cache;
    A.counts[t] = e;
    A.flags[t] = flags;
}

// The zero-witness continuation (scene.cpp:3543-3593), one wave per
// deferred pair (one wave per workgroup): every lane runs GJK and the
// wave-parallel EPA with the same values; lane 0 writes the outputs.
__global__ __launch_bounds__(64) void k_pcm_gjk(PairArgs A) {
    const u32 n = min(*A.gjk_count, A.n_pairs);
    const bool writer = threadIdx.x == 0;
    for (u32 w = blockIdx.x; w < n; w += gridDim.x) {
        const u32 t = A.gjk_list[w];
        const Box3 f = load_box(A.pos, A.quat, A.half, A.first[t]);
        const Box3 s = load_box(A.pos, A.quat, A.half, A.second[t]);
        Pt wit;
        const bool hit = box_box_gjk_epa(f, s, A.cd, wit);
        if (!writer) continue;
        Cache& cref = A.cache[A.cache_index ? A.cache_index[t] : t];
        Cache cache = cref;
        u32 flags = A.flags[t];
        const float mm = smin(pcm_margin(f.half), pcm_margin(s.half));
        if (!hit) {
            c_set(cache, 0, true);
            cref = cache;
            A.counts[t] = 0;
            continue;
        }
        const float rb = mm * 0.05f;
        const float rsq = rb * rb;
        bool replaced = false;
        const u32 cnt = c_count(cache);
#pragma unroll
        for (u32 i = 0; i < 4; ++i) {
            const Pt& o = cache.pts[i];
            const V da = vsub(o.a, wit.a), db = vsub(o.b, wit.b);
            const bool hit = !replaced && i < cnt && smin(vdot(da, da), vdot(db, db)) < rsq;
            cache.pts[i] = psel(hit, wit, cache.pts[i]);
            replaced |= hit;
        }
        if (!replaced) {
            if (cnt < 4) {
                put4(cache.pts, cnt, wit);
                c_set(cache, cnt + 1, true);
            } else {
                reduce_five(cache, wit);
            }
        }
        const u32 e = emit_box_box(cache, s, A.cd, false, true, wit.n, A.contacts + 28ull * t, flags);
        cref = cache;
        A.counts[t] = e;
        A.flags[t] = flags;
    }
}

struct PlaneArgs {
    u32 n;
    const u32* box;  // slot per plane pair
    const u32* cache_index;
    const float* pos;
    const float* quat;
    const float* half;
    V plane_pos;
    float cd;
    Cache* cache;
    float* contacts;
    u32* counts;
    u32* flags;
    const uint8_t* sphere = nullptr;  // per slot: non-zero for a sphere (radius in half[3 s]), or null
};

// scene.cpp:3774-3955 (box geometry), followed by the scene's flip of the
// Gu plane-first normal (scene.cpp:4652-4657): published normal is +y.
__global__ __launch_bounds__(128) void k_pcm_plane_box(PlaneArgs A) {
    const u32 t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= A.n) return;
    const Box3 bx = load_box(A.pos, A.quat, A.half, A.box ? A.box[t] : t);
    Cache& cref = A.cache[A.cache_index ? A.cache_index[t] : t];
    if (A.sphere && A.sphere[A.box ? A.box[t] : t]) {
        // a sphere on the plane (scene.cpp:4502-4513): one stateless contact
        SPt o{};
        u32 sflags = 0;
        const u32 e = sphere_plane(bx.pos, bx.half.x, A.plane_pos.y, A.cd, o);
        if (e) {
            const float v[7] = {o.p.x, o.p.y, o.p.z, o.n.x, o.n.y, o.n.z, o.s};
            float* out = A.contacts + 28ull * t;
            for (int k = 0; k < 7; ++k) {
                out[k] = v[k];
                if (!isfinite(v[k])) sflags |= kInvalid;
            }
        }
        A.counts[t] = e;
        A.flags[t] = sflags;
        cref = Cache{};
        return;
    }
    Cache cache = cref;
    const V rel{bx.pos.x - A.plane_pos.x, bx.pos.y - A.plane_pos.y, bx.pos.z - A.plane_pos.z};
    const V h = bx.half;
    const float margin = smin(0.15f * smin(smin(h.x, h.y), h.z), 0.15f);
    const float brk = margin * 0.2f;
    const u32 prev = c_count(cache);
    refresh(cache, rel, bx.rot, brk);
    const bool init = c_init(cache);
    const float md = init ? smax(smax(fabsf(rel.x - cache.rel_pos.x), fabsf(rel.y - cache.rel_pos.y)),
                                 fabsf(rel.z - cache.rel_pos.z))
                          : __builtin_inff();
    const Q q = bx.rot;
    const Q old = cache.rel_rot;
    const float rd = q.x * old.x + q.y * old.y + q.z * old.z + q.w * old.w;
    const bool regen = !init || c_count(cache) != prev || md > brk || rd < 0.99996f;
    u32 flags = regen ? kRegenerated : 0u;
    if (regen) {
        cache.rel_pos = rel;
        cache.rel_rot = q;
        // The corners below the contact distance, in corner order, are the
        // candidates; every selection loop below visits them in that order
        // (unrolled over the eight corners, so nothing leaves registers).
        V loc[8], ip[8];
        unsigned below = 0;
#pragma unroll
        for (unsigned corner = 0; corner < 8u; ++corner) {
            loc[corner] = {(corner & 4u) ? -h.x : h.x, (corner & 2u) ? -h.y : h.y, (corner & 1u) ? -h.z : h.z};
            const V r = srot(q, loc[corner]);
            ip[corner] = {rel.x + r.x, rel.y + r.y, rel.z + r.z};
            if (!(ip[corner].y >= A.cd)) below |= 1u << corner;
        }
        auto cand = [&](unsigned corner) -> Pt {
            return {loc[corner], {ip[corner].x, 0.0f, ip[corner].z}, {0.0f, 1.0f, 0.0f}, ip[corner].y};
        };
        auto cand_of = [&](unsigned id) -> Pt {
            Pt r = cand(0);
#pragma unroll
            for (unsigned corner = 1; corner < 8u; ++corner) r = psel(id == corner, cand(corner), r);
            return r;
        };
        const int n = __popc(below);
        if (n <= 4) {
            u32 k = 0;
#pragma unroll
            for (unsigned corner = 0; corner < 8u; ++corner)
                if (below & (1u << corner)) put4(cache.pts, k++, cand(corner));
            c_set(cache, static_cast<u32>(n), true);
        } else {
            // deepest (first on ties)
            unsigned first = 8;
            float sf = 0.0f;
            V bf{0, 0, 0};
#pragma unroll
            for (unsigned corner = 0; corner < 8u; ++corner)
                if ((below & (1u << corner)) && (first == 8 || ip[corner].y < sf)) {
                    first = corner;
                    sf = ip[corner].y;
                    bf = {ip[corner].x, 0.0f, ip[corner].z};
                }
            // farthest from it (the first candidate counts, even if it is the deepest)
            unsigned far = 8;
            float maxd = 0.0f;
            V bfar{0, 0, 0};
#pragma unroll
            for (unsigned corner = 0; corner < 8u; ++corner)
                if (below & (1u << corner)) {
                    const V b{ip[corner].x, 0.0f, ip[corner].z};
                    const V d = vsub(b, bf);
                    const float dd = vdot(d, d);
                    if (far == 8 || dd > maxd) {
                        maxd = dd;
                        far = corner;
                        bfar = b;
                    }
                }
            unsigned chosen = (1u << first) | (1u << far);
            const V segv = vsub(bfar, bf);
            V side{-segv.z, 0.0f, segv.x};
            const float ns = vdot(side, side);
            if (ns > 0.0f) {
                const float inv = 1.0f / sqrtf(ns);
                side = {side.x * inv, 0.0f, side.z * inv};
            } else {
                side = {0.0f, 1.0f, 0.0f};
            }
            // the lowest candidate stands in when no side value compares
            const unsigned lowest = static_cast<unsigned>(__ffs(below) - 1);
            float largest = -kFltMax, smallest = kFltMax;
            unsigned third = lowest, fourth = lowest;
#pragma unroll
            for (unsigned corner = 0; corner < 8u; ++corner) {
                if (!(below & (1u << corner)) || (chosen & (1u << corner))) continue;
                const V b{ip[corner].x, 0.0f, ip[corner].z};
                const float d = vdot(vsub(b, bf), side);
                if (d > largest) {
                    largest = d;
                    third = corner;
                }
                if (d < smallest) {
                    smallest = d;
                    fourth = corner;
                }
            }
            chosen |= 1u << third;
            if (smallest * largest > 0.0f) {
                largest = -kFltMax;
#pragma unroll
                for (unsigned corner = 0; corner < 8u; ++corner) {
                    if (!(below & (1u << corner)) || (chosen & (1u << corner))) continue;
                    const V b{ip[corner].x, 0.0f, ip[corner].z};
                    const float d = vdot(vsub(b, bf), side);
                    if (d > largest) {
                        largest = d;
                        fourth = corner;
                    }
                }
            }
            chosen |= 1u << fourth;
            unsigned idx[4] = {first, far, third, fourth};
            V centers[4];
            float sidx[4];
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const Pt ck = cand_of(idx[k]);
                centers[k] = ck.b;
                sidx[k] = ck.s;
            }
            // every other candidate replaces the nearest selected one if deeper
#pragma unroll
            for (unsigned corner = 0; corner < 8u; ++corner) {
                if (!(below & (1u << corner)) || (chosen & (1u << corner))) continue;
                const V b{ip[corner].x, 0.0f, ip[corner].z};
                float nearest = kFltMax;
                int slot = 0;
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    const V d = vsub(b, centers[j]);
                    const float dd = vdot(d, d);
                    if (dd < nearest) {
                        nearest = dd;
                        slot = j;
                    }
                }
#pragma unroll
                for (int j = 0; j < 4; ++j)
                    if (j == slot && ip[corner].y < sidx[j]) {
                        idx[j] = corner;
                        sidx[j] = ip[corner].y;
                    }
            }
#pragma unroll
            for (int k = 0; k < 4; ++k) cache.pts[k] = cand_of(idx[k]);
            c_set(cache, 4, true);
        }
    }
    u32 e = 0;
    float* out = A.contacts + 28ull * t;
    const u32 count = c_count(cache);
#pragma unroll
    for (u32 k = 0; k < 4; ++k) {
        if (k >= count) continue;
        const Pt& pt = cache.pts[k];
        if (pt.s > A.cd) continue;
        float* o = out + 7 * e;
        o[0] = A.plane_pos.x + pt.b.x;
        o[1] = A.plane_pos.y;
        o[2] = A.plane_pos.z + pt.b.z;
        o[3] = 0.0f;
        o[4] = 1.0f;
        o[5] = 0.0f;
        o[6] = pt.s;
        if (!isfinite(o[0]) || !isfinite(o[2]) || !isfinite(o[6])) flags |= kInvalid;
        ++e;
    }
    cref = cache;
    A.counts[t] = e;
    A.flags[t] = flags;
}

__global__ void k_zero1(u32* p) { *p = 0u; }

// Stateless parity entries (one thread per pose pair).
__global__ __launch_bounds__(kWitnessThreads) void k_candidates(u32 n, const float* a, const float* b, float cd,
                                                                float* out, u32* counts, u32* overlapping,
                                                                u32* selection) {
    __shared__ float wit_lds[3 * kMaxCand * kWitnessThreads];
    const u32 t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const float* pa = a + 10ull * t;
    const float* pb = b + 10ull * t;
    const Box3 ba{{pa[0], pa[1], pa[2]}, {pa[3], pa[4], pa[5]}, {pa[6], pa[7], pa[8], pa[9]}};
    const Box3 bb{{pb[0], pb[1], pb[2]}, {pb[3], pb[4], pb[5]}, {pb[6], pb[7], pb[8], pb[9]}};
    Witnesses c;
    c.lds = wit_lds + threadIdx.x;
    c.stride = blockDim.x;
    const bool ov = box_box_candidates(ba, bb, cd, c);
    const int cnt = ov ? c.count : 0;
    overlapping[t] = ov ? 1u : 0u;
    counts[t] = static_cast<u32>(cnt);
    float* o = out + 160ull * t;
    for (int i = 0; i < cnt; ++i) {
        const Pt w = c.pt(i);
        o[10 * i + 0] = w.a.x;
        o[10 * i + 1] = w.a.y;
        o[10 * i + 2] = w.a.z;
        o[10 * i + 3] = w.b.x;
        o[10 * i + 4] = w.b.y;
        o[10 * i + 5] = w.b.z;
        o[10 * i + 6] = w.n.x;
        o[10 * i + 7] = w.n.y;
        o[10 * i + 8] = w.n.z;
        o[10 * i + 9] = w.s;
    }
    int sel[4] = {255, 255, 255, 255};
    if (cnt > 4) batch_selection(c, cnt, 1.0f, sel);
    for (int k = 0; k < 4; ++k) selection[4ull * t + k] = static_cast<u32>(sel[k]);
}

// One wave per pair (one wave per workgroup), as in the pipeline.
__global__ __launch_bounds__(64) void k_gjk_epa(u32 n, const float* a, const float* b, float cd, float* out,
                                                u32* ok) {
    const u32 t = blockIdx.x;
    if (t >= n) return;
    const float* pa = a + 10ull * t;
    const float* pb = b + 10ull * t;
    const Box3 ba{{pa[0], pa[1], pa[2]}, {pa[3], pa[4], pa[5]}, {pa[6], pa[7], pa[8], pa[9]}};
    const Box3 bb{{pb[0], pb[1], pb[2]}, {pb[3], pb[4], pb[5]}, {pb[6], pb[7], pb[8], pb[9]}};
    Pt c{{0, 0, 0}, {0, 0, 0}, {0, 0, 0}, 0.0f};
    const bool r = box_box_gjk_epa(ba, bb, cd, c);
    if (threadIdx.x != 0) return;
    ok[t] = r ? 1u : 0u;
    float* o = out + 10ull * t;
    o[0] = c.a.x; o[1] = c.a.y; o[2] = c.a.z;
    o[3] = c.b.x; o[4] = c.b.y; o[5] = c.b.z;
    o[6] = c.n.x; o[7] = c.n.y; o[8] = c.n.z;
    o[9] = c.s;
}

}  // namespace fyzz_hip::pcm

using namespace fyzz_hip;

extern "C" {

uint32_t fyzz_hip_pcm_cache_bytes(void) { return sizeof(pcm::Cache); }

int fyzz_hip_pcm_box_box(fyzz_hip_context* ctx, uint32_t n_pairs, const uint32_t* first,
                         const uint32_t* second, const uint32_t* cache_index, const float* pos,
                         const float* quat, const float* half, float contact_distance, void* cache,
                         float* contacts, uint32_t* counts, uint32_t* flags, uint32_t* scratch) {
    return guarded(ctx, [&] {
        if (!n_pairs) return;
        FYZZ_REQUIRE(first && second && pos && quat && half && cache && contacts && counts && flags && scratch,
                     kInvalidArgument, "pcm_box_box: null device pointer");
        pcm::PairArgs A{n_pairs, first, second, cache_index, pos, quat, half, contact_distance,
                        static_cast<pcm::Cache*>(cache), contacts, counts, flags, scratch + 1, scratch};
        hipStream_t st = ctx->stream;
        hipLaunchKernelGGL(pcm::k_zero1, dim3(1), dim3(1), 0, st, scratch);
        ProfScope ps(&ctx->prof, "pcm.box_box", st);
        hipLaunchKernelGGL(pcm::k_pcm_box_box, dim3(grid_for(n_pairs, 128)), dim3(128), 0, st, A);
        hipLaunchKernelGGL(pcm::k_pcm_gjk, dim3(std::min<u32>(std::max<u32>(n_pairs, 1u), 8192u)), dim3(64), 0, st, A);
        FYZZ_HIP_CHECK(hipGetLastError());
    });
}

int fyzz_hip_pcm_plane_box(fyzz_hip_context* ctx, uint32_t n, const uint32_t* box_slot,
                           const uint32_t* cache_index, const float* pos, const float* quat,
                           const float* half, const float* plane_position, float contact_distance,
                           void* cache, float* contacts, uint32_t* counts, uint32_t* flags) {
    return guarded(ctx, [&] {
        if (!n) return;
        FYZZ_REQUIRE(box_slot && pos && quat && half && plane_position && cache && contacts && counts && flags,
                     kInvalidArgument, "pcm_plane_box: null pointer");
        pcm::PlaneArgs A{n, box_slot, cache_index, pos, quat, half,
                         {plane_position[0], plane_position[1], plane_position[2]}, contact_distance,
                         static_cast<pcm::Cache*>(cache), contacts, counts, flags};
        ProfScope ps(&ctx->prof, "pcm.plane_box", ctx->stream);
        hipLaunchKernelGGL(pcm::k_pcm_plane_box, dim3(grid_for(n, 128)), dim3(128), 0, ctx->stream, A);
        FYZZ_HIP_CHECK(hipGetLastError());
    });
}

}  // extern "C"

namespace fyzz_hip::pcm {

void box_box_list(fyzz_hip_context* ctx, const ListPairs& L, const float* pos, const float* quat,
                  const float* half, float cd, float* contacts, u32* counts, u32* flags, u32* scratch,
                  const uint8_t* sphere) {
    if (!L.cap) return;
    PairArgs A{L.cap, L.first, L.second, nullptr, pos, quat, half, cd, static_cast<Cache*>(L.cache_dst),
               contacts, counts, flags, scratch + 1, scratch};
    A.d_n = L.d_n;
    A.sphere = sphere;
    A.cache_src = static_cast<const Cache*>(L.cache_src);
    A.src_index = L.src_index;
    hipStream_t st = ctx->stream;
    hipLaunchKernelGGL(k_zero1, dim3(1), dim3(1), 0, st, scratch);
    ProfScope ps(&ctx->prof, "pcm.box_box", st);
    hipLaunchKernelGGL(k_pcm_box_box, dim3(grid_for(L.cap, 128)), dim3(128), 0, st, A);
    hipLaunchKernelGGL(k_pcm_gjk, dim3(std::min<u32>(std::max<u32>(L.cap, 1u), 8192u)), dim3(64), 0, st, A);
    FYZZ_HIP_CHECK(hipGetLastError());
}

void plane_all(fyzz_hip_context* ctx, u32 n, const float* pos, const float* quat, const float* half,
               const float plane_position[3], float cd, void* cache, float* contacts, u32* counts, u32* flags,
               const uint8_t* sphere) {
    if (!n) return;
    PlaneArgs A{n, nullptr, nullptr, pos, quat, half, {plane_position[0], plane_position[1], plane_position[2]}, cd,
                static_cast<Cache*>(cache), contacts, counts, flags};
    A.sphere = sphere;
    ProfScope ps(&ctx->prof, "pcm.plane_box", ctx->stream);
    hipLaunchKernelGGL(k_pcm_plane_box, dim3(grid_for(n, 128)), dim3(128), 0, ctx->stream, A);
    FYZZ_HIP_CHECK(hipGetLastError());
}

}  // namespace fyzz_hip::pcm

extern "C" {

int fyzz_hip_pcm_candidates(fyzz_hip_context* ctx, uint32_t n, const float* a, const float* b,
                            float contact_distance, float* out, uint32_t* counts, uint32_t* overlapping,
                            uint32_t* selection) {
    return guarded(ctx, [&] {
        if (!n) return;
        FYZZ_REQUIRE(a && b && out && counts && overlapping && selection, kInvalidArgument,
                     "pcm_candidates: null pointer");
        hipLaunchKernelGGL(pcm::k_candidates, dim3(grid_for(n, 64)), dim3(64), 0, ctx->stream, n, a, b,
                           contact_distance, out, counts, overlapping, selection);
        FYZZ_HIP_CHECK(hipGetLastError());
    });
}

int fyzz_hip_pcm_gjk_epa(fyzz_hip_context* ctx, uint32_t n, const float* a, const float* b,
                         float contact_distance, float* out, uint32_t* ok) {
    return guarded(ctx, [&] {
        if (!n) return;
        FYZZ_REQUIRE(a && b && out && ok, kInvalidArgument, "pcm_gjk_epa: null pointer");
        hipLaunchKernelGGL(pcm::k_gjk_epa, dim3(n), dim3(64), 0, ctx->stream, n, a, b,
                           contact_distance, out, ok);
        FYZZ_HIP_CHECK(hipGetLastError());
    });
}

}  // extern "C"
