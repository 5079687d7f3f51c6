// Oracle PCM narrowphase restatement: box/box SAT + incident-face clipping,
// four-point batch selection, GJK/EPA continuation, the persistent
// manifold cache (refresh, regeneration thresholds, five-point reduction)
// and plane/box corner manifolds. FP32, operation order of the reference.
// Test infrastructure only.
//
// Two quaternion rotations appear in the reference with different
// association:  scene.cpp:57-66 computes (v + w t) + u x t, while the Gu
// translation units (GuPCMContactBoxBox.cpp:33-37, GuPCMGjkEpa.cpp:34-38)
// compute v + (t w + u x t). Both are restated as written.
#include <algorithm>
#include <atomic>
#include <array>
#include <cmath>
#include <cstdint>
#include <limits>

#include "orc.h"
#include "orc_math.h"

namespace {
using V = orc::V3<float>;
using Q = orc::Q4<float>;
using orc::cross;
using orc::dot;
using orc::parallel_ranges;

V add(V a, V b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
V sub(V a, V b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
V mul(V a, float f) { return {a.x * f, a.y * f, a.z * f}; }
V neg(V a) { return {-a.x, -a.y, -a.z}; }
float len2(V v) { return dot(v, v); }
float comp(V v, unsigned i) { return i == 0 ? v.x : (i == 1 ? v.y : v.z); }

// Gu-file rotation: v + (t w + u x t).
V gu_rotate(Q q, V v) {
    const V u{q.x, q.y, q.z};
    const V t = mul(cross(u, v), 2.0f);
    return add(v, add(mul(t, q.w), cross(u, t)));
}
V gu_inverse_rotate(Q q, V v) {
    const V u{-q.x, -q.y, -q.z};
    const V t = mul(cross(u, v), 2.0f);
    return add(v, add(mul(t, q.w), cross(u, t)));
}

V load3(const float* p) { return {p[0], p[1], p[2]}; }
Q load4(const float* p) { return {p[0], p[1], p[2], p[3]}; }
void store3(float* p, V v) {
    p[0] = v.x;
    p[1] = v.y;
    p[2] = v.z;
}

struct Cand {
    V a, b, n;
    float s;
};

// ---------------- GuPCMContactBoxBox.cpp ----------------
struct Box {
    V half, pos;
    Q rot;
    V axis[3];
    Box(V h, V p, Q q) : half(h), pos(p), rot(q) {
        axis[0] = gu_rotate(q, {1, 0, 0});
        axis[1] = gu_rotate(q, {0, 1, 0});
        axis[2] = gu_rotate(q, {0, 0, 1});
    }
};
V local_of(const Box& b, V w) { return gu_inverse_rotate(b.rot, sub(w, b.pos)); }
struct Frame {
    V x, y, z, p;
    V project(V w) const {
        const V d = sub(w, p);
        return {dot(d, x), dot(d, y), dot(d, z)};
    }
    V unproject(V v) const { return add(p, add(add(mul(x, v.x), mul(y, v.y)), mul(z, v.z))); }
};

// GuPCMContactBoxBox.cpp:63-83
bool contains(const V poly[4], V p, V low, V high) {
    if (p.x < low.x || p.y < low.y || p.x > high.x || p.y > high.y) return false;
    unsigned hits = 0;
    const float eps = std::numeric_limits<float>::epsilon();
    for (unsigned i = 0, j = 3; i < 4; j = i++) {
        const V a = poly[j], b = poly[i];
        if ((p.x == a.x && p.y == a.y) || (p.x == b.x && p.y == b.y)) return true;
        if ((a.y > p.y) != (b.y > p.y)) {
            const float dx = b.x - a.x, dy = b.y - a.y;
            const float part = (p.y - a.y) * dx + (a.x + eps) * dy;
            const float target = p.x * dy;
            const bool hit = dy > 0.0f ? part >= target : target >= part;
            if (hit && ++hits == 2) return false;
        }
    }
    return hits > 0;
}

// GuPCMContactBoxBox.cpp:87-104
bool segment_slab(V p, V d, V high, V low, float& enter, float& exit) {
    enter = 0.0f;
    exit = 1.0f;
    for (unsigned a = 0; a < 3; ++a) {
        const float x = comp(p, a), dd = comp(d, a);
        const float mn = comp(low, a), mx = comp(high, a);
        if (std::abs(dd) < 1.0e-6f) {
            if (x < mn || x > mx) return false;
            continue;
        }
        const float one = (mn - x) / dd, two = (mx - x) / dd;
        enter = std::max(enter, std::min(one, two));
        exit = std::min(exit, std::max(one, two));
    }
    return enter <= exit && enter <= 1.0f;
}

struct Gen {
    const Box& a;
    const Box& b;
    Cand* out;
    std::size_t cap;
    std::size_t count = 0;
    bool overflow = false;
    void emit(const Frame& f, V ref, V inc, V normal, float sep, bool flipped) {
        if (count >= cap) {
            overflow = true;
            return;
        }
        const V w0 = f.unproject(flipped ? inc : ref);
        const V w1 = f.unproject(flipped ? ref : inc);
        out[count++] = {local_of(a, w0), local_of(b, w1), gu_inverse_rotate(b.rot, normal), sep};
    }
    // GuPCMContactBoxBox.cpp:119-178
    void calculate(const Frame& f, const V poly[4], V inc_n, float fx, float fy, V normal, float cd,
                   bool flipped) {
        fx *= 1.0001f;
        fy *= 1.0001f;
        bool pen[4] = {}, in[4] = {};
        const float big = std::numeric_limits<float>::max();
        V mn{big, big, big};
        V mx = neg(mn);
        const V high{fx, fy, big};
        for (unsigned i = 0; i < 4; ++i) {
            const V p = poly[i];
            mn.x = std::min(mn.x, p.x);
            mn.y = std::min(mn.y, p.y);
            mx.x = std::max(mx.x, p.x);
            mx.y = std::max(mx.y, p.y);
            if (-p.z < cd) {
                pen[i] = true;
                in[i] = std::abs(p.x) <= fx && std::abs(p.y) <= fy;
                if (in[i]) emit(f, {p.x, p.y, 0}, p, normal, -p.z, flipped);
            }
        }
        if (count == 4) return;
        if (inc_n.z != 0.0f) {
            const V corners[4] = {{fx, fy, 0}, {fx, -fy, 0}, {-fx, fy, 0}, {-fx, -fy, 0}};
            for (V c : corners) {
                if (!contains(poly, c, mn, mx)) continue;
                const float nom = dot(inc_n, sub(poly[0], c));
                const float t = nom / inc_n.z;
                const float sep = -t;
                if (cd > sep) emit(f, c, {c.x, c.y, t}, normal, sep, flipped);
            }
        }
        const V low{-fx, -fy, -cd - std::numeric_limits<float>::epsilon()};
        for (unsigned s = 0, e = 3; s < 4; e = s++) {
            if (!pen[s] && !pen[e]) continue;
            const bool c0 = pen[s] && in[s];
            const bool c1 = pen[e] && in[e];
            if (c0 && c1) continue;
            const V edge = sub(poly[e], poly[s]);
            float t0, t1;
            if (!segment_slab(poly[s], edge, high, low, t0, t1)) continue;
            if (!c0) {
                const V p = add(poly[s], mul(edge, t0));
                emit(f, {p.x, p.y, 0}, p, normal, -p.z, flipped);
            }
            if (!c1) {
                const V p = add(poly[s], mul(edge, t1));
                emit(f, {p.x, p.y, 0}, p, normal, -p.z, flipped);
            }
        }
    }
};

// GuPCMContactBoxBox.cpp:301-436
bool box_box_candidates(V h0, V p0, Q q0, V h1, V p1, Q q1, float cd, Cand* out, std::size_t cap,
                        std::size_t& count, bool& overflow) {
    const Box a{h0, p0, q0}, b{h1, p1, q1};
    Gen gen{a, b, out, cap};
    const V delta = sub(b.pos, a.pos);
    float sgn[6], ov[6];
    constexpr float se = 1.0e-6f;
    auto radius = [&](const Box& bx, V ax) {
        return (std::abs(dot(ax, bx.axis[0])) + se) * bx.half.x +
               (std::abs(dot(ax, bx.axis[1])) + se) * bx.half.y +
               (std::abs(dot(ax, bx.axis[2])) + se) * bx.half.z;
    };
    overflow = false;
    for (unsigned i = 0; i < 6; ++i) {
        const bool first = i < 3;
        const Box& ref = first ? a : b;
        const Box& inc = first ? b : a;
        const unsigned ax = i % 3;
        const V n = ref.axis[ax];
        sgn[i] = dot(delta, n);
        ov[i] = comp(ref.half, ax) + radius(inc, n) - std::abs(sgn[i]) + cd;
        if (ov[i] < 0.0f) {
            count = 0;
            return false;
        }
    }
    for (unsigned i = 0; i < 3; ++i)
        for (unsigned j = 0; j < 3; ++j) {
            const V ca = cross(a.axis[i], b.axis[j]);
            float ra = 0.0f, rb = 0.0f;
            for (unsigned k = 0; k < 3; ++k) {
                if (k != i) ra += (std::abs(dot(ca, a.axis[k])) + se) * comp(a.half, k);
                if (k != j) rb += (std::abs(dot(ca, b.axis[k])) + se) * comp(b.half, k);
            }
            if (std::abs(dot(delta, ca)) > ra + rb + cd) {
                count = 0;
                return false;
            }
        }
    unsigned feature = 0;
    float mo = ov[0];
    for (unsigned i = 1; i < 6; ++i)
        if (mo > ov[i]) {
            mo = ov[i];
            feature = i;
        }
    const bool flipped = feature >= 3;
    const Box& inc = flipped ? a : b;
    const bool np = sgn[feature] <= 0.0f;
    Frame fr{};
    V normal{};
    float fx = 0.0f, fy = 0.0f;
    switch (feature) {
        case 0:
            normal = np ? a.axis[0] : neg(a.axis[0]);
            fr = {np ? neg(a.axis[2]) : a.axis[2], a.axis[1], normal,
                  add(a.pos, mul(a.axis[0], np ? -a.half.x : a.half.x))};
            fx = a.half.z;
            fy = a.half.y;
            break;
        case 1:
            normal = np ? a.axis[1] : neg(a.axis[1]);
            fr = {a.axis[0], np ? neg(a.axis[2]) : a.axis[2], normal,
                  add(a.pos, mul(a.axis[1], np ? -a.half.y : a.half.y))};
            fx = a.half.x;
            fy = a.half.z;
            break;
        case 2:
            normal = np ? a.axis[2] : neg(a.axis[2]);
            fr = {a.axis[0], np ? a.axis[1] : neg(a.axis[1]), normal,
                  add(a.pos, mul(a.axis[2], np ? -a.half.z : a.half.z))};
            fx = a.half.x;
            fy = a.half.y;
            break;
        case 3:
            normal = np ? b.axis[0] : neg(b.axis[0]);
            fr = {np ? b.axis[2] : neg(b.axis[2]), b.axis[1], neg(normal),
                  add(b.pos, mul(b.axis[0], np ? b.half.x : -b.half.x))};
            fx = b.half.z;
            fy = b.half.y;
            break;
        case 4:
            normal = np ? b.axis[1] : neg(b.axis[1]);
            fr = {b.axis[0], np ? b.axis[2] : neg(b.axis[2]), neg(normal),
                  add(b.pos, mul(b.axis[1], np ? b.half.y : -b.half.y))};
            fx = b.half.x;
            fy = b.half.z;
            break;
        default:
            normal = np ? b.axis[2] : neg(b.axis[2]);
            fr = {b.axis[0], np ? neg(b.axis[1]) : b.axis[1], neg(normal),
                  add(b.pos, mul(b.axis[2], np ? b.half.z : -b.half.z))};
            fx = b.half.x;
            fy = b.half.y;
            break;
    }
    const V search = neg(fr.z);
    const float dx = std::abs(dot(search, inc.axis[0]));
    const float dy = std::abs(dot(search, inc.axis[1]));
    const float dz = std::abs(dot(search, inc.axis[2]));
    const unsigned ia = dx >= dy && dx >= dz ? 0 : (dy >= dz ? 1 : 2);
    const float facing = dot(search, inc.axis[ia]);
    const float se2 = facing > 0.0f ? -comp(inc.half, ia) : comp(inc.half, ia);
    const V fc = add(inc.pos, mul(inc.axis[ia], se2));
    const unsigned ua = ia == 0 ? 1 : 0;
    const unsigned va = ia == 2 ? 1 : 2;
    const V u = mul(inc.axis[ua], comp(inc.half, ua));
    const V v = mul(inc.axis[va], comp(inc.half, va));
    const V pw[4] = {add(fc, add(u, v)), add(fc, sub(u, v)), sub(fc, add(u, v)), sub(fc, sub(u, v))};
    V poly[4];
    for (unsigned i = 0; i < 4; ++i) poly[i] = fr.project(pw[i]);
    const V fnorm = fr.project(add(fr.p, facing > 0.0f ? neg(inc.axis[ia]) : inc.axis[ia]));
    gen.calculate(fr, poly, fnorm, fx, fy, normal, cd, flipped);
    count = gen.count;
    overflow = gen.overflow;
    return true;
}

// GuPCMContactBoxBox.cpp:185-299
void batch_selection(const Cand* c, std::size_t n, float tol, std::uint8_t sel[4]) {
    std::uint8_t avail[64];
    for (std::size_t i = 0; i < n; ++i) avail[i] = static_cast<std::uint8_t>(i);
    std::size_t rem = n;
    float deepest_d = c[0].s, shallow = deepest_d;
    std::size_t deepest = 0;
    for (std::size_t i = 1; i < rem; ++i) {
        const float d = c[i].s;
        shallow = std::max(shallow, d);
        if (deepest_d > d) {
            deepest_d = d;
            deepest = i;
        }
    }
    sel[0] = static_cast<std::uint8_t>(deepest);
    avail[deepest] = avail[--rem];
    std::size_t fp = 0;
    float far2 = len2(sub(c[avail[0]].b, c[sel[0]].b));
    for (std::size_t i = 1; i < rem; ++i) {
        const float v = len2(sub(c[avail[i]].b, c[sel[0]].b));
        if (v > far2) {
            far2 = v;
            fp = i;
        }
    }
    sel[1] = avail[fp];
    avail[fp] = avail[--rem];
    const V base = sub(c[sel[1]].b, c[sel[0]].b);
    V tr = cross(base, c[sel[0]].n);
    const float tr2 = len2(tr);
    if (tr2 > 0.0f) {
        const float inv = 1.0f / std::sqrt(tr2);
        tr = {tr.x * inv, tr.y * inv, tr.z * inv};
    } else {
        tr = c[sel[0]].n;
    }
    float greatest = -std::numeric_limits<float>::max();
    float least = std::numeric_limits<float>::max();
    std::size_t gp = 0, lp = 0;
    for (std::size_t i = 0; i < rem; ++i) {
        const float side = dot(sub(c[avail[i]].b, c[sel[0]].b), tr);
        if (side > greatest) {
            greatest = side;
            gp = i;
        }
        if (side < least) {
            least = side;
            lp = i;
        }
    }
    sel[2] = avail[gp];
    avail[gp] = avail[--rem];
    if (lp == rem) lp = gp;
    if (least * greatest > 0.0f) {
        greatest = -std::numeric_limits<float>::max();
        for (std::size_t i = 0; i < rem; ++i) {
            const float side = dot(sub(c[avail[i]].b, c[sel[0]].b), tr);
            if (side > greatest) {
                greatest = side;
                lp = i;
            }
        }
    }
    sel[3] = avail[lp];
    avail[lp] = avail[--rem];
    const float th = tol * 0.02f;
    if (deepest_d < th && shallow > th) {
        for (std::size_t i = 0; i < 4; ++i) {
            float depth = c[sel[i]].s;
            if (depth > th) {
                std::size_t rep = rem;
                for (std::size_t j = 0; j < rem; ++j) {
                    const float cdp = c[avail[j]].s;
                    if (depth > cdp && th > cdp) {
                        depth = cdp;
                        rep = j;
                    }
                }
                if (rep < rem) std::swap(sel[i], avail[rep]);
            }
        }
    }
}

// ---------------- GuPCMGjkEpa.cpp (box/box only) ----------------
float box_min_margin(V h) { return 0.05f * std::min({h.x, h.y, h.z}); }
struct Sup {
    V a, b, q;
    std::uint32_t ai, bi;
};
struct Pair2 {
    V ha, hb, a_in_b;
    Q rel;
    static V vertex(V e, V d, std::uint32_t& idx) {
        const bool x = d.x > 0.f, y = d.y > 0.f, z = d.z > 0.f;
        idx = (x ? 1u : 0u) | (y ? 2u : 0u) | (z ? 4u : 0u);
        return {x ? e.x : -e.x, y ? e.y : -e.y, z ? e.z : -e.z};
    }
    Sup support(V dir) const {
        const V da = gu_rotate({-rel.x, -rel.y, -rel.z, rel.w}, dir);
        Sup s{};
        const V pa = vertex(ha, da, s.ai);
        s.b = vertex(hb, neg(dir), s.bi);
        s.a = add(a_in_b, gu_rotate(rel, pa));
        s.q = sub(s.a, s.b);
        return s;
    }
    V local_a(V pb) const { return gu_rotate({-rel.x, -rel.y, -rel.z, rel.w}, sub(pb, a_in_b)); }
};
// GuPCMGjkEpa.cpp:87-112
bool solve(int n, float a[3][4], float r[3]) {
    float scale = 0.0f;
    for (int i = 0; i < n; ++i) scale = std::max(scale, std::abs(a[i][i]));
    if (!(scale > 0.0f) || !std::isfinite(scale)) return false;
    const float tol = std::numeric_limits<float>::epsilon() * scale;
    for (int col = 0; col < n; ++col) {
        int piv = col;
        for (int row = col + 1; row < n; ++row)
            if (std::abs(a[row][col]) > std::abs(a[piv][col])) piv = row;
        if (!(std::abs(a[piv][col]) > tol)) return false;
        for (int j = col; j <= n; ++j) std::swap(a[col][j], a[piv][j]);
        const float div = a[col][col];
        for (int j = col; j <= n; ++j) a[col][j] /= div;
        for (int row = 0; row < n; ++row)
            if (row != col) {
                const float f = a[row][col];
                for (int j = col; j <= n; ++j) a[row][j] = std::fma(-f, a[col][j], a[row][j]);
            }
    }
    for (int i = 0; i < n; ++i) r[i] = a[i][n];
    return true;
}
struct Closest {
    V p{}, a{}, b{};
    float w[4] = {0, 0, 0, 0};
    unsigned mask = 0;
    float d2 = std::numeric_limits<float>::max();
};
// GuPCMGjkEpa.cpp:119-161
Closest closest_simplex(const Sup* sx, int count) {
    Closest best;
    for (unsigned mask = 1; mask < (1u << count); ++mask) {
        int ids[4] = {}, n = 0;
        for (int i = 0; i < count; ++i)
            if (mask & (1u << i)) ids[n++] = i;
        float w[4] = {0, 0, 0, 0};
        w[ids[0]] = 1.0f;
        if (n > 1) {
            const V base = sx[ids[0]].q;
            V dirs[3] = {};
            float m[3][4] = {};
            for (int i = 1; i < n; ++i) dirs[i - 1] = sub(sx[ids[i]].q, base);
            for (int i = 1; i < n; ++i) {
                for (int j = 1; j < n; ++j) m[i - 1][j - 1] = dot(dirs[i - 1], dirs[j - 1]);
                m[i - 1][n - 1] = -dot(dirs[i - 1], base);
            }
            float co[3] = {};
            if (!solve(n - 1, m, co)) continue;
            for (int i = 1; i < n; ++i) {
                w[ids[i]] = co[i - 1];
                w[ids[0]] -= co[i - 1];
            }
        }
        bool feasible = true;
        for (int i = 0; i < n; ++i)
            if (w[ids[i]] < -1.0e-5f) feasible = false;
        if (!feasible) continue;
        V q{0, 0, 0}, pa{0, 0, 0}, pb{0, 0, 0};
        for (int i = 0; i < count; ++i) {
            q = add(q, mul(sx[i].q, w[i]));
            pa = add(pa, mul(sx[i].a, w[i]));
            pb = add(pb, mul(sx[i].b, w[i]));
        }
        const float d = len2(q);
        if (d < best.d2) {
            best.p = q;
            best.a = pa;
            best.b = pb;
            for (int k = 0; k < 4; ++k) best.w[k] = w[k];
            best.mask = mask;
            best.d2 = d;
        }
    }
    return best;
}
void compact(Sup* sx, int& count, const Closest& c) {
    int w = 0;
    for (int r = 0; r < count; ++r)
        if ((c.mask & (1u << r)) && c.w[r] > 1.0e-6f) sx[w++] = sx[r];
    if (w == 0 && count > 0) w = 1;
    count = w;
}
bool contact_from_gjk(const Pair2& pr, const Closest& c, Cand& out) {
    const float dist = std::sqrt(c.d2);
    if (!(dist > 0.0f) || !std::isfinite(dist)) return false;
    const V n = mul(c.p, 1.0f / dist);
    out = {pr.local_a(c.a), c.b, n, dist};
    return true;
}
struct Facet {
    int i = 0, j = 0, k = 0;
    V n{};
    float d = 0.0f;
    bool valid = false;
    bool in_heap = true;
};
Facet make_facet(const Sup* v, int i, int j, int k) {
    const V a = v[i].q, b = v[j].q, c = v[k].q;
    const V e0 = sub(b, a), e1 = sub(c, a);
    V n = cross(e0, e1);
    const float n2 = len2(n);
    if (!(n2 > std::numeric_limits<float>::epsilon())) return {};
    n = mul(n, 1.0f / std::sqrt(n2));
    Facet f;
    f.i = i;
    f.j = j;
    f.k = k;
    f.n = n;
    f.d = dot(n, a);
    f.valid = true;
    f.in_heap = true;
    return f;
}
// GuPCMGjkEpa.cpp:222-268
bool init_epa(const Pair2& pr, Sup* v, int& count, Facet* fa, int& fc) {
    if (count <= 0 || count > 4) return false;
    if (count == 1) {
        const Sup nx = pr.support({1, 0, 0});
        if (len2(sub(nx.q, v[0].q)) == 0.0f) return false;
        v[count++] = nx;
    }
    if (count == 2) {
        const V e = sub(v[1].q, v[0].q);
        const V ab{std::abs(e.x), std::abs(e.y), std::abs(e.z)};
        V ax{1, 0, 0};
        if (ab.x > ab.y && ab.z > ab.y)
            ax = {0, 1, 0};
        else if (ab.x > ab.z)
            ax = {0, 0, 1};
        V perp = cross(ax, e);
        const float sq = len2(perp);
        if (!(sq > 0.0f)) return false;
        perp = mul(perp, 1.0f / std::sqrt(sq));
        v[count++] = pr.support(perp);
    }
    if (count == 3) {
        const Facet f0 = make_facet(v, 0, 1, 2);
        const Facet f1 = make_facet(v, 1, 0, 2);
        if (!f0.valid || !f1.valid) return false;
        fa[fc++] = f0;
        fa[fc++] = f1;
        return true;
    }
    const V p0 = v[0].q;
    const V v1 = sub(v[1].q, p0), v2 = sub(v[2].q, p0);
    if (dot(cross(v1, v2), sub(v[3].q, p0)) > 0.0f) std::swap(v[1], v[2]);
    const int faces[4][3] = {{0, 1, 2}, {0, 3, 1}, {0, 2, 3}, {1, 3, 2}};
    for (const auto& f : faces) {
        const Facet x = make_facet(v, f[0], f[1], f[2]);
        if (!x.valid) return false;
        fa[fc++] = x;
    }
    return true;
}
// GuPCMGjkEpa.cpp:269-301
bool contact_from_facet(const Pair2& pr, const Sup* v, const Facet& f, Cand& out) {
    const Sup& p = v[f.i];
    const Sup& q = v[f.j];
    const Sup& r = v[f.k];
    const V e0 = sub(q.q, p.q), e1 = sub(r.q, p.q);
    const float a = dot(e0, e0), b = dot(e0, e1), c = dot(e1, e1);
    const float den = a * c - b * b;
    if (!(den > std::numeric_limits<float>::epsilon() * a * c)) return false;
    const float x = -dot(p.q, e0), y = -dot(p.q, e1);
    const float u = (x * c - y * b) / den;
    const float vv = (y * a - x * b) / den;
    const float w = 1.0f - u - vv;
    const V pa = add(add(mul(p.a, w), mul(q.a, u)), mul(r.a, vv));
    const V pb = add(add(mul(p.b, w), mul(q.b, u)), mul(r.b, vv));
    if (!std::isfinite(pa.x) || !std::isfinite(pb.x) || !std::isfinite(f.d)) return false;
    const V n = neg(f.n);
    const float depth = std::abs(f.d);
    out = {pr.local_a(pa), pb, n, -depth};
    const V diff = sub(pa, pb);
    return std::sqrt(len2(diff)) <= depth + 1.0e-3f;
}
// GuPCMGjkEpa.cpp:302-395
bool epa(const Pair2& pr, const Sup* sx, int sc, Cand& out) {
    Sup v[64];
    int vc = sc;
    for (int i = 0; i < sc; ++i) v[i] = sx[i];
    Facet fa[64];
    int fc = 0;
    if (!init_epa(pr, v, vc, fa, fc)) return false;
    const float eps = 0.1f * std::min(box_min_margin(pr.ha), box_min_margin(pr.hb));
    struct Edge {
        int from, to;
    };
    Edge bd[32];
    float upper = std::numeric_limits<float>::max();
    for (int it = 0; it < 60; ++it) {
        int closest = -1;
        float dist = std::numeric_limits<float>::max();
        for (int i = 0; i < fc; ++i)
            if (fa[i].valid && fa[i].in_heap && fa[i].d < dist) {
                closest = i;
                dist = fa[i].d;
            }
        if (closest < 0) return false;
        const Facet sel = fa[closest];
        const Sup np = pr.support(sel.n);
        const float proj = dot(np.q, sel.n);
        if (std::abs(proj - dist) <= eps) return contact_from_facet(pr, v, sel, out);
        if (vc == 64) return false;
        upper = std::min(upper, proj);
        bool dup = false;
        for (int i = 0; i < vc; ++i)
            if (v[i].ai == np.ai && v[i].bi == np.bi) {
                dup = true;
                break;
            }
        if (dup) return contact_from_facet(pr, v, sel, out);
        const int ni = vc++;
        v[ni] = np;
        int bc = 0;
        for (int i = 0; i < fc; ++i) {
            Facet& f = fa[i];
            if (!f.valid) continue;
            if (dot(f.n, sub(np.q, v[f.i].q)) < 0.0f) continue;
            f.valid = false;
            const Edge es[3] = {{f.i, f.j}, {f.j, f.k}, {f.k, f.i}};
            for (const Edge e : es) {
                int rev = -1;
                for (int j = 0; j < bc; ++j)
                    if (bd[j].from == e.to && bd[j].to == e.from) {
                        rev = j;
                        break;
                    }
                if (rev >= 0)
                    bd[rev] = bd[--bc];
                else {
                    if (bc == 32) return false;
                    bd[bc++] = e;
                }
            }
        }
        if (bc == 0) return false;
        int avail = 64 - fc;
        for (int i = 0; i < fc; ++i) avail += !fa[i].valid;
        if (bc > avail) return false;
        for (int i = 0; i < bc; ++i) {
            const Edge e = bd[i];
            Facet f = make_facet(v, e.from, e.to, ni);
            if (!f.valid) return false;
            f.in_heap = f.d <= upper;
            int slot = -1;
            for (int j = 0; j < fc; ++j)
                if (!fa[j].valid) {
                    slot = j;
                    break;
                }
            if (slot < 0) slot = fc++;
            fa[slot] = f;
        }
    }
    return false;
}
// GuPCMGjkEpa.cpp:396-515 (polyhedral cores only: no sphere radius)
bool gjk_epa(const Pair2& pr, float cd, Cand& out) {
    const float eps = 0.1f * std::min(box_min_margin(pr.ha), box_min_margin(pr.hb));
    constexpr float progress = 1.0f - 0.000225f;
    Sup sx[4];
    int count = 0;
    V closest = len2(pr.a_in_b) > 0.0f ? pr.a_in_b : V{1, 0, 0};
    float previous = std::numeric_limits<float>::max();
    for (int it = 0; it < 32; ++it) {
        const float dist = std::sqrt(len2(closest));
        if (count > 0 && dist <= eps) return epa(pr, sx, count, out);
        if (dist <= 0.0f) return false;
        const V n = mul(closest, 1.0f / dist);
        const Sup s = pr.support(neg(closest));
        const float proj = dot(n, s.q);
        if (proj > cd) return false;
        if (count > 0 && proj > dist * progress) {
            const Closest cur = closest_simplex(sx, count);
            return cur.mask && contact_from_gjk(pr, cur, out);
        }
        for (int i = 0; i < count; ++i)
            if (s.ai == sx[i].ai && s.bi == sx[i].bi) {
                if (proj > 0.f) {
                    const Closest cur = closest_simplex(sx, count);
                    return cur.mask && contact_from_gjk(pr, cur, out) && out.s <= cd;
                }
                return false;
            }
        if (count == 4) return epa(pr, sx, count, out);
        sx[count++] = s;
        const Closest nx = closest_simplex(sx, count);
        if (!nx.mask) return false;
        closest = nx.p;
        const float nd = std::sqrt(nx.d2);
        compact(sx, count, nx);
        if (nd >= previous) return false;
        previous = nd;
    }
    return false;
}
// GuPCMGjkEpa.cpp:527-566 (box/box entry)
bool box_box_gjk_epa(V h0, V p0, Q q0, V h1, V p1, Q q1, float cd, Cand& out) {
    const V rel = gu_inverse_rotate(q1, sub(p0, p1));
    const Q ib{-q1.x, -q1.y, -q1.z, q1.w};
    const Q& r0 = q0;
    const Q qa{ib.w * r0.x + r0.w * ib.x + ib.y * r0.z - ib.z * r0.y,
               ib.w * r0.y + r0.w * ib.y + ib.z * r0.x - ib.x * r0.z,
               ib.w * r0.z + r0.w * ib.z + ib.x * r0.y - ib.y * r0.x,
               ib.w * r0.w - ib.x * r0.x - ib.y * r0.y - ib.z * r0.z};
    Pair2 pr{h0, h1, rel, qa};
    if (!gjk_epa(pr, cd, out)) return false;
    return out.s <= cd;
}

// ---------------- scene.cpp persistent manifolds ----------------
struct Cache {
    Cand pts[4];
    V rel_pos;
    Q rel_rot, orient_a, orient_b;
    std::uint32_t count, initialized;
};
static_assert(sizeof(Cand) == 40, "PCM point layout");

// scene.cpp:3432-3465 (scene rotate)
void refresh(Cache& c, V pos_ab, Q rot_ab, float brk) {
    const float b2 = brk * brk;
    for (std::uint32_t i = c.count; i > 0; --i) {
        Cand& p = c.pts[i - 1];
        const V r = orc::rotate(rot_ab, p.a);
        const V ainb{pos_ab.x + r.x, pos_ab.y + r.y, pos_ab.z + r.z};
        const V v{ainb.x - p.b.x, ainb.y - p.b.y, ainb.z - p.b.z};
        const float d = v.x * p.n.x + v.y * p.n.y + v.z * p.n.z;
        const V pd{v.x - p.n.x * d, v.y - p.n.y * d, v.z - p.n.z * d};
        const float t2 = pd.x * pd.x + pd.y * pd.y + pd.z * pd.z;
        if (t2 > b2)
            c.pts[i - 1] = c.pts[--c.count];
        else
            p.s = d;
    }
}

// scene.cpp:3317-3431
void reduce_five(Cache& cache, const Cand& nc) {
    Cand cd[5];
    for (int i = 0; i < 4; ++i) cd[i] = cache.pts[i];
    cd[4] = nc;
    bool chosen[5] = {};
    int sel[4] = {};
    auto sq = [](V a) { return dot(a, a); };
    auto seg = [&](V p, V a, V b) {
        const V ab = sub(b, a), ap = sub(p, a);
        const float den = sq(ab);
        const float t = den == 0.0f ? 0.0f : std::clamp(dot(ap, ab) / den, 0.0f, 1.0f);
        const V d{ap.x - t * ab.x, ap.y - t * ab.y, ap.z - t * ab.z};
        return sq(d);
    };
    auto tri = [&](V p, V a, V b, V c) {
        const V ab = sub(b, a), ac = sub(c, a);
        const V ap = sub(p, a), bp = sub(p, b), cp = sub(p, c);
        const float d1 = dot(ab, ap), d2 = dot(ac, ap);
        const float d3 = dot(ab, bp), d4 = dot(ac, bp);
        const float d5 = dot(ab, cp), d6 = dot(ac, cp);
        if (d1 < 0.0f && d2 < 0.0f) return sq(ap);
        if (d3 >= 0.0f && d3 >= d4) return sq(bp);
        if (d6 >= 0.0f && d6 >= d5) return sq(cp);
        const float vc = d1 * d4 - d3 * d2;
        if (vc < 0.0f && d1 >= 0.0f && d3 < 0.0f) return seg(p, a, b);
        const float va = d3 * d6 - d5 * d4;
        if (va < 0.0f && d4 >= d3 && d5 >= d6) return seg(p, b, c);
        const float vb = d5 * d2 - d1 * d6;
        if (vb < 0.0f && d2 >= 0.0f && d6 < 0.0f) return seg(p, a, c);
        const V n{ab.y * ac.z - ab.z * ac.y, ab.z * ac.x - ab.x * ac.z, ab.x * ac.y - ab.y * ac.x};
        const float den = sq(n);
        if (den == 0.0f) return 0.0f;
        const float pr = dot(n, ap);
        return pr * pr / den;
    };
    int deepest = 4;
    for (int i = 0; i < 4; ++i)
        if (cd[i].s < cd[deepest].s) deepest = i;
    sel[0] = deepest;
    chosen[deepest] = true;
    int far = 0;
    float largest = sq(sub(cd[0].b, cd[deepest].b));
    for (int i = 1; i < 5; ++i)
        if (!chosen[i]) {
            const float v = sq(sub(cd[i].b, cd[deepest].b));
            if (v > largest) {
                largest = v;
                far = i;
            }
        }
    sel[1] = far;
    chosen[far] = true;
    largest = -std::numeric_limits<float>::max();
    int third = 0;
    for (int i = 0; i < 5; ++i)
        if (!chosen[i]) {
            const float v = seg(cd[i].b, cd[deepest].b, cd[far].b);
            if (v > largest) {
                largest = v;
                third = i;
            }
        }
    sel[2] = third;
    chosen[third] = true;
    largest = -std::numeric_limits<float>::max();
    int fourth = third;
    for (int i = 0; i < 5; ++i)
        if (!chosen[i]) {
            const float v = tri(cd[i].b, cd[sel[0]].b, cd[sel[1]].b, cd[sel[2]].b);
            if (v > largest) {
                largest = v;
                fourth = i;
            }
        }
    if (chosen[fourth]) {
        cache.count = 3;
        for (int i = 0; i < 3; ++i) cache.pts[i] = cd[sel[i]];
        return;
    }
    sel[3] = fourth;
    chosen[fourth] = true;
    for (int i = 0; i < 4; ++i) cache.pts[i] = cd[sel[i]];
    cache.count = 4;
    int uns = -1;
    for (int i = 0; i < 5; ++i)
        if (!chosen[i]) {
            uns = i;
            break;
        }
    if (uns < 0) return;
    float closest = std::numeric_limits<float>::max();
    int rep = 0;
    for (int i = 0; i < 4; ++i) {
        const float d = sq(sub(cache.pts[i].a, cd[uns].a));
        if (d < closest) {
            closest = d;
            rep = i;
        }
    }
    if (cd[uns].s < cache.pts[rep].s) cache.pts[rep] = cd[uns];
}

float pcm_margin(V h) { return std::min(0.15f * std::min({h.x, h.y, h.z}), 0.15f); }

// scene.cpp:3466-3654. Returns the number of emitted contacts (-1 on the
// reference's overflow/logic errors).
int box_box(V h0, V p0, Q q0, V h1, V p1, Q q1, float cd, Cache& cache, float* out_pts,
            std::uint32_t* flags) {
    const V delta{p0.x - p1.x, p0.y - p1.y, p0.z - p1.z};
    const V rel = orc::rotate_inverse(q1, delta);
    const Q inv_b{-q1.x, -q1.y, -q1.z, q1.w};
    const Q rot = orc::multiply(inv_b, q0);
    const float mm = std::min(pcm_margin(h0), pcm_margin(h1));
    const std::uint32_t initial = cache.count;
    refresh(cache, rel, rot, mm * 0.8f);
    const bool lost = cache.count != initial;
    const float th[5] = {0.5f, 0.125f, 0.25f, 0.375f, 0.375f};
    const float qth[5] = {0.9998f, 0.9999f, 0.9999f, 0.9999f, 0.9999f};
    const float td = cache.initialized
                         ? std::max({std::abs(rel.x - cache.rel_pos.x), std::abs(rel.y - cache.rel_pos.y),
                                     std::abs(rel.z - cache.rel_pos.z)})
                         : std::numeric_limits<float>::infinity();
    auto qdot = [](Q a, Q b) { return a.x * b.x + a.y * b.y + a.z * b.z + a.w * b.w; };
    const float qa = qdot(q0, cache.orient_a);
    const float qb = qdot(q1, cache.orient_b);
    const float tp = mm * th[cache.count];
    bool regen = !cache.initialized || lost || td > tp || qa < qth[cache.count] || qb < qth[cache.count];
    if (!regen) {
        auto travel = [](float d, V e) {
            const float ang = d < 1.0f ? std::acos(std::max(-1.0f, d)) : 0.0f;
            return ang * std::sqrt(e.x * e.x + e.y * e.y + e.z * e.z);
        };
        regen = travel(qa, h0) > tp || travel(qb, h1) > tp;
    }
    bool direct = false, gjk = false;
    V gjk_n{};
    *flags = regen ? 1u : 0u;
    if (regen) {
        cache.rel_pos = rel;
        cache.rel_rot = rot;
        cache.orient_a = q0;
        cache.orient_b = q1;
        cache.initialized = 1;
        Cand cands[64];
        std::size_t count = 0;
        bool overflow = false;
        const bool overlapping = box_box_candidates(h0, p0, q0, h1, p1, q1, cd, cands, 64, count, overflow);
        if (overflow) return -1;
        if (!overlapping) {
            cache.count = 0;
            return 0;
        }
        if (count == 0) {
            *flags |= 2u;
            Cand w{};
            if (!box_box_gjk_epa(h0, p0, q0, h1, p1, q1, cd, w)) {
                cache.count = 0;
                return 0;
            }
            gjk = true;
            gjk_n = w.n;
            const float rb = mm * 0.05f;
            const float rsq = rb * rb;
            bool replaced = false;
            for (std::uint32_t i = 0; i < cache.count; ++i) {
                const Cand& o = cache.pts[i];
                if (std::min(len2(sub(o.a, w.a)), len2(sub(o.b, w.b))) < rsq) {
                    cache.pts[i] = w;
                    replaced = true;
                    break;
                }
            }
            if (!replaced) {
                if (cache.count < 4)
                    cache.pts[cache.count++] = w;
                else
                    reduce_five(cache, w);
            }
        } else if (count <= 4) {
            for (std::size_t i = 0; i < count; ++i) cache.pts[i] = cands[i];
            cache.count = static_cast<std::uint32_t>(count);
        } else {
            std::uint8_t sel[4];
            batch_selection(cands, count, 1.0f, sel);
            for (int i = 0; i < 4; ++i) cache.pts[i] = cands[sel[i]];
            cache.count = 4;
        }
        direct = count > 0;
    }
    if (cache.count == 0) return 0;
    V nb{0, 0, 0};
    if (gjk)
        nb = gjk_n;
    else if (direct)
        nb = cache.pts[0].n;
    else
        for (std::uint32_t i = 0; i < cache.count; ++i) {
            nb.x += cache.pts[i].n.x;
            nb.y += cache.pts[i].n.y;
            nb.z += cache.pts[i].n.z;
        }
    const float n2 = nb.x * nb.x + nb.y * nb.y + nb.z * nb.z;
    if (!(n2 > 0.0f) || !std::isfinite(n2)) return -1;
    const float il = 1.0f / std::sqrt(n2);
    const V nn{nb.x * il, nb.y * il, nb.z * il};
    const V normal = orc::rotate(q1, nn);
    int emitted = 0;
    for (std::uint32_t i = 0; i < cache.count; ++i) {
        const Cand& p = cache.pts[i];
        if (!direct && p.s > cd) continue;
        const V r = orc::rotate(q1, p.b);
        float* o = out_pts + 7 * emitted;
        o[0] = r.x + p1.x;
        o[1] = r.y + p1.y;
        o[2] = r.z + p1.z;
        o[3] = normal.x;
        o[4] = normal.y;
        o[5] = normal.z;
        o[6] = p.s;
        ++emitted;
    }
    return emitted;
}

// scene.cpp:3774-3955 (box geometry). plane: position (x,h,z), normal +y.
int plane_box(V plane_pos, V h, V p, Q q, float cd, Cache& cache, float* out_pts, std::uint32_t* flags) {
    const V rel{p.x - plane_pos.x, p.y - plane_pos.y, p.z - plane_pos.z};
    const float margin = std::min(0.15f * std::min({h.x, h.y, h.z}), 0.15f);
    const float brk = margin * 0.2f;
    auto sq = [](V v) { return v.x * v.x + v.y * v.y + v.z * v.z; };
    const std::uint32_t prev = cache.count;
    refresh(cache, rel, q, brk);
    const float md = cache.initialized
                         ? std::max({std::abs(rel.x - cache.rel_pos.x), std::abs(rel.y - cache.rel_pos.y),
                                     std::abs(rel.z - cache.rel_pos.z)})
                         : std::numeric_limits<float>::infinity();
    const Q& old = cache.rel_rot;
    const float rd = q.x * old.x + q.y * old.y + q.z * old.z + q.w * old.w;
    const bool regen = !cache.initialized || cache.count != prev || md > brk || rd < 0.99996f;
    *flags = regen ? 1u : 0u;
    if (regen) {
        cache.rel_pos = rel;
        cache.rel_rot = q;
        cache.initialized = 1;
        cache.count = 0;
        Cand c[8];
        std::size_t n = 0;
        for (unsigned corner = 0; corner < 8u; ++corner) {
            const V local{(corner & 4u) ? -h.x : h.x, (corner & 2u) ? -h.y : h.y, (corner & 1u) ? -h.z : h.z};
            const V r = orc::rotate(q, local);
            const V ip{rel.x + r.x, rel.y + r.y, rel.z + r.z};
            if (ip.y >= cd) continue;
            c[n++] = {local, {ip.x, 0.0f, ip.z}, {0.0f, 1.0f, 0.0f}, ip.y};
        }
        if (n <= 4) {
            for (std::size_t i = 0; i < n; ++i) cache.pts[i] = c[i];
            cache.count = static_cast<std::uint32_t>(n);
        } else {
            bool chosen[8] = {};
            std::size_t idx[4] = {};
            std::size_t first = 0;
            for (std::size_t i = 1; i < n; ++i)
                if (c[i].s < c[first].s) first = i;
            idx[0] = first;
            chosen[first] = true;
            std::size_t far = 0;
            float maxd = sq(sub(c[0].b, c[first].b));
            for (std::size_t i = 1; i < n; ++i) {
                const float d = sq(sub(c[i].b, c[first].b));
                if (d > maxd) {
                    maxd = d;
                    far = i;
                }
            }
            idx[1] = far;
            chosen[far] = true;
            const V segv = sub(c[far].b, c[first].b);
            V side{-segv.z, 0.0f, segv.x};
            const float ns = sq(side);
            if (ns > 0.0f) {
                const float inv = 1.0f / std::sqrt(ns);
                side = {side.x * inv, 0.0f, side.z * inv};
            } else {
                side = {0.0f, 1.0f, 0.0f};
            }
            float largest = -std::numeric_limits<float>::max();
            float smallest = std::numeric_limits<float>::max();
            std::size_t third = 0, fourth = 0;
            for (std::size_t i = 0; i < n; ++i) {
                if (chosen[i]) continue;
                const float d = dot(sub(c[i].b, c[first].b), side);
                if (d > largest) {
                    largest = d;
                    third = i;
                }
                if (d < smallest) {
                    smallest = d;
                    fourth = i;
                }
            }
            idx[2] = third;
            chosen[third] = true;
            if (smallest * largest > 0.0f) {
                largest = -std::numeric_limits<float>::max();
                for (std::size_t i = 0; i < n; ++i) {
                    if (chosen[i]) continue;
                    const float d = dot(sub(c[i].b, c[first].b), side);
                    if (d > largest) {
                        largest = d;
                        fourth = i;
                    }
                }
            }
            idx[3] = fourth;
            chosen[fourth] = true;
            Cand centers[4];
            for (int k = 0; k < 4; ++k) centers[k] = c[idx[k]];
            for (std::size_t i = 0; i < n; ++i) {
                if (chosen[i]) continue;
                float nearest = std::numeric_limits<float>::max();
                std::size_t slot = 0;
                for (std::size_t j = 0; j < 4; ++j) {
                    const float d = sq(sub(c[i].b, centers[j].b));
                    if (d < nearest) {
                        nearest = d;
                        slot = j;
                    }
                }
                if (c[i].s < c[idx[slot]].s) idx[slot] = i;
            }
            for (int k = 0; k < 4; ++k) cache.pts[k] = c[idx[k]];
            cache.count = 4;
        }
    }
    int emitted = 0;
    for (std::uint32_t k = 0; k < cache.count; ++k) {
        const Cand& pt = cache.pts[k];
        if (pt.s > cd) continue;
        float* o = out_pts + 7 * emitted;
        o[0] = plane_pos.x + pt.b.x;
        o[1] = plane_pos.y;
        o[2] = plane_pos.z + pt.b.z;
        o[3] = 0.0f;
        o[4] = -1.0f;
        o[5] = 0.0f;
        o[6] = pt.s;
        ++emitted;
    }
    return emitted;
}

void cand_out(const Cand& c, float* o) {
    store3(o, c.a);
    store3(o + 3, c.b);
    store3(o + 6, c.n);
    o[9] = c.s;
}
Cand cand_in(const float* o) { return {load3(o), load3(o + 3), load3(o + 6), o[9]}; }

}  // namespace

extern "C" {

// Box layout: half(3) position(3) quat(4). Candidate layout: local_a(3)
// local_b(3) normal_b(3) separation(1).
int orc_pcm_box_box_candidates(const float* a, const float* b, float cd, float* out, std::uint32_t cap,
                               std::uint32_t* count, int* overlapping) {
    Cand c[64];
    std::size_t n = 0;
    bool overflow = false;
    const bool ov = box_box_candidates(load3(a), load3(a + 3), load4(a + 6), load3(b), load3(b + 3), load4(b + 6),
                                       cd, c, 64, n, overflow);
    if (overflow) return -1;
    *overlapping = ov ? 1 : 0;
    *count = static_cast<std::uint32_t>(n);
    for (std::size_t i = 0; i < n && i < cap; ++i) cand_out(c[i], out + 10 * i);
    return 0;
}

int orc_pcm_batch_selection(const float* cands, std::uint32_t n, float tol, std::uint8_t* sel4) {
    if (n <= 4 || n > 64) return -1;
    Cand c[64];
    for (std::uint32_t i = 0; i < n; ++i) c[i] = cand_in(cands + 10 * i);
    batch_selection(c, n, tol, sel4);
    return 0;
}

int orc_pcm_gjk_epa(const float* a, const float* b, float cd, float* out, int* ok) {
    Cand c{};
    const bool r = box_box_gjk_epa(load3(a), load3(a + 3), load4(a + 6), load3(b), load3(b + 3), load4(b + 6),
                                   cd, c);
    *ok = r ? 1 : 0;
    cand_out(c, out);
    return 0;
}

// Persistent manifolds. cache: orc_pcm_cache_floats() floats per pair in
// the layout of the device cache (points[4] x 10, rel_pos 3, rel_rot 4,
// orient_a 4, orient_b 4, count, initialized). out_pts: 4 x 7 floats.
// Returns #contacts or -1; *flags bit0 regenerated, bit1 GJK/EPA ran.
std::uint32_t orc_pcm_cache_floats(void) { return 56; }

static void cache_in(const float* f, Cache& c) {
    for (int i = 0; i < 4; ++i) c.pts[i] = cand_in(f + 10 * i);
    c.rel_pos = load3(f + 40);
    c.rel_rot = load4(f + 43);
    c.orient_a = load4(f + 47);
    c.orient_b = load4(f + 51);
    std::uint32_t u[2];
    std::memcpy(u, f + 55, 4);
    c.count = u[0] & 0xffu;
    c.initialized = (u[0] >> 8) & 1u;
}
static void cache_out(const Cache& c, float* f) {
    for (int i = 0; i < 4; ++i) cand_out(c.pts[i], f + 10 * i);
    store3(f + 40, c.rel_pos);
    f[43] = c.rel_rot.x; f[44] = c.rel_rot.y; f[45] = c.rel_rot.z; f[46] = c.rel_rot.w;
    f[47] = c.orient_a.x; f[48] = c.orient_a.y; f[49] = c.orient_a.z; f[50] = c.orient_a.w;
    f[51] = c.orient_b.x; f[52] = c.orient_b.y; f[53] = c.orient_b.z; f[54] = c.orient_b.w;
    const std::uint32_t u = (c.count & 0xffu) | (c.initialized ? 0x100u : 0u);
    std::memcpy(f + 55, &u, 4);
}

// first/second: box layout as above (Gu order: second's frame holds the
// normal). cd = contact distance (sum of contact offsets).
int orc_pcm_box_box(const float* first, const float* second, float cd, float* cache, float* out_pts,
                    std::uint32_t* flags) {
    Cache c{};
    cache_in(cache, c);
    const int r = box_box(load3(first), load3(first + 3), load4(first + 6), load3(second), load3(second + 3),
                          load4(second + 6), cd, c, out_pts, flags);
    cache_out(c, cache);
    return r;
}

int orc_pcm_plane_box(const float* plane_pos, const float* box, float cd, float* cache, float* out_pts,
                      std::uint32_t* flags) {
    Cache c{};
    cache_in(cache, c);
    const int r = plane_box(load3(plane_pos), load3(box), load3(box + 3), load4(box + 6), cd, c, out_pts, flags);
    cache_out(c, cache);
    return r;
}

// Many pairs at once (a whole step's pair list): pair t is orc_pcm_box_box
// of box slots first[t] (Gu first) and second[t] read from the slot arrays
// pos[3n], quat[4n], half[3n]; caches[t] (56 floats) in and out,
// out_pts[t] 4 x 7 floats, counts[t], flags[t]. With threads > 1 the pairs
// are split over that many host threads (pairs are independent). Returns
// 0, or -1 - t for the first pair whose manifold is invalid.
int orc_pcm_box_box_many(std::uint32_t n_pairs, const std::uint32_t* first, const std::uint32_t* second,
                         const float* pos, const float* quat, const float* half, float cd, float* caches,
                         float* out_pts, std::uint32_t* counts, std::uint32_t* flags, std::uint32_t threads) {
    std::atomic<std::int64_t> bad{-1};
    auto run = [&](std::uint32_t begin, std::uint32_t end) {
        for (std::uint32_t t = begin; t < end; ++t) {
            float a[10], b[10];
            for (int k = 0; k < 3; ++k) {
                a[k] = half[3ull * first[t] + k];
                a[3 + k] = pos[3ull * first[t] + k];
                b[k] = half[3ull * second[t] + k];
                b[3 + k] = pos[3ull * second[t] + k];
            }
            for (int k = 0; k < 4; ++k) {
                a[6 + k] = quat[4ull * first[t] + k];
                b[6 + k] = quat[4ull * second[t] + k];
            }
            const int r = orc_pcm_box_box(a, b, cd, caches + 56ull * t, out_pts + 28ull * t, flags + t);
            if (r < 0) {
                std::int64_t cur = bad.load();
                while ((cur < 0 || cur > t) && !bad.compare_exchange_weak(cur, t)) {
                }
                counts[t] = 0;
            } else {
                counts[t] = static_cast<std::uint32_t>(r);
            }
        }
    };
    parallel_ranges(n_pairs, threads, run);
    return bad.load() < 0 ? 0 : static_cast<int>(-1 - bad.load());
}

// Every box slot t < n against the plane at plane_pos: orc_pcm_plane_box
// with caches[t] in place; outputs as orc_pcm_box_box_many.
int orc_pcm_plane_box_many(std::uint32_t n, const float* plane_pos, const float* pos, const float* quat,
                           const float* half, float cd, float* caches, float* out_pts, std::uint32_t* counts,
                           std::uint32_t* flags, std::uint32_t threads) {
    std::atomic<std::int64_t> bad{-1};
    auto run = [&](std::uint32_t begin, std::uint32_t end) {
        for (std::uint32_t t = begin; t < end; ++t) {
            float b[10];
            for (int k = 0; k < 3; ++k) {
                b[k] = half[3ull * t + k];
                b[3 + k] = pos[3ull * t + k];
            }
            for (int k = 0; k < 4; ++k) b[6 + k] = quat[4ull * t + k];
            const int r = orc_pcm_plane_box(plane_pos, b, cd, caches + 56ull * t, out_pts + 28ull * t, flags + t);
            if (r < 0) {
                std::int64_t cur = bad.load();
                while ((cur < 0 || cur > t) && !bad.compare_exchange_weak(cur, t)) {
                }
                counts[t] = 0;
            } else {
                counts[t] = static_cast<std::uint32_t>(r);
            }
        }
    };
    parallel_ranges(n, threads, run);
    return bad.load() < 0 ? 0 : static_cast<int>(-1 - bad.load());
}

}  // extern "C"
