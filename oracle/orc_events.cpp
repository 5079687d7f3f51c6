// Oracle restatement of Scene::publish_contact_events (reference
// src/native/scene.cpp:9149-9259) and the touch transitions of
// Scene::update_contacts (4640-4731) for the admitted class: every dynamic
// body awake, so every touching manager reports. Test infrastructure only.
//
// Inputs: the step's touching managers in Gu order (canonical shape pair
// a < b ascending) with their published contact points, per-point normal
// impulses (the solver's float applied impulses, written back as double,
// scene.cpp:9313-9333) and the manager's Gu orientation flag
// (gu_order_reversed, scene.cpp:3280-3287); and the previous step's
// touching pairs (sorted). Outputs: ContactEvent records (world.h:178-190)
// as ev[n][6] = body0, body1, kind (0 found, 1 persists, 2 lost), count,
// shape0, shape1, with point offsets and summed impulses, and
// ContactEventPoint records (world.h:194-199).
#include <cstddef>
#include <cstdint>
#include <vector>

namespace {
struct Pair {
    std::uint32_t a, b;
    bool operator<(const Pair& o) const { return a < o.a || (a == o.a && b < o.b); }
    bool operator==(const Pair& o) const { return a == o.a && b == o.b; }
};
}  // namespace

extern "C" {

// cur: n_cur records of 36 u32/f32 words each:
//   [0] shape a, [1] shape b, [2] body of a, [3] body of b, [4] reversed,
//   [5] count, [6..33] 4 points x (position, normal, separation) as f32,
//   [34..35] unused; imp: n_cur x 4 float impulses (Gu point order).
// prev: n_prev sorted pairs [n][2]. Returns 0; -1 if an output is short.
int orc_contact_events(std::uint32_t n_cur, const std::uint32_t* cur, const float* imp, std::uint32_t n_prev,
                       const std::uint32_t* prev, std::uint32_t cap, std::uint32_t* ev, std::uint64_t* ev_off,
                       double* ev_imp, std::uint32_t pt_cap, float* pt, double* pt_imp, std::uint32_t* n_ev,
                       std::uint32_t* n_pt) {
    std::vector<Pair> touching(n_cur), before(n_prev);
    for (std::uint32_t i = 0; i < n_cur; ++i) touching[i] = {cur[36 * i], cur[36 * i + 1]};
    for (std::uint32_t i = 0; i < n_prev; ++i) before[i] = {prev[2 * i], prev[2 * i + 1]};
    // scene.cpp:4660-4699: found = touching now, not before; lost = before, not now
    std::vector<Pair> found, lost;
    {
        std::size_t j = 0;
        for (const Pair& p : touching) {
            while (j < before.size() && before[j] < p) lost.push_back(before[j++]);
            if (j < before.size() && before[j] == p)
                ++j;
            else
                found.push_back(p);
        }
        while (j < before.size()) lost.push_back(before[j++]);
    }
    std::uint32_t ne = 0, np = 0;
    bool short_out = false;
    // Lost events carry the pair's bodies; bodies come from the previous
    // touching record, so the caller passes them through `prev` order: a
    // lost pair's bodies equal its shapes' bodies, which the admitted class
    // (one shape per body, body id = shape id) makes the shapes themselves.
    auto append = [&](Pair pair, std::uint32_t body_a, std::uint32_t body_b, std::uint32_t kind,
                      std::uint32_t count, std::uint64_t offset, double impulse) {
        if (ne >= cap) {
            short_out = true;
            return;
        }
        const bool reverse = body_a > body_b;
        std::uint32_t* e = ev + 6ull * ne;
        e[0] = reverse ? body_b : body_a;
        e[1] = reverse ? body_a : body_b;
        e[2] = kind;
        e[3] = count;
        e[4] = reverse ? pair.b : pair.a;
        e[5] = reverse ? pair.a : pair.b;
        ev_off[ne] = offset;
        ev_imp[ne] = impulse;
        ++ne;
    };
    std::size_t li = 0, fi = 0;
    for (std::uint32_t i = 0; i < n_cur; ++i) {
        const std::uint32_t* r = cur + 36ull * i;
        const Pair pair{r[0], r[1]};
        while (li < lost.size() && lost[li] < pair) {
            append(lost[li], lost[li].a, lost[li].b, 2, 0, np, 0.0);
            ++li;
        }
        while (fi < found.size() && found[fi] < pair) ++fi;
        const bool is_new = fi < found.size() && found[fi] == pair;
        const std::uint32_t count = r[5];
        const float* pts = reinterpret_cast<const float*>(r + 6);
        double impulse = 0.0;
        const std::uint64_t offset = np;
        const bool flip = (r[4] != 0) != (r[2] > r[3]);
        for (std::uint32_t k = 0; k < count; ++k) {
            const double f = static_cast<double>(imp[4ull * i + k]);
            impulse += f;
            if (np >= pt_cap) {
                short_out = true;
                continue;
            }
            float* o = pt + 7ull * np;
            for (int c = 0; c < 3; ++c) o[c] = pts[7 * k + c];
            for (int c = 3; c < 6; ++c) o[c] = flip ? -pts[7 * k + c] : pts[7 * k + c];
            o[6] = pts[7 * k + 6];
            pt_imp[np] = f;
            ++np;
        }
        append(pair, r[2], r[3], is_new ? 0u : 1u, count, offset, impulse);
    }
    for (; li < lost.size(); ++li) append(lost[li], lost[li].a, lost[li].b, 2, 0, np, 0.0);
    *n_ev = ne;
    *n_pt = np;
    return short_out ? -1 : 0;
}

}  // extern "C"
