// Oracle SAP restatement: box bounds + encoding, sorted endpoints, the
// x-axis sweep pair set and found/lost. Test infrastructure only.
#include <algorithm>
#include <cmath>
#include <vector>

#include "orc.h"
#include "orc_math.h"

using orc::Q4;
using orc::V3;

extern "C" {

// scene.cpp:1233-1316 (update_shape_bounds, box branch): world extent is
// |R| h (identity orientation takes h directly), plus the shape contact
// offset; min/max keys via encode_min/encode_max.
int orc_box_bounds(std::size_t n, const float* pos, const float* quat,
                   const float* half, const float* contact_offset,
                   std::uint32_t* bounds) {
    for (std::size_t i = 0; i < n; ++i) {
        const V3<float> p{pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]};
        const V3<float> h{half[3 * i], half[3 * i + 1], half[3 * i + 2]};
        const Q4<float> q{quat[4 * i], quat[4 * i + 1], quat[4 * i + 2], quat[4 * i + 3]};
        V3<float> e = h;
        if (q.x != 0.0f || q.y != 0.0f || q.z != 0.0f || q.w != 1.0f) {
            const V3<float> x = orc::rotate(q, V3<float>{1.0f, 0.0f, 0.0f});
            const V3<float> y = orc::rotate(q, V3<float>{0.0f, 1.0f, 0.0f});
            const V3<float> z = orc::rotate(q, V3<float>{0.0f, 0.0f, 1.0f});
            e = {std::abs(x.x) * h.x + std::abs(y.x) * h.y + std::abs(z.x) * h.z,
                 std::abs(x.y) * h.x + std::abs(y.y) * h.y + std::abs(z.y) * h.z,
                 std::abs(x.z) * h.x + std::abs(y.z) * h.y + std::abs(z.z) * h.z};
        }
        const float c = contact_offset[i];
        e.x += c;
        e.y += c;
        e.z += c;
        const float b[6] = {p.x - e.x, p.y - e.y, p.z - e.z, p.x + e.x, p.y + e.y, p.z + e.z};
        for (float v : b)
            if (!std::isfinite(v)) return -1;  // scene.cpp:1298-1300 overflow_error
        for (unsigned k = 0; k < 3; ++k) {
            bounds[k * n + i] = orc::encode_min(b[k]);
            bounds[(k + 3) * n + i] = orc::encode_max(b[k + 3]);
        }
    }
    return 0;
}

// scene.cpp:1341-1346 comparator; slots are in shape-id order, so the slot
// stands in for the id.
void orc_sort_endpoints(std::size_t n, const std::uint32_t* bounds, unsigned axis,
                        std::uint32_t* value, std::uint32_t* payload) {
    struct E {
        std::uint32_t v, slot;
        bool is_max;
    };
    std::vector<E> e;
    e.reserve(2 * n);
    for (std::size_t i = 0; i < n; ++i) {
        e.push_back({bounds[axis * n + i], std::uint32_t(i), false});
        e.push_back({bounds[(axis + 3) * n + i], std::uint32_t(i), true});
    }
    std::sort(e.begin(), e.end(), [](const E& a, const E& b) {
        if (a.v != b.v) return a.v < b.v;
        if (a.is_max != b.is_max) return !a.is_max;
        return a.slot < b.slot;
    });
    for (std::size_t i = 0; i < e.size(); ++i) {
        value[i] = e[i].v;
        payload[i] = (e[i].slot << 1) | (e[i].is_max ? 1u : 0u);
    }
}

// scene.cpp:2642-2703: x sweep with separate dynamic/static active lists,
// y/z overlap test on encoded keys, same-body exclusion, then the plane x
// dynamic-shape product; sort + unique.
std::size_t orc_sap_pairs(std::size_t n, const std::uint32_t* bounds,
                          const std::uint32_t* shape_id, const std::uint32_t* body_id,
                          const std::uint8_t* dynamic, std::size_t n_planes,
                          const std::uint32_t* plane_ids, std::uint32_t* out,
                          std::size_t cap) {
    std::vector<std::uint32_t> value(2 * n), payload(2 * n);
    orc_sort_endpoints(n, bounds, 0, value.data(), payload.data());
    std::vector<std::uint64_t> pairs;
    std::vector<std::uint32_t> act_dyn, act_sta, where(n);
    auto key = [](std::uint32_t a, std::uint32_t b) {
        const std::uint32_t lo = a < b ? a : b, hi = a < b ? b : a;
        return (std::uint64_t(lo) << 32) | hi;
    };
    for (std::size_t i = 0; i < 2 * n; ++i) {
        const std::uint32_t s = payload[i] >> 1;
        const bool is_max = payload[i] & 1u;
        auto& own = dynamic[s] ? act_dyn : act_sta;
        if (is_max) {
            const std::uint32_t idx = where[s];
            const std::uint32_t tail = own.back();
            own[idx] = tail;
            where[tail] = idx;
            own.pop_back();
            continue;
        }
        auto compare = [&](const std::vector<std::uint32_t>& list) {
            for (std::uint32_t o : list) {
                if (body_id[s] == body_id[o]) continue;
                if (bounds[1 * n + s] > bounds[4 * n + o] || bounds[1 * n + o] > bounds[4 * n + s] ||
                    bounds[2 * n + s] > bounds[5 * n + o] || bounds[2 * n + o] > bounds[5 * n + s])
                    continue;
                pairs.push_back(key(shape_id[s], shape_id[o]));
            }
        };
        compare(act_dyn);
        if (dynamic[s]) compare(act_sta);
        where[s] = std::uint32_t(own.size());
        own.push_back(s);
    }
    for (std::size_t p = 0; p < n_planes; ++p)
        for (std::size_t s = 0; s < n; ++s)
            if (dynamic[s]) pairs.push_back(key(plane_ids[p], shape_id[s]));
    std::sort(pairs.begin(), pairs.end());
    pairs.erase(std::unique(pairs.begin(), pairs.end()), pairs.end());
    for (std::size_t i = 0; i < pairs.size() && i < cap; ++i) {
        out[2 * i] = std::uint32_t(pairs[i] >> 32);
        out[2 * i + 1] = std::uint32_t(pairs[i]);
    }
    return pairs.size();
}

// scene.cpp:3245-3290: one linear merge of the sorted prior and new lists.
void orc_found_lost(std::size_t n_old, const std::uint32_t* old_pairs,
                    std::size_t n_new, const std::uint32_t* new_pairs,
                    std::uint32_t* found, std::size_t* n_found,
                    std::uint32_t* lost, std::size_t* n_lost) {
    auto k = [](const std::uint32_t* p, std::size_t i) {
        return (std::uint64_t(p[2 * i]) << 32) | p[2 * i + 1];
    };
    std::size_t o = 0, w = 0, nf = 0, nl = 0;
    while (o < n_old || w < n_new) {
        if (w == n_new || (o < n_old && k(old_pairs, o) < k(new_pairs, w))) {
            lost[2 * nl] = old_pairs[2 * o];
            lost[2 * nl + 1] = old_pairs[2 * o + 1];
            ++nl;
            ++o;
        } else if (o == n_old || k(new_pairs, w) < k(old_pairs, o)) {
            found[2 * nf] = new_pairs[2 * w];
            found[2 * nf + 1] = new_pairs[2 * w + 1];
            ++nf;
            ++w;
        } else {
            ++o;
            ++w;
        }
    }
    *n_found = nf;
    *n_lost = nl;
}

}  // extern "C"
