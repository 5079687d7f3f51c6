// Oracle: CPU restatement of the reference NativeCpu hot path.
//
// TEST INFRASTRUCTURE ONLY. This library is the checker: tests/, the
// smoke entry and the CPU-baseline leg of bench.py may load it; the
// product never links, loads or falls back to it.
//
// Every function cites the reference file:line it restates (paths are
// relative to the reference repository root). Layout conventions shared
// with the device C-ABI:
//   * boxes are indexed by slot 0..n-1 in increasing shape-id order;
//   * encoded bounds are six planes of n u32: plane k<3 is the min key of
//     axis k, plane 3+k the max key (scene.h Shape::encoded_bounds order);
//   * an endpoint payload is (slot << 1) | is_max;
//   * a pair is two u32 shape ids (a < b), sorted lexicographically.
#pragma once
#include <cstddef>
#include <cstdint>

extern "C" {

// ---- SAP (broadphase) -------------------------------------------------
// Encoded AABB of n boxes: pos[3n], quat[4n] (x,y,z,w), half[3n],
// contact offset[n] -> bounds[6][n].   scene.cpp:26-35, 1233-1316.
int orc_box_bounds(std::size_t n, const float* pos, const float* quat,
                   const float* half, const float* contact_offset,
                   std::uint32_t* bounds);

// Sorted endpoints of one axis: value[2n], payload[2n], ordered by
// (value, min-before-max, shape id).   scene.cpp:1341-1346, 1394-1400.
void orc_sort_endpoints(std::size_t n, const std::uint32_t* bounds, unsigned axis,
                        std::uint32_t* value, std::uint32_t* payload);

// Full overlap pair set by the x-axis sweep over sorted endpoints, with the
// dynamic filter and the plane Cartesian product.   scene.cpp:2642-2703.
// shape_id[n], body_id[n], dynamic[n] per slot; planes: plane_ids[n_planes]
// with their body ids plane_body[n_planes]. Returns #pairs (writes at most
// cap pairs to out[2*cap]).
std::size_t orc_sap_pairs(std::size_t n, const std::uint32_t* bounds,
                          const std::uint32_t* shape_id, const std::uint32_t* body_id,
                          const std::uint8_t* dynamic, std::size_t n_planes,
                          const std::uint32_t* plane_ids, std::uint32_t* out,
                          std::size_t cap);

// Found/lost between two sorted pair lists (publish_pairs merge,
// scene.cpp:3230-3296). Returns counts through n_found / n_lost.
void orc_found_lost(std::size_t n_old, const std::uint32_t* old_pairs,
                    std::size_t n_new, const std::uint32_t* new_pairs,
                    std::uint32_t* found, std::size_t* n_found,
                    std::uint32_t* lost, std::size_t* n_lost);

}  // extern "C"

#ifdef __cplusplus
#include <thread>
#include <vector>
namespace orc {
// Runs fn(begin, end) over [0, n) split into `threads` contiguous ranges
// on host threads (inline when threads <= 1 or n is small). The work items
// must be independent.
template <class F>
inline void parallel_ranges(std::uint32_t n, std::uint32_t threads, F&& fn) {
    if (threads <= 1 || n < 1024) {
        fn(0u, n);
        return;
    }
    std::vector<std::thread> pool;
    const std::uint32_t step = (n + threads - 1) / threads;
    for (std::uint32_t b = 0; b < n; b += step) pool.emplace_back(fn, b, b + step < n ? b + step : n);
    for (auto& t : pool) t.join();
}
}  // namespace orc
#endif
