// Host-side helper of the synthetic workloads: the sphere-rejection
// placement of the box pile (fyzz_amd/scenes.py pile_scene), in C++ for
// large N. Radii arrive in placement order. Same draws and acceptance
// rule as the Python loop:
//   attempt k draws u_a = uniform(seed, 3 + a, index k), a = 0..2,
//   center = ((u0 - 0.5) * footprint, 1 + u1 * height, (u2 - 0.5) * footprint),
//   accepted when no accepted center lies closer than r_b + r_o and
//   center.y - r_b > 0; accepted centers live in a dense grid of cubic
//   cells of side 2 sqrt(3) (at least twice the largest radius).
#include <cmath>
#include <cstdint>
#include <vector>

namespace {

std::uint64_t splitmix64(std::uint64_t x) {
    x += 0x9E3779B97F4A7C15ull;
    std::uint64_t z = x;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

double uniform(std::uint64_t seed, std::uint64_t stream, std::uint64_t index) {
    const std::uint64_t base = seed * 0x9E3779B97F4A7C15ull + stream * 0xD1B54A32D192ED03ull;
    return static_cast<double>(splitmix64(index + base) >> 11) * 0x1.0p-53;
}

}  // namespace

extern "C" int fyzz_scene_pile_centers(std::uint64_t seed, std::uint32_t n, const double* radius, double footprint,
                                       double height, double* out_pos, std::uint64_t max_attempts,
                                       std::uint64_t* attempts_used) {
    const double cell = 2.0 * std::sqrt(3.0);
    // cell index ranges of every possible center, plus one cell each side
    const std::int64_t x0 = static_cast<std::int64_t>(std::floor(-0.5 * footprint / cell)) - 1;
    const std::int64_t x1 = static_cast<std::int64_t>(std::floor(0.5 * footprint / cell)) + 1;
    const std::int64_t y0 = static_cast<std::int64_t>(std::floor(1.0 / cell)) - 1;
    const std::int64_t y1 = static_cast<std::int64_t>(std::floor((1.0 + height) / cell)) + 1;
    const std::int64_t nx = x1 - x0 + 1, ny = y1 - y0 + 1;
    if (nx <= 0 || ny <= 0 || nx * ny * nx > (1ll << 28)) return -2;
    std::vector<std::vector<std::uint32_t>> grid(static_cast<std::size_t>(nx * ny * nx));
    auto at = [&](std::int64_t x, std::int64_t y, std::int64_t z) -> std::vector<std::uint32_t>* {
        if (x < x0 || x > x1 || y < y0 || y > y1 || z < x0 || z > x1) return nullptr;
        return &grid[static_cast<std::size_t>(((x - x0) * ny + (y - y0)) * nx + (z - x0))];
    };
    std::uint64_t attempt = 0;
    for (std::uint32_t b = 0; b < n; ++b) {
        for (;;) {
            if (attempt >= max_attempts) return -1;
            const double u0 = uniform(seed, 3, attempt), u1 = uniform(seed, 4, attempt),
                         u2 = uniform(seed, 5, attempt);
            ++attempt;
            const double c[3] = {(u0 - 0.5) * footprint, 1.0 + u1 * height, (u2 - 0.5) * footprint};
            const std::int64_t k[3] = {static_cast<std::int64_t>(std::floor(c[0] / cell)),
                                       static_cast<std::int64_t>(std::floor(c[1] / cell)),
                                       static_cast<std::int64_t>(std::floor(c[2] / cell))};
            bool ok = c[1] - radius[b] > 0.0;
            for (int dx = -1; dx <= 1 && ok; ++dx)
                for (int dy = -1; dy <= 1 && ok; ++dy)
                    for (int dz = -1; dz <= 1 && ok; ++dz) {
                        const auto* list = at(k[0] + dx, k[1] + dy, k[2] + dz);
                        if (!list) continue;
                        for (std::uint32_t o : *list) {
                            const double d0 = c[0] - out_pos[3 * o], d1 = c[1] - out_pos[3 * o + 1],
                                         d2 = c[2] - out_pos[3 * o + 2];
                            const double rr = radius[b] + radius[o];
                            if (d0 * d0 + d1 * d1 + d2 * d2 < rr * rr) {
                                ok = false;
                                break;
                            }
                        }
                    }
            if (ok) {
                auto* list = at(k[0], k[1], k[2]);
                if (!list) return -2;
                out_pos[3 * b] = c[0];
                out_pos[3 * b + 1] = c[1];
                out_pos[3 * b + 2] = c[2];
                list->push_back(b);
                break;
            }
        }
    }
    if (attempts_used) *attempts_used = attempt;
    return 0;
}
