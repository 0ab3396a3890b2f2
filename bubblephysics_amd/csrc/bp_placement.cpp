// bp_placement.cpp -- host-side random placement of non-overlapping discs.
//
// Spawning bubbles without overlap is the host's job in the reference (placement rules of the
// transaction stager, Sources/BubblePhysicsRuntime/PlacementGeometry.swift:42-50, and the
// rejection of a body over a bubble centre, TransactionStager.swift:424-428). This helper does
// random sequential addition with a uniform cell grid so that worlds of millions of bubbles can
// be laid out in seconds; it is used to build synthetic scenes.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <unordered_map>
#include <vector>

#include "../../include/bubblephysics.h"

namespace {

struct SplitMix {
    uint64_t s;
    uint64_t next() {
        uint64_t z = (s += 0x9e3779b97f4a7c15ull);
        z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
        z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
        return z ^ (z >> 31);
    }
    double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
};

}  // namespace

extern "C" {

// Places discs of the given radii (in the given order) uniformly at random inside
// region = (min.x, min.y, max.x, max.y) shrunk by `inset`, keeping every pair at least
// r_a + r_b + gap apart and every disc clear of the fixed circles `obstacles` (x, y, r triples).
// out_xy receives centres; discs that found no place after `attempts` tries get NaN.
// Returns the number placed in *out_placed.
bp_status bp_place_discs(const float *radii, uint32_t count, const float *region, float inset, float gap,
                         const float *obstacles, uint32_t obstacle_count, uint64_t seed, uint32_t attempts,
                         float *out_xy, uint32_t *out_placed) {
    if ((count && (!radii || !out_xy)) || !region || !out_placed || (obstacle_count && !obstacles)) return BP_ERR_INVALID_ARGUMENT;
    double rmax = 0.0;
    for (uint32_t i = 0; i < count; ++i) rmax = std::max(rmax, (double)radii[i]);
    for (uint32_t k = 0; k < obstacle_count; ++k) rmax = std::max(rmax, (double)obstacles[3 * k + 2]);
    double cell = std::max(2.0 * rmax + gap, 1e-3);
    auto key = [](int64_t x, int64_t y) { return (uint64_t)(x * 73856093) ^ ((uint64_t)y * 19349663ull << 1); };
    std::unordered_map<uint64_t, std::vector<uint32_t>> grid;
    grid.reserve((size_t)count + obstacle_count);
    std::vector<double> px, py, pr;
    px.reserve((size_t)count + obstacle_count); py.reserve(px.capacity()); pr.reserve(px.capacity());
    auto insert = [&](double x, double y, double r) {
        uint32_t id = (uint32_t)px.size();
        px.push_back(x); py.push_back(y); pr.push_back(r);
        grid[key((int64_t)std::floor(x / cell), (int64_t)std::floor(y / cell))].push_back(id);
    };
    for (uint32_t k = 0; k < obstacle_count; ++k) insert(obstacles[3 * k], obstacles[3 * k + 1], obstacles[3 * k + 2]);
    SplitMix rng{seed};
    uint32_t placed = 0;
    for (uint32_t i = 0; i < count; ++i) {
        double r = radii[i];
        double x0 = region[0] + inset + r, x1 = region[2] - inset - r;
        double y0 = region[1] + inset + r, y1 = region[3] - inset - r;
        out_xy[2 * i] = NAN; out_xy[2 * i + 1] = NAN;
        if (!(x1 >= x0) || !(y1 >= y0)) continue;
        for (uint32_t a = 0; a < attempts; ++a) {
            double x = x0 + (x1 - x0) * rng.uniform(), y = y0 + (y1 - y0) * rng.uniform();
            int64_t cx = (int64_t)std::floor(x / cell), cy = (int64_t)std::floor(y / cell);
            bool clear = true;
            for (int dx = -1; dx <= 1 && clear; ++dx)
                for (int dy = -1; dy <= 1 && clear; ++dy) {
                    auto it = grid.find(key(cx + dx, cy + dy));
                    if (it == grid.end()) continue;
                    for (uint32_t j : it->second) {
                        double ddx = x - px[j], ddy = y - py[j], need = r + pr[j] + gap;
                        if (ddx * ddx + ddy * ddy < need * need) { clear = false; break; }
                    }
                }
            if (!clear) continue;
            insert(x, y, r);
            out_xy[2 * i] = (float)x; out_xy[2 * i + 1] = (float)y;
            ++placed;
            break;
        }
    }
    *out_placed = placed;
    return BP_OK;
}

}  // extern "C"
