// Host-side BVH construction for the MI355X tracer.
//
// build_canonical_bvh restates acoustic_agent/bvh.py:8-77 (the reference's
// deterministic builder): node bounds are the union of primitive bounds, the
// split axis is the argmax of the centroid extent (first axis on ties), the
// primitives are stably sorted by centroid on that axis and split at n/2,
// leaves (n <= leaf_size) store their ids in ascending order, nodes are laid
// out in preorder with escape links.  Identical inputs give the identical tree.
//
// collapse_to_wide turns that tree into the 4-wide, breadth-first device
// layout the HIP traversal reads (no reference counterpart: the reference
// walks the binary tree directly, steam_rt.py:4481-4506).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <deque>
#include <numeric>

#include "aa_internal.h"

namespace aa {

namespace {

struct Builder {
    const double* lower;
    const double* upper;
    std::vector<double> centroid;
    int64_t max_leaf;
    CanonicalBVH* out;

    void append(std::vector<int64_t>& idx, size_t begin, size_t end, int64_t depth) {
        const int64_t node = out->node_count();
        double lo[3] = {INFINITY, INFINITY, INFINITY};
        double hi[3] = {-INFINITY, -INFINITY, -INFINITY};
        for (size_t k = begin; k < end; ++k) {
            const int64_t p = idx[k];
            for (int a = 0; a < 3; ++a) {
                lo[a] = std::min(lo[a], lower[p * 3 + a]);
                hi[a] = std::max(hi[a], upper[p * 3 + a]);
            }
        }
        for (int a = 0; a < 3; ++a) {
            out->lo.push_back(lo[a]);
            out->hi.push_back(hi[a]);
        }
        out->start.push_back(-1);
        out->count.push_back(0);
        out->escape.push_back(-1);
        out->max_depth = std::max(out->max_depth, depth);
        const int64_t n = (int64_t)(end - begin);
        if (n <= max_leaf) {
            std::vector<int64_t> ordered(idx.begin() + (long)begin, idx.begin() + (long)end);
            std::sort(ordered.begin(), ordered.end());
            out->start[node] = (int64_t)out->prims.size();
            out->count[node] = n;
            out->prims.insert(out->prims.end(), ordered.begin(), ordered.end());
            out->leaf_count += 1;
        } else {
            double cmin[3] = {INFINITY, INFINITY, INFINITY};
            double cmax[3] = {-INFINITY, -INFINITY, -INFINITY};
            for (size_t k = begin; k < end; ++k) {
                for (int a = 0; a < 3; ++a) {
                    const double c = centroid[idx[k] * 3 + a];
                    cmin[a] = std::min(cmin[a], c);
                    cmax[a] = std::max(cmax[a], c);
                }
            }
            int axis = 0;
            double best = cmax[0] - cmin[0];
            for (int a = 1; a < 3; ++a) {
                if (cmax[a] - cmin[a] > best) {
                    best = cmax[a] - cmin[a];
                    axis = a;
                }
            }
            const double* cen = centroid.data();
            std::stable_sort(idx.begin() + (long)begin, idx.begin() + (long)end,
                             [cen, axis](int64_t x, int64_t y) { return cen[x * 3 + axis] < cen[y * 3 + axis]; });
            const size_t split = begin + (size_t)(n / 2);
            append(idx, begin, split, depth + 1);
            append(idx, split, end, depth + 1);
        }
        out->escape[node] = out->node_count();
    }
};

double half_area(const double* lo, const double* hi) {
    const double ex = hi[0] - lo[0], ey = hi[1] - lo[1], ez = hi[2] - lo[2];
    return ex * ey + ey * ez + ez * ex;
}

// Binned surface-area-heuristic builder producing the same preorder /
// escape layout as the canonical builder.
struct SahBuilder {
    static constexpr int kBins = 16;
    // Relative costs of a node visit and a triangle test.  On gfx950 a leaf's
    // triangles are tested in a loop that only the lanes holding that leaf
    // run (a few of 64), so a test costs the wave about as much as several
    // node visits; 4.0 gave the shortest traces of 1.0 / 2.0 / 4.0 on MI355X.
    static constexpr double kTraversalCost = 1.2;
    static constexpr double kTriangleCost = 4.0;
    const double* lower;
    const double* upper;
    std::vector<double> centroid;
    int64_t max_leaf;
    CanonicalBVH* out;

    void bounds_of(const std::vector<int64_t>& idx, size_t b, size_t e, double* lo, double* hi) const {
        for (int a = 0; a < 3; ++a) {
            lo[a] = INFINITY;
            hi[a] = -INFINITY;
        }
        for (size_t k = b; k < e; ++k)
            for (int a = 0; a < 3; ++a) {
                lo[a] = std::min(lo[a], lower[idx[k] * 3 + a]);
                hi[a] = std::max(hi[a], upper[idx[k] * 3 + a]);
            }
    }

    void make_leaf(std::vector<int64_t>& idx, size_t begin, size_t end, int64_t node) {
        std::vector<int64_t> ordered(idx.begin() + (long)begin, idx.begin() + (long)end);
        std::sort(ordered.begin(), ordered.end());
        out->start[node] = (int64_t)out->prims.size();
        out->count[node] = (int64_t)ordered.size();
        out->prims.insert(out->prims.end(), ordered.begin(), ordered.end());
        out->leaf_count += 1;
    }

    void append(std::vector<int64_t>& idx, size_t begin, size_t end, int64_t depth) {
        const int64_t node = out->node_count();
        double lo[3], hi[3];
        bounds_of(idx, begin, end, lo, hi);
        for (int a = 0; a < 3; ++a) {
            out->lo.push_back(lo[a]);
            out->hi.push_back(hi[a]);
        }
        out->start.push_back(-1);
        out->count.push_back(0);
        out->escape.push_back(-1);
        out->max_depth = std::max(out->max_depth, depth);
        const int64_t n = (int64_t)(end - begin);
        if (n <= 1) {
            make_leaf(idx, begin, end, node);
            out->escape[node] = out->node_count();
            return;
        }
        double cmin[3] = {INFINITY, INFINITY, INFINITY}, cmax[3] = {-INFINITY, -INFINITY, -INFINITY};
        for (size_t k = begin; k < end; ++k)
            for (int a = 0; a < 3; ++a) {
                cmin[a] = std::min(cmin[a], centroid[idx[k] * 3 + a]);
                cmax[a] = std::max(cmax[a], centroid[idx[k] * 3 + a]);
            }
        const double parent_area = std::max(half_area(lo, hi), 1e-300);
        double best_cost = INFINITY;
        int best_axis = -1, best_bin = -1;
        for (int a = 0; a < 3; ++a) {
            const double extent = cmax[a] - cmin[a];
            if (!(extent > 1e-12)) continue;
            int64_t cnt[kBins] = {0};
            double blo[kBins][3], bhi[kBins][3];
            for (int b = 0; b < kBins; ++b)
                for (int c = 0; c < 3; ++c) {
                    blo[b][c] = INFINITY;
                    bhi[b][c] = -INFINITY;
                }
            for (size_t k = begin; k < end; ++k) {
                const int64_t p = idx[k];
                int b = (int)((centroid[p * 3 + a] - cmin[a]) / extent * kBins);
                b = std::min(std::max(b, 0), kBins - 1);
                cnt[b] += 1;
                for (int c = 0; c < 3; ++c) {
                    blo[b][c] = std::min(blo[b][c], lower[p * 3 + c]);
                    bhi[b][c] = std::max(bhi[b][c], upper[p * 3 + c]);
                }
            }
            double right_area[kBins];
            int64_t right_cnt[kBins];
            double rlo[3] = {INFINITY, INFINITY, INFINITY}, rhi[3] = {-INFINITY, -INFINITY, -INFINITY};
            int64_t rc = 0;
            for (int b = kBins - 1; b > 0; --b) {
                rc += cnt[b];
                for (int c = 0; c < 3; ++c) {
                    rlo[c] = std::min(rlo[c], blo[b][c]);
                    rhi[c] = std::max(rhi[c], bhi[b][c]);
                }
                right_cnt[b] = rc;
                right_area[b] = rc ? half_area(rlo, rhi) : 0.0;
            }
            double llo[3] = {INFINITY, INFINITY, INFINITY}, lhi[3] = {-INFINITY, -INFINITY, -INFINITY};
            int64_t lc = 0;
            for (int b = 0; b < kBins - 1; ++b) {
                lc += cnt[b];
                for (int c = 0; c < 3; ++c) {
                    llo[c] = std::min(llo[c], blo[b][c]);
                    lhi[c] = std::max(lhi[c], bhi[b][c]);
                }
                if (lc == 0 || right_cnt[b + 1] == 0) continue;
                const double cost = kTraversalCost + kTriangleCost * (half_area(llo, lhi) * (double)lc +
                                                                      right_area[b + 1] * (double)right_cnt[b + 1]) /
                                                         parent_area;
                if (cost < best_cost) {
                    best_cost = cost;
                    best_axis = a;
                    best_bin = b;
                }
            }
        }
        const double leaf_cost = kTriangleCost * (double)n;
        if (n <= max_leaf && !(best_cost < leaf_cost)) {
            make_leaf(idx, begin, end, node);
            out->escape[node] = out->node_count();
            return;
        }
        size_t mid;
        if (best_axis >= 0) {
            const double extent = cmax[best_axis] - cmin[best_axis];
            const double* cen = centroid.data();
            const int axis = best_axis, split = best_bin;
            auto it = std::stable_partition(idx.begin() + (long)begin, idx.begin() + (long)end, [&](int64_t p) {
                int b = (int)((cen[p * 3 + axis] - cmin[axis]) / extent * kBins);
                b = std::min(std::max(b, 0), kBins - 1);
                return b <= split;
            });
            mid = (size_t)(it - idx.begin());
        } else {
            mid = begin + (size_t)(n / 2);  // coincident centroids: split the list
        }
        if (mid == begin || mid == end) mid = begin + (size_t)(n / 2);
        append(idx, begin, mid, depth + 1);
        append(idx, mid, end, depth + 1);
        out->escape[node] = out->node_count();
    }
};

float round_down(double x) {
    float f = (float)x;
    if ((double)f > x) f = std::nextafter(f, -INFINITY);
    return f;
}

float round_up(double x) {
    float f = (float)x;
    if ((double)f < x) f = std::nextafter(f, INFINITY);
    return f;
}

}  // namespace

std::string build_canonical_bvh(const double* lower, const double* upper, int64_t n, int64_t leaf_size,
                                CanonicalBVH& out) {
    out = CanonicalBVH();
    if (n < 1) return "BVH requires at least one surface";
    for (int64_t k = 0; k < n * 3; ++k) {
        if (!std::isfinite(lower[k]) || !std::isfinite(upper[k])) return "BVH bounds must be finite";
        if (upper[k] < lower[k]) return "BVH maximum bounds must not be below minimum bounds";
    }
    Builder b;
    b.lower = lower;
    b.upper = upper;
    b.max_leaf = std::max<int64_t>(1, leaf_size);
    b.centroid.resize((size_t)(n * 3));
    for (int64_t k = 0; k < n * 3; ++k) b.centroid[(size_t)k] = 0.5 * (lower[k] + upper[k]);
    b.out = &out;
    out.lo.reserve((size_t)(6 * n));
    out.hi.reserve((size_t)(6 * n));
    std::vector<int64_t> idx((size_t)n);
    std::iota(idx.begin(), idx.end(), 0);
    b.append(idx, 0, (size_t)n, 1);
    return std::string();
}

std::string build_sah_bvh(const double* lower, const double* upper, int64_t n, int64_t leaf_size, CanonicalBVH& out) {
    out = CanonicalBVH();
    if (n < 1) return "BVH requires at least one primitive";
    for (int64_t k = 0; k < n * 3; ++k) {
        if (!std::isfinite(lower[k]) || !std::isfinite(upper[k])) return "BVH bounds must be finite";
        if (upper[k] < lower[k]) return "BVH maximum bounds must not be below minimum bounds";
    }
    SahBuilder b;
    b.lower = lower;
    b.upper = upper;
    b.max_leaf = std::max<int64_t>(1, leaf_size);
    b.centroid.resize((size_t)(n * 3));
    for (int64_t k = 0; k < n * 3; ++k) b.centroid[(size_t)k] = 0.5 * (lower[k] + upper[k]);
    b.out = &out;
    std::vector<int64_t> idx((size_t)n);
    std::iota(idx.begin(), idx.end(), 0);
    b.append(idx, 0, (size_t)n, 1);
    return std::string();
}

void collapse_to_wide(const CanonicalBVH& bin, WideBVH& out) {
    out = WideBVH();
    struct Item {
        int64_t bin_node;
        int32_t wide;
        int32_t level;
    };
    auto is_leaf = [&](int64_t node) { return bin.count[(size_t)node] > 0; };
    auto left = [&](int64_t node) { return node + 1; };
    auto right = [&](int64_t node) { return bin.escape[(size_t)(node + 1)]; };

    auto fill_slot = [&](WideNode& w, int slot, int64_t child, int32_t code) {
        w.v[0][slot] = round_down(bin.lo[(size_t)child * 3 + 0]);
        w.v[1][slot] = round_up(bin.hi[(size_t)child * 3 + 0]);
        w.v[2][slot] = round_down(bin.lo[(size_t)child * 3 + 1]);
        w.v[3][slot] = round_up(bin.hi[(size_t)child * 3 + 1]);
        w.v[4][slot] = round_down(bin.lo[(size_t)child * 3 + 2]);
        w.v[5][slot] = round_up(bin.hi[(size_t)child * 3 + 2]);
        int32_t bits = code;
        std::memcpy(&w.v[6][slot], &bits, 4);
    };
    auto empty_slot = [&](WideNode& w, int slot) {
        for (int r = 0; r < 6; ++r) w.v[r][slot] = 1e30f;
        int32_t bits = kEmptyChild;
        std::memcpy(&w.v[6][slot], &bits, 4);
    };

    std::deque<Item> queue;
    out.nodes.emplace_back();
    std::memset(&out.nodes[0], 0, sizeof(WideNode));
    queue.push_back({0, 0, 0});
    int32_t current_level = 0;
    while (!queue.empty()) {
        Item it = queue.front();
        queue.pop_front();
        if (it.level != current_level) {
            out.level_end.push_back(it.wide);
            current_level = it.level;
        }
        std::vector<int64_t> kids;
        if (is_leaf(it.bin_node)) {
            kids.push_back(it.bin_node);  // only for a leaf root
        } else {
            kids.push_back(left(it.bin_node));
            kids.push_back(right(it.bin_node));
            while (kids.size() < 4) {
                int pick = -1;
                double best = -1.0;
                for (size_t k = 0; k < kids.size(); ++k) {
                    if (is_leaf(kids[k])) continue;
                    const double a = half_area(&bin.lo[(size_t)kids[k] * 3], &bin.hi[(size_t)kids[k] * 3]);
                    if (a > best) {
                        best = a;
                        pick = (int)k;
                    }
                }
                if (pick < 0) break;
                const int64_t node = kids[(size_t)pick];
                kids[(size_t)pick] = left(node);
                kids.insert(kids.begin() + pick + 1, right(node));
            }
        }
        for (int slot = 0; slot < 4; ++slot) {
            if (slot >= (int)kids.size()) {
                empty_slot(out.nodes[(size_t)it.wide], slot);
                continue;
            }
            const int64_t child = kids[(size_t)slot];
            if (is_leaf(child)) {
                fill_slot(out.nodes[(size_t)it.wide], slot, child,
                          encode_leaf(bin.start[(size_t)child], bin.count[(size_t)child]));
            } else {
                const int32_t wide = (int32_t)out.nodes.size();
                out.nodes.emplace_back();
                std::memset(&out.nodes.back(), 0, sizeof(WideNode));
                fill_slot(out.nodes[(size_t)it.wide], slot, child, wide);
                queue.push_back({child, wide, it.level + 1});
            }
        }
        out.depth = std::max(out.depth, it.level + 1);
    }
    out.level_end.push_back((int32_t)out.nodes.size());
}

}  // namespace aa
