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
