/*
 * Oracle (test infrastructure only; never linked into the product).
 *
 * Precision-independent pieces of the CPU restatement of the reference
 * reflection tracer:
 *   - the canonical binary BVH builder, restating acoustic_agent/bvh.py:8-77
 *     (median split on the largest centroid extent, stable argsort, leaf <= 4,
 *     preorder layout with stackless escape links);
 *   - Philox4x32-10 (Salmon et al., "Parallel random numbers: as easy as
 *     1, 2, 3", SC'11), the counter-based generator of the "counter" ray
 *     stream, and the per-ray / per-bounce draws built from it.
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load
 * this code, and only as the checker.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "aa_oracle.h"

/* ------------------------------------------------------------------ */
/* Philox4x32-10                                                       */
/* ------------------------------------------------------------------ */

void ao_philox4x32_10(const uint32_t ctr_in[4], const uint32_t key_in[2], uint32_t out[4])
{
    uint32_t c0 = ctr_in[0], c1 = ctr_in[1], c2 = ctr_in[2], c3 = ctr_in[3];
    uint32_t k0 = key_in[0], k1 = key_in[1];
    for (int round = 0; round < 10; ++round) {
        if (round > 0) {
            k0 += 0x9E3779B9u;
            k1 += 0xBB67AE85u;
        }
        uint64_t p0 = (uint64_t)0xD2511F53u * (uint64_t)c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * (uint64_t)c2;
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        uint32_t n0 = hi1 ^ c1 ^ k0;
        uint32_t n2 = hi0 ^ c3 ^ k1;
        c0 = n0;
        c1 = lo1;
        c2 = n2;
        c3 = lo0;
    }
    out[0] = c0;
    out[1] = c1;
    out[2] = c2;
    out[3] = c3;
}

/*
 * Counter stream layout (shared with the HIP kernels, which implement it
 * independently): key = (seed low word, seed high word);
 *   initial direction of ray i : counter (0xFFFFFFFF, i, 0, 0), word 0 -> jitter
 *   bounce k of ray i          : counter (k, i, 0, 0), word 0 -> U, word 1 -> J
 * U = (w >> 8) * 2^-24 (exact in binary32), J = (w * M) >> 32.
 * The jitter mirrors rng.uniform(-0.02, 0.02) of steam_rt.py:3964 and the
 * Fibonacci direction formula of steam_rt.py:3958-3968.
 */
#define AO_GOLDEN_QUARTER 0x30E44323405AC1F5ull /* round((3 - sqrt 5) / 4 * 2^64) */

void ao_counter_direction(uint64_t seed, int64_t ray, int64_t num_rays, double out[3])
{
    uint32_t ctr[4] = {0xFFFFFFFFu, (uint32_t)ray, 0u, 0u};
    uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    uint32_t r[4];
    ao_philox4x32_10(ctr, key, r);
    double u = (double)(r[0] >> 8) * (1.0 / 16777216.0);
    double jitter = -0.02 + 0.04 * u;
    double n = (double)(num_rays > 0 ? num_rays : 1);
    double i = (double)ray + 0.5;
    double z = 1.0 - 2.0 * i / n;
    double rr = 1.0 - z * z;
    if (rr < 0.0) rr = 0.0;
    if (rr > 1.0) rr = 1.0;
    double rad = sqrt(rr);
    /* golden-angle step (i + 1/2) pi (3 - sqrt 5), reduced modulo 2 pi
       exactly: 2 pi frac((2 ray + 1) K / 2^64) with K = (3 - sqrt 5) / 4
       in 64-bit fixed point (the product wraps modulo 2^64) */
    const uint64_t f = (2u * (uint64_t)ray + 1u) * AO_GOLDEN_QUARTER;
    double theta = 6.283185307179586 * ((double)(f >> 11) * 0x1p-53) + jitter;
    out[0] = cos(theta) * rad;
    out[1] = sin(theta) * rad;
    out[2] = z;
}

void ao_counter_bounce(uint64_t seed, int64_t ray, int32_t bounce, int32_t num_samples,
                       uint32_t *u24, int32_t *sample_index)
{
    uint32_t ctr[4] = {(uint32_t)bounce, (uint32_t)ray, 0u, 0u};
    uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    uint32_t r[4];
    ao_philox4x32_10(ctr, key, r);
    *u24 = r[0] >> 8;
    *sample_index = (int32_t)(((uint64_t)r[1] * (uint64_t)(uint32_t)num_samples) >> 32);
}

/* ------------------------------------------------------------------ */
/* Canonical BVH (acoustic_agent/bvh.py:8-77)                          */
/* ------------------------------------------------------------------ */

typedef struct {
    double key;
    int64_t pos;
    int64_t id;
} SortItem;

static int sort_item_cmp(const void *a, const void *b)
{
    const SortItem *x = (const SortItem *)a;
    const SortItem *y = (const SortItem *)b;
    if (x->key < y->key) return -1;
    if (x->key > y->key) return 1;
    /* equal keys keep their current order: a stable argsort (bvh.py:57) */
    return (x->pos > y->pos) - (x->pos < y->pos);
}

static int id_cmp(const void *a, const void *b)
{
    int64_t x = *(const int64_t *)a, y = *(const int64_t *)b;
    return (x > y) - (x < y);
}

typedef struct {
    const double *lower, *upper;
    double *centroids;
    int64_t max_leaf;
    double *node_min, *node_max;
    int64_t *start, *count, *escape, *prims;
    int64_t node_count, prim_count, leaf_count, max_depth;
    SortItem *scratch;
} BuildState;

static void append_node(BuildState *st, int64_t *indices, int64_t n, int64_t depth)
{
    int64_t node = st->node_count++;
    double lo[3] = {INFINITY, INFINITY, INFINITY};
    double hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int64_t k = 0; k < n; ++k) {
        int64_t p = indices[k];
        for (int a = 0; a < 3; ++a) {
            if (st->lower[p * 3 + a] < lo[a]) lo[a] = st->lower[p * 3 + a];
            if (st->upper[p * 3 + a] > hi[a]) hi[a] = st->upper[p * 3 + a];
        }
    }
    for (int a = 0; a < 3; ++a) {
        st->node_min[node * 3 + a] = lo[a];
        st->node_max[node * 3 + a] = hi[a];
    }
    st->start[node] = -1;
    st->count[node] = 0;
    st->escape[node] = -1;
    if (depth > st->max_depth) st->max_depth = depth;

    if (n <= st->max_leaf) {
        int64_t tmp[64];
        int64_t *ordered = n <= 64 ? tmp : (int64_t *)malloc(sizeof(int64_t) * (size_t)n);
        memcpy(ordered, indices, sizeof(int64_t) * (size_t)n);
        qsort(ordered, (size_t)n, sizeof(int64_t), id_cmp);
        st->start[node] = st->prim_count;
        st->count[node] = n;
        for (int64_t k = 0; k < n; ++k) st->prims[st->prim_count++] = ordered[k];
        st->leaf_count++;
        if (ordered != tmp) free(ordered);
    } else {
        double cmin[3] = {INFINITY, INFINITY, INFINITY};
        double cmax[3] = {-INFINITY, -INFINITY, -INFINITY};
        for (int64_t k = 0; k < n; ++k) {
            int64_t p = indices[k];
            for (int a = 0; a < 3; ++a) {
                double c = st->centroids[p * 3 + a];
                if (c < cmin[a]) cmin[a] = c;
                if (c > cmax[a]) cmax[a] = c;
            }
        }
        int axis = 0;
        double best = cmax[0] - cmin[0];
        for (int a = 1; a < 3; ++a) {
            double e = cmax[a] - cmin[a];
            if (e > best) { /* np.argmax keeps the first maximum */
                best = e;
                axis = a;
            }
        }
        SortItem *items = st->scratch;
        for (int64_t k = 0; k < n; ++k) {
            items[k].key = st->centroids[indices[k] * 3 + axis];
            items[k].pos = k;
            items[k].id = indices[k];
        }
        qsort(items, (size_t)n, sizeof(SortItem), sort_item_cmp);
        for (int64_t k = 0; k < n; ++k) indices[k] = items[k].id;
        int64_t split = n / 2;
        append_node(st, indices, split, depth + 1);
        append_node(st, indices + split, n - split, depth + 1);
    }
    st->escape[node] = st->node_count;
}

int ao_build_bvh(const double *lower, const double *upper, int64_t n, int64_t leaf_size,
                 double *node_min, double *node_max, int64_t *start, int64_t *count,
                 int64_t *escape, int64_t *prims, int64_t *stats)
{
    if (n < 1) return -1;
    for (int64_t k = 0; k < n * 3; ++k) {
        if (!isfinite(lower[k]) || !isfinite(upper[k])) return -2;
        if (upper[k] < lower[k]) return -3;
    }
    BuildState st;
    memset(&st, 0, sizeof(st));
    st.lower = lower;
    st.upper = upper;
    st.max_leaf = leaf_size > 1 ? leaf_size : 1;
    st.centroids = (double *)malloc(sizeof(double) * (size_t)(n * 3));
    st.scratch = (SortItem *)malloc(sizeof(SortItem) * (size_t)n);
    int64_t *indices = (int64_t *)malloc(sizeof(int64_t) * (size_t)n);
    if (!st.centroids || !st.scratch || !indices) {
        free(st.centroids);
        free(st.scratch);
        free(indices);
        return -4;
    }
    for (int64_t k = 0; k < n * 3; ++k) st.centroids[k] = 0.5 * (lower[k] + upper[k]);
    for (int64_t k = 0; k < n; ++k) indices[k] = k;
    st.node_min = node_min;
    st.node_max = node_max;
    st.start = start;
    st.count = count;
    st.escape = escape;
    st.prims = prims;
    append_node(&st, indices, n, 1);
    stats[0] = st.node_count;
    stats[1] = st.leaf_count;
    stats[2] = st.max_depth;
    free(st.centroids);
    free(st.scratch);
    free(indices);
    return 0;
}
