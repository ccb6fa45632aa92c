/*
 * Oracle (test infrastructure only): CPU restatement of the reference
 * reflection tracer.  Compiled twice, with -DAO_FP32 (binary32, mirrors the
 * reference FP32 CUDA kernel steam_rt_cuda.py:491-710 and its epsilons) and
 * without (binary64, the reference Numba kernel steam_rt.py:4708-4886).
 *
 * Function-by-function citations (reference file:line):
 *   point_in_polygon   steam_rt.py:4258-4274
 *   box_hit            steam_rt.py:4277-4358
 *   surface_hit        steam_rt.py:4361-4392  (analytic kinds 0/1/2)
 *   tri_hit            no reference counterpart: Moller-Trumbore test of the
 *                      triangle tessellation of the same surfaces; it carries
 *                      the reference surface id of the primitive it came from.
 *   aabb_hit           steam_rt.py:4406-4426 (steam_rt_cuda.py:412-428 for FP32)
 *   closest_hit        steam_rt.py:4429-4478 (linear) and 4481-4506 (BVH);
 *                      tie rule steam_rt.py:4499, orientation 4395-4403
 *   any_hit            steam_rt.py:4517-4520 / 4788-4792: the closest-hit
 *                      shadow test `EPS < t < D - EPS` equals an any-hit query
 *                      over (EPS, D - EPS); shadow_closest = 1 runs it literally.
 *   sphere_before      steam_rt.py:4522-4533
 *   normalize3         steam_rt.py:4556-4561
 *   diffuse_direction  steam_rt.py:4564-4582
 *   ao_trace           steam_rt.py:4708-4886 (bounce loop 4743-4848,
 *                      NEE + deposit 4770-4826, reduction 4851-4871); the
 *                      energy expression follows the operand order of the FP32
 *                      kernel steam_rt_cuda.py:663-677 in both precisions.
 *                      Visual candidates (hit points, surfaces, strongest
 *                      contribution) follow steam_rt.py:4718-4731, 4765-4769
 *                      and 4821-4824.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "aa_oracle.h"

#ifdef AO_FP32
#define REAL float
#define SUF f32
#define RS(x) ((float)(x))
#define RSQRT sqrtf
#define RPOW powf
#define RFLOOR floorf
#define RABS fabsf
#define RFMA fmaf
#define BVH_EPS 1.0e-5f
#define PAR_EPS 1.0e-12f
#else
#define REAL double
#define SUF f64
#define RS(x) ((double)(x))
#define RSQRT sqrt
#define RPOW pow
#define RFLOOR floor
#define RABS fabs
#define RFMA fma
#define BVH_EPS 2.0e-6
#define PAR_EPS 1.0e-15
#endif

#define CAT2(a, b) a##_##b
#define CAT(a, b) CAT2(a, b)
#define SCENE CAT(ao_scene, SUF)
#define PARAMS CAT(ao_params, SUF)
#define OUTS CAT(ao_out, SUF)

#define EPS RS(1e-6)
#define HIT_OFFSET RS(1e-2)
#define BIG RS(1e30)
#define NB AO_NUM_BANDS

typedef struct {
    int64_t nodes, prims;
} Counters;

static int point_in_polygon(REAL x, REAL y, const REAL *corners, int n)
{
    int inside = 0;
    int j = n - 1;
    for (int i = 0; i < n; ++i) {
        REAL xi = corners[i * 2], yi = corners[i * 2 + 1];
        REAL xj = corners[j * 2], yj = corners[j * 2 + 1];
        if ((yi > y) != (yj > y)) {
            REAL denom = yj - yi;
            if (RABS(denom) > RS(1e-12)) {
                REAL x_cross = (xj - xi) * (y - yi) / denom + xi;
                if (x < x_cross) inside = !inside;
            }
        }
        j = i;
    }
    return inside;
}

/* Returns the hit distance (BIG on a miss) and the face normal. */
static REAL box_hit(const REAL o_[3], const REAL d_[3], const REAL *center, const REAL *axis_u,
                    const REAL *axis_v, const REAL *half, const REAL *zr, REAL n_out[3])
{
    REAL relx = o_[0] - center[0];
    REAL rely = o_[1] - center[1];
    REAL ox = relx * axis_u[0] + rely * axis_u[1];
    REAL oy = relx * axis_v[0] + rely * axis_v[1];
    REAL oz = o_[2];
    REAL dx = d_[0] * axis_u[0] + d_[1] * axis_u[1];
    REAL dy = d_[0] * axis_v[0] + d_[1] * axis_v[1];
    REAL dz = d_[2];
    REAL t_min = -BIG, t_max = BIG;
    REAL nx = 0, ny = 0, nz = 0, ex_nx = 0, ex_ny = 0, ex_nz = 0;
    for (int axis_i = 0; axis_i < 3; ++axis_i) {
        REAL o, d, lo, hi, ax, ay, az;
        if (axis_i == 0) {
            o = ox; d = dx; lo = -half[0]; hi = half[0]; ax = axis_u[0]; ay = axis_u[1]; az = 0;
        } else if (axis_i == 1) {
            o = oy; d = dy; lo = -half[1]; hi = half[1]; ax = axis_v[0]; ay = axis_v[1]; az = 0;
        } else {
            o = oz; d = dz; lo = zr[0]; hi = zr[1]; ax = 0; ay = 0; az = 1;
        }
        if (RABS(d) <= RS(1e-12)) {
            if (o < lo || o > hi) return BIG;
            continue;
        }
        REAL enter, exit_, en_x, en_y, en_z, out_x, out_y, out_z;
        if (d > 0) {
            enter = (lo - o) / d; exit_ = (hi - o) / d;
            en_x = -ax; en_y = -ay; en_z = -az; out_x = ax; out_y = ay; out_z = az;
        } else {
            enter = (hi - o) / d; exit_ = (lo - o) / d;
            en_x = ax; en_y = ay; en_z = az; out_x = -ax; out_y = -ay; out_z = -az;
        }
        if (enter > t_min) { t_min = enter; nx = en_x; ny = en_y; nz = en_z; }
        if (exit_ < t_max) { t_max = exit_; ex_nx = out_x; ex_ny = out_y; ex_nz = out_z; }
        if (t_min > t_max) return BIG;
    }
    if (t_max <= EPS) return BIG;
    if (t_min > EPS) {
        n_out[0] = nx; n_out[1] = ny; n_out[2] = nz;
        return t_min;
    }
    n_out[0] = ex_nx; n_out[1] = ex_ny; n_out[2] = ex_nz;
    return t_max;
}

static REAL surface_hit(const SCENE *sc, int si, const REAL o[3], const REAL d[3], REAL n[3])
{
    REAL t = BIG;
    n[0] = sc->normals[si * 3];
    n[1] = sc->normals[si * 3 + 1];
    n[2] = sc->normals[si * 3 + 2];
    int64_t kind = sc->kinds[si];
    if (kind == 0) {
        REAL sx = sc->wall_delta[si * 2], sy = sc->wall_delta[si * 2 + 1];
        REAL det = sx * d[1] - sy * d[0];
        if (RABS(det) > RS(1e-12)) {
            REAL relx = o[0] - sc->wall_a[si * 2];
            REAL rely = o[1] - sc->wall_a[si * 2 + 1];
            REAL cand_t = (relx * sy - rely * sx) / det;
            REAL u = (relx * d[1] - rely * d[0]) / det;
            REAL z = o[2] + cand_t * d[2];
            if (cand_t > EPS && u >= RS(-1e-6) && u <= RS(1.0 + 1e-6) &&
                z >= sc->wall_z[si * 2] - RS(1e-6) && z <= sc->wall_z[si * 2 + 1] + RS(1e-6))
                t = cand_t;
        }
    } else if (kind == 2) {
        REAL bn[3];
        REAL cand_t = box_hit(o, d, sc->box_center + si * 2, sc->box_axis_u + si * 2,
                              sc->box_axis_v + si * 2, sc->box_half + si * 2, sc->box_z + si * 2, bn);
        if (cand_t > EPS && cand_t < RS(1e29)) {
            t = cand_t;
            n[0] = bn[0]; n[1] = bn[1]; n[2] = bn[2];
        }
    } else if (RABS(d[2]) > RS(1e-12)) {
        REAL cand_t = (sc->z_values[si] - o[2]) / d[2];
        REAL px = o[0] + cand_t * d[0];
        REAL py = o[1] + cand_t * d[1];
        if (cand_t > EPS && point_in_polygon(px, py, sc->corners, sc->num_corners)) t = cand_t;
    }
    return t;
}

/* Moller-Trumbore, double-sided, barycentric tolerance 1e-6 (the reference's
 * wall-edge tolerance, steam_rt.py:4380).  Cross products and dot products
 * are evaluated with fused multiply-adds: a*b - c*d as fma(a, b, -(c*d)) and
 * a.b as fma(ax, bx, fma(ay, by, az*bz)).  The HIP kernel uses the same
 * operations in the same order, so the binary32 build reproduces it
 * exactly. */
static REAL tri_hit(const SCENE *sc, int ti, const REAL o[3], const REAL d[3])
{
    const REAL *v = sc->tri_v + (size_t)ti * 9;
    REAL e1x = v[3] - v[0], e1y = v[4] - v[1], e1z = v[5] - v[2];
    REAL e2x = v[6] - v[0], e2y = v[7] - v[1], e2z = v[8] - v[2];
    REAL px = RFMA(d[1], e2z, -(d[2] * e2y));
    REAL py = RFMA(d[2], e2x, -(d[0] * e2z));
    REAL pz = RFMA(d[0], e2y, -(d[1] * e2x));
    REAL det = RFMA(e1x, px, RFMA(e1y, py, e1z * pz));
    if (!(RABS(det) > RS(1e-12))) return BIG;
    REAL inv = RS(1) / det;
    REAL sx = o[0] - v[0], sy = o[1] - v[1], sz = o[2] - v[2];
    REAL u = RFMA(sx, px, RFMA(sy, py, sz * pz)) * inv;
    if (u < RS(-1e-6) || u > RS(1.0 + 1e-6)) return BIG;
    REAL qx = RFMA(sy, e1z, -(sz * e1y));
    REAL qy = RFMA(sz, e1x, -(sx * e1z));
    REAL qz = RFMA(sx, e1y, -(sy * e1x));
    REAL vv = RFMA(d[0], qx, RFMA(d[1], qy, d[2] * qz)) * inv;
    if (vv < RS(-1e-6) || u + vv > RS(1.0 + 1e-6)) return BIG;
    REAL t = RFMA(e2x, qx, RFMA(e2y, qy, e2z * qz)) * inv;
    return t > EPS ? t : BIG;
}

/* One primitive: t (BIG on a miss), its surface id and unoriented normal. */
static REAL prim_hit(const SCENE *sc, int prim, const REAL o[3], const REAL d[3], int *surf, REAL n[3])
{
    if (sc->mode == 1) {
        *surf = sc->tri_surface[prim];
        n[0] = sc->tri_normal[prim * 3];
        n[1] = sc->tri_normal[prim * 3 + 1];
        n[2] = sc->tri_normal[prim * 3 + 2];
        return tri_hit(sc, prim, o, d);
    }
    *surf = prim;
    return surface_hit(sc, prim, o, d, n);
}

static int aabb_hit(const REAL o[3], const REAL d[3], const REAL *lo, const REAL *hi, REAL max_distance)
{
    REAL enter = 0;
    REAL exit_ = max_distance + BVH_EPS;
    for (int a = 0; a < 3; ++a) {
        REAL value = d[a];
        if (RABS(value) <= PAR_EPS) {
            if (o[a] < lo[a] || o[a] > hi[a]) return 0;
            continue;
        }
        REAL first = (lo[a] - o[a]) / value;
        REAL second = (hi[a] - o[a]) / value;
        if (first > second) {
            REAL tmp = first;
            first = second;
            second = tmp;
        }
        if (first > enter) enter = first;
        if (second < exit_) exit_ = second;
        if (enter > exit_) return 0;
    }
    return exit_ > EPS;
}

static inline int better(REAL t, int surf, int prim, REAL best_t, int best_surf, int best_prim)
{
    if (t < best_t) return 1;
    if (t == best_t && t < BIG) {
        if (best_surf < 0 || surf < best_surf) return 1;
        if (surf == best_surf && prim < best_prim) return 1;
    }
    return 0;
}

static void closest_hit(const SCENE *sc, const REAL o[3], const REAL d[3], int *best_prim, int *best_surf,
                        REAL *best_t_out, REAL bn[3], Counters *cnt)
{
    REAL best_t = BIG;
    int bs = -1, bp = -1;
    REAL n[3], tn[3] = {0, 0, 0};
    if (sc->use_bvh) {
        int node = 0;
        while (node < sc->num_nodes) {
            cnt->nodes++;
            if (!aabb_hit(o, d, sc->bvh_min + node * 3, sc->bvh_max + node * 3, best_t)) {
                node = (int)sc->bvh_escape[node];
                continue;
            }
            int64_t count = sc->bvh_count[node];
            if (count > 0) {
                int64_t start = sc->bvh_start[node];
                for (int64_t k = 0; k < count; ++k) {
                    int prim = (int)sc->bvh_prims[start + k];
                    int surf;
                    cnt->prims++;
                    REAL t = prim_hit(sc, prim, o, d, &surf, n);
                    if (better(t, surf, prim, best_t, bs, bp)) {
                        best_t = t; bs = surf; bp = prim;
                        tn[0] = n[0]; tn[1] = n[1]; tn[2] = n[2];
                    }
                }
            }
            node += 1;
        }
    } else {
        int count = sc->mode == 1 ? sc->num_triangles : sc->num_surfaces;
        for (int prim = 0; prim < count; ++prim) {
            int surf;
            cnt->prims++;
            REAL t = prim_hit(sc, prim, o, d, &surf, n);
            if (better(t, surf, prim, best_t, bs, bp)) {
                best_t = t; bs = surf; bp = prim;
                tn[0] = n[0]; tn[1] = n[1]; tn[2] = n[2];
            }
        }
    }
    if (bs >= 0) {
        REAL dot = tn[0] * d[0] + tn[1] * d[1] + tn[2] * d[2];
        if (dot > 0) {
            tn[0] = -tn[0]; tn[1] = -tn[1]; tn[2] = -tn[2];
        }
    }
    *best_prim = bp;
    *best_surf = bs;
    *best_t_out = best_t;
    bn[0] = tn[0]; bn[1] = tn[1]; bn[2] = tn[2];
}

static int any_hit(const SCENE *sc, const REAL o[3], const REAL d[3], REAL max_distance, Counters *cnt)
{
    REAL n[3];
    int surf;
    if (sc->use_bvh) {
        int node = 0;
        while (node < sc->num_nodes) {
            cnt->nodes++;
            if (!aabb_hit(o, d, sc->bvh_min + node * 3, sc->bvh_max + node * 3, max_distance)) {
                node = (int)sc->bvh_escape[node];
                continue;
            }
            int64_t count = sc->bvh_count[node];
            if (count > 0) {
                int64_t start = sc->bvh_start[node];
                for (int64_t k = 0; k < count; ++k) {
                    cnt->prims++;
                    REAL t = prim_hit(sc, (int)sc->bvh_prims[start + k], o, d, &surf, n);
                    if (t > EPS && t < max_distance - EPS) return 1;
                }
            }
            node += 1;
        }
        return 0;
    }
    int count = sc->mode == 1 ? sc->num_triangles : sc->num_surfaces;
    for (int prim = 0; prim < count; ++prim) {
        cnt->prims++;
        REAL t = prim_hit(sc, prim, o, d, &surf, n);
        if (t > EPS && t < max_distance - EPS) return 1;
    }
    return 0;
}

static int sphere_before(const REAL o[3], const REAL d[3], const REAL c[3], REAL radius, REAL max_distance)
{
    REAL ox = o[0] - c[0], oy = o[1] - c[1], oz = o[2] - c[2];
    REAL b = ox * d[0] + oy * d[1] + oz * d[2];
    REAL cc = ox * ox + oy * oy + oz * oz - radius * radius;
    REAL disc = b * b - cc;
    if (disc < 0) return 0;
    REAL root = -b - RSQRT(disc > 0 ? disc : RS(0));
    return root >= 0 && root < max_distance;
}

static void normalize3(REAL *x, REAL *y, REAL *z)
{
    REAL n = RSQRT(*x * *x + *y * *y + *z * *z);
    if (n <= RS(1e-12)) {
        *x = 0; *y = 0; *z = 0;
        return;
    }
    *x = *x / n;
    *y = *y / n;
    *z = *z / n;
}

static void diffuse_direction(const REAL s[3], REAL nx, REAL ny, REAL nz, REAL out[3])
{
    REAL hx = 1, hy = 0, hz = 0;
    if (RABS(nx) >= RS(0.9)) {
        hx = 0;
        hy = 1;
    }
    REAL tx = hy * nz - hz * ny;
    REAL ty = hz * nx - hx * nz;
    REAL tz = hx * ny - hy * nx;
    normalize3(&tx, &ty, &tz);
    REAL bx = ny * tz - nz * ty;
    REAL by = nz * tx - nx * tz;
    REAL bz = nx * ty - ny * tx;
    REAL dx = s[0] * tx + s[1] * bx + s[2] * nx;
    REAL dy = s[0] * ty + s[1] * by + s[2] * ny;
    REAL dz = s[0] * tz + s[1] * bz + s[2] * nz;
    normalize3(&dx, &dy, &dz);
    out[0] = dx; out[1] = dy; out[2] = dz;
}

void CAT(ao_diffuse_direction, SUF)(const REAL sample[3], const REAL normal[3], REAL out[3])
{
    diffuse_direction(sample, normal[0], normal[1], normal[2], out);
}

int CAT(ao_closest_hit, SUF)(const SCENE *sc, const REAL origin[3], const REAL direction[3], int32_t *prim,
                             int32_t *surface, REAL *t, REAL normal[3])
{
    Counters cnt = {0, 0};
    int p, s;
    closest_hit(sc, origin, direction, &p, &s, t, normal, &cnt);
    *prim = p;
    *surface = s;
    return 0;
}

typedef struct {
    REAL *echogram, *ambisonic, *surface_energy;
    int64_t *hit_counts, *contrib_counts;
    int64_t actual_bounces, active_count, bounce_queries, shadow_queries;
    int64_t contributions, escaped;
    Counters cb, cs;
} Local;

int CAT(ao_trace, SUF)(const SCENE *sc, const PARAMS *p, OUTS *out)
{
    const int N = p->num_rays, B = p->num_bounces, K = p->num_bins;
    const int S = sc->num_surfaces;
    const int M = p->num_diffuse_samples;
    if (N <= 0 || B < 0 || K <= 0 || S <= 0 || M <= 0) return -1;
    if (p->rng_mode == 0 && (!p->directions || !p->diffuse_random || !p->diffuse_indices)) return -2;
    int ray_begin = p->ray_begin, ray_end = p->ray_end > 0 ? p->ray_end : N;
    if (ray_begin < 0 || ray_end > N || ray_begin > ray_end) return -3;
    int nthreads = 1;
#ifdef _OPENMP
    nthreads = p->num_threads > 0 ? p->num_threads : omp_get_max_threads();
#endif
    Local *loc = (Local *)calloc((size_t)nthreads, sizeof(Local));
    const int amb = p->render_ambisonics && out->ambisonic;
    for (int t = 0; t < nthreads; ++t) {
        loc[t].echogram = (REAL *)calloc((size_t)NB * K, sizeof(REAL));
        loc[t].ambisonic = amb ? (REAL *)calloc((size_t)NB * 4 * K, sizeof(REAL)) : NULL;
        loc[t].surface_energy = (REAL *)calloc((size_t)S, sizeof(REAL));
        loc[t].hit_counts = (int64_t *)calloc((size_t)S, sizeof(int64_t));
        loc[t].contrib_counts = (int64_t *)calloc((size_t)S, sizeof(int64_t));
    }
    const REAL ray_scale = RS(4.0 * M_PI) / (REAL)(N > 1 ? N : 1);
    const REAL INV_PI = RS(1.0 / M_PI), INV_4PI = RS(1.0 / (4.0 * M_PI)), INV_8PI = RS(1.0 / (8.0 * M_PI));
    const REAL *L = p->listener, *Sp = p->source;
    const REAL source_distance =
        RSQRT((L[0] - Sp[0]) * (L[0] - Sp[0]) + (L[1] - Sp[1]) * (L[1] - Sp[1]) + (L[2] - Sp[2]) * (L[2] - Sp[2]));

    /* visual candidates (steam_rt.py:4718-4731): ray ri % stride == 0 owns
       slot ri / stride while that is below the limit */
    const int V = (out->visual_points && p->visual_limit > 0) ? p->visual_limit : 0;
    const int vstride = p->visual_stride > 0 ? p->visual_stride : 1;
    for (int v = 0; v < V; ++v) {
        for (int b = 0; b < B; ++b) {
            out->visual_points[((size_t)v * B + b) * 3] = 0;
            out->visual_points[((size_t)v * B + b) * 3 + 1] = 0;
            out->visual_points[((size_t)v * B + b) * 3 + 2] = 0;
            out->visual_surfaces[(size_t)v * B + b] = -1;
        }
        out->visual_rays[v] = -1;
        out->visual_orders[v] = 0;
        out->visual_distances[v] = 0;
        out->visual_gains[v] = 0;
    }

#ifdef _OPENMP
#pragma omp parallel for schedule(static) num_threads(nthreads)
#endif
    for (int ri = ray_begin; ri < ray_end; ++ri) {
        int tid = 0;
#ifdef _OPENMP
        tid = omp_get_thread_num();
#endif
        Local *lc = &loc[tid];
        REAL origin[3] = {L[0], L[1], L[2]};
        REAL dir[3], dir0[3];
        if (p->rng_mode == 0) {
            dir[0] = p->directions[(size_t)ri * 3];
            dir[1] = p->directions[(size_t)ri * 3 + 1];
            dir[2] = p->directions[(size_t)ri * 3 + 2];
        } else {
            double dd[3];
            ao_counter_direction(p->seed, ri, N, dd);
            dir[0] = (REAL)dd[0];
            dir[1] = (REAL)dd[1];
            dir[2] = (REAL)dd[2];
        }
        dir0[0] = dir[0]; dir0[1] = dir[1]; dir0[2] = dir[2];
        REAL accum_distance = 0;
        REAL accum_energy[NB];
        for (int b = 0; b < NB; ++b) accum_energy[b] = 1;
        int alive = 1;
        const int logging = out->log_prim && ri < p->log_rays;
        const int vslot = (V > 0 && ri % vstride == 0 && ri / vstride < V) ? ri / vstride : -1;
        if (vslot >= 0) out->visual_rays[vslot] = ri;
        for (int bounce = 0; bounce < B; ++bounce) {
            if (bounce + 1 > lc->actual_bounces) lc->actual_bounces = bounce + 1;
            int prim, surf;
            REAL t, n[3];
            lc->bounce_queries++;
            closest_hit(sc, origin, dir, &prim, &surf, &t, n, &lc->cb);
            if (logging && bounce < p->log_bounces) {
                size_t li = (size_t)ri * p->log_bounces + bounce;
                out->log_prim[li] = prim;
                out->log_surface[li] = surf;
                out->log_t[li] = (double)t;
                if (out->log_point) {
                    /* the hit point before the offset (no hit: t = BIG, logged as is) */
                    out->log_point[li * 3] = (double)(origin[0] + t * dir[0]);
                    out->log_point[li * 3 + 1] = (double)(origin[1] + t * dir[1]);
                    out->log_point[li * 3 + 2] = (double)(origin[2] + t * dir[2]);
                }
            }
            if (surf < 0) lc->escaped++;
            if (surf < 0 || t <= p->listener_radius || accum_distance > p->max_path_len) {
                alive = 0;
                break;
            }
            if (bounce > 0) {
                if (sphere_before(origin, dir, L, p->listener_radius, t)) {
                    alive = 0;
                    break;
                }
                if (source_distance > p->source_radius && sphere_before(origin, dir, Sp, p->source_radius, t)) {
                    alive = 0;
                    break;
                }
            }
            lc->hit_counts[surf] += 1;
            REAL hx = origin[0] + t * dir[0] + HIT_OFFSET * n[0];
            REAL hy = origin[1] + t * dir[1] + HIT_OFFSET * n[1];
            REAL hz = origin[2] + t * dir[2] + HIT_OFFSET * n[2];
            if (vslot >= 0) {
                REAL *vp = out->visual_points + ((size_t)vslot * B + bounce) * 3;
                vp[0] = hx;
                vp[1] = hy;
                vp[2] = hz;
                out->visual_surfaces[(size_t)vslot * B + bounce] = surf;
            }
            REAL tsx = Sp[0] - hx, tsy = Sp[1] - hy, tsz = Sp[2] - hz;
            REAL dist_to_source = RSQRT(tsx * tsx + tsy * tsy + tsz * tsz);
            if (dist_to_source > p->irradiance_min_distance) {
                REAL sdx = tsx / dist_to_source, sdy = tsy / dist_to_source, sdz = tsz / dist_to_source;
                REAL facing = n[0] * sdx + n[1] * sdy + n[2] * sdz;
                if (facing > 0) {
                    REAL so[3] = {hx, hy, hz}, sd[3] = {sdx, sdy, sdz};
                    int occluded;
                    lc->shadow_queries++;
                    if (p->shadow_closest) {
                        int sp_, ss_;
                        REAL st_, sn_[3];
                        closest_hit(sc, so, sd, &sp_, &ss_, &st_, sn_, &lc->cs);
                        occluded = ss_ >= 0 && st_ > EPS && st_ < dist_to_source - EPS;
                    } else {
                        occluded = any_hit(sc, so, sd, dist_to_source, &lc->cs);
                    }
                    if (!occluded) {
                        REAL scat = sc->scattering[surf];
                        REAL diffuse = INV_PI * scat * facing;
                        REAL halfx = sdx - dir[0], halfy = sdy - dir[1], halfz = sdz - dir[2];
                        normalize3(&halfx, &halfy, &halfz);
                        REAL cos_half = halfx * n[0] + halfy * n[1] + halfz * n[2];
                        if (cos_half < 0) cos_half = 0;
                        REAL specular = (p->specular_exponent + RS(2)) * INV_8PI * (RS(1) - scat) *
                                        RPOW(cos_half, p->specular_exponent);
                        REAL dm = dist_to_source > p->irradiance_min_distance ? dist_to_source
                                                                              : p->irradiance_min_distance;
                        REAL distance_term = INV_4PI / (dm * dm);
                        REAL source_cosine = -(p->source_forward[0] * sdx + p->source_forward[1] * sdy +
                                               p->source_forward[2] * sdz);
                        REAL source_gain =
                            RPOW(RABS((RS(1) - p->dipole_weight) + p->dipole_weight * source_cosine), p->dipole_power);
                        REAL rel_delay = (accum_distance + t + dist_to_source) / p->c - p->direct_delay;
                        REAL q = rel_delay / p->bin_dur;
                        if (q >= 0 && q < (REAL)K) {
                            int bin = (int)RFLOOR(q);
                            REAL energy_sum = 0;
                            for (int b = 0; b < NB; ++b) {
                                REAL energy = ray_scale * source_gain * distance_term * (diffuse + specular) *
                                              sc->reflection[surf * NB + b] * accum_energy[b];
                                lc->echogram[(size_t)b * K + bin] += energy;
                                if (amb) {
                                    REAL *a = lc->ambisonic + (size_t)b * 4 * K;
                                    a[bin] += energy;
                                    a[K + bin] += energy * dir0[0];
                                    a[2 * K + bin] += energy * dir0[1];
                                    a[3 * K + bin] += energy * dir0[2];
                                }
                                energy_sum += energy;
                            }
                            /* the strongest contribution (steam_rt.py:4821-4824) */
                            if (vslot >= 0 && energy_sum > out->visual_gains[vslot] * out->visual_gains[vslot]) {
                                out->visual_orders[vslot] = bounce + 1;
                                out->visual_distances[vslot] = accum_distance + t + dist_to_source;
                                out->visual_gains[vslot] = RSQRT(energy_sum > 0 ? energy_sum : 0);
                            }
                            lc->contrib_counts[surf] += 1;
                            lc->surface_energy[surf] += energy_sum;
                            lc->contributions++;
                        }
                    }
                }
            }
            for (int b = 0; b < NB; ++b) accum_energy[b] *= sc->reflection[surf * NB + b];
            accum_distance += t;
            origin[0] = hx;
            origin[1] = hy;
            origin[2] = hz;
            REAL u;
            int32_t sample_index;
            if (p->rng_mode == 0) {
                u = p->diffuse_random[(size_t)bounce * N + ri];
                sample_index = p->diffuse_indices[(size_t)bounce * N + ri];
            } else {
                uint32_t u24;
                ao_counter_bounce(p->seed, ri, bounce, M, &u24, &sample_index);
                u = (REAL)((double)u24 * (1.0 / 16777216.0));
            }
            if (u < sc->scattering[surf]) {
                REAL nd[3];
                diffuse_direction(p->diffuse_bank + (size_t)sample_index * 3, n[0], n[1], n[2], nd);
                dir[0] = nd[0]; dir[1] = nd[1]; dir[2] = nd[2];
            } else {
                REAL dot = dir[0] * n[0] + dir[1] * n[1] + dir[2] * n[2];
                REAL dx = dir[0] - RS(2) * dot * n[0];
                REAL dy = dir[1] - RS(2) * dot * n[1];
                REAL dz = dir[2] - RS(2) * dot * n[2];
                normalize3(&dx, &dy, &dz);
                dir[0] = dx; dir[1] = dy; dir[2] = dz;
            }
            if (accum_distance > p->max_path_len) {
                alive = 0;
                break;
            }
        }
        if (alive) lc->active_count++;
    }

    /* Serial reduction in thread order (steam_rt.py:4858-4871). */
    memset(out->echogram, 0, sizeof(REAL) * (size_t)NB * K);
    if (amb) memset(out->ambisonic, 0, sizeof(REAL) * (size_t)NB * 4 * K);
    memset(out->hit_counts, 0, sizeof(int64_t) * (size_t)S);
    memset(out->contrib_counts, 0, sizeof(int64_t) * (size_t)S);
    memset(out->surface_energy, 0, sizeof(REAL) * (size_t)S);
    out->actual_bounces = 0;
    out->active_count = 0;
    out->bounce_queries = out->shadow_queries = 0;
    out->nodes_bounce = out->nodes_shadow = out->prims_bounce = out->prims_shadow = 0;
    out->contributions = out->escaped = 0;
    for (int t = 0; t < nthreads; ++t) {
        Local *lc = &loc[t];
        out->active_count += lc->active_count;
        if (lc->actual_bounces > out->actual_bounces) out->actual_bounces = lc->actual_bounces;
        out->bounce_queries += lc->bounce_queries;
        out->shadow_queries += lc->shadow_queries;
        out->nodes_bounce += lc->cb.nodes;
        out->prims_bounce += lc->cb.prims;
        out->nodes_shadow += lc->cs.nodes;
        out->prims_shadow += lc->cs.prims;
        out->contributions += lc->contributions;
        out->escaped += lc->escaped;
        for (int s = 0; s < S; ++s) {
            out->hit_counts[s] += lc->hit_counts[s];
            out->contrib_counts[s] += lc->contrib_counts[s];
            out->surface_energy[s] += lc->surface_energy[s];
        }
        for (size_t k = 0; k < (size_t)NB * K; ++k) out->echogram[k] += lc->echogram[k];
        if (amb)
            for (size_t k = 0; k < (size_t)NB * 4 * K; ++k) out->ambisonic[k] += lc->ambisonic[k];
        free(lc->echogram);
        free(lc->ambisonic);
        free(lc->surface_energy);
        free(lc->hit_counts);
        free(lc->contrib_counts);
    }
    free(loc);
    return 0;
}
