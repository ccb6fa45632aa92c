/*
 * Oracle (test infrastructure only): CPU restatement of the reference
 * reflection tracer `_trace_energy_kernel` (acoustic_agent/steam_rt.py:4708-4886)
 * and its helpers (steam_rt.py:4256-4582, bvh.py:8-77).
 *
 * ao_trace.c is compiled twice: with REAL = double (suffix _f64, the FP64
 * oracle) and REAL = float (suffix _f32, the same algorithm in binary32,
 * which mirrors the reference's FP32 CUDA backend steam_rt_cuda.py:491-710).
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use
 * this library, and only as the checker.
 */
#ifndef AA_ORACLE_H
#define AA_ORACLE_H

#include <stdint.h>

#define AO_NUM_BANDS 6

#define AO_DECLARE(REALT, SUF)                                                          \
    typedef struct {                                                                    \
        int32_t mode; /* 0: analytic primitives (kinds 0/1/2), 1: triangles */          \
        int32_t num_surfaces;                                                           \
        const int64_t *kinds;                                                           \
        const REALT *wall_a, *wall_delta, *wall_z, *z_values;                           \
        const REALT *box_center, *box_axis_u, *box_axis_v, *box_half, *box_z;           \
        const REALT *normals; /* [S,3] */                                               \
        const REALT *corners; /* [P,2] */                                               \
        int32_t num_corners;                                                            \
        int32_t num_triangles;                                                          \
        const REALT *tri_v;          /* [T,9] v0 v1 v2 */                               \
        const int32_t *tri_surface;  /* [T] */                                          \
        const REALT *tri_normal;     /* [T,3] */                                        \
        int32_t use_bvh;                                                                \
        int32_t num_nodes;                                                              \
        const REALT *bvh_min, *bvh_max; /* [Nn,3] */                                    \
        const int64_t *bvh_start, *bvh_count, *bvh_escape, *bvh_prims;                  \
        const REALT *reflection; /* [S,6] energy reflection 1 - alpha */                \
        const REALT *scattering; /* [S] */                                              \
    } ao_scene_##SUF;                                                                   \
                                                                                        \
    typedef struct {                                                                    \
        int32_t num_rays, num_bounces, num_bins, num_diffuse_samples;                   \
        int32_t rng_mode;       /* 0: arrays below, 1: Philox counter stream */         \
        int32_t render_ambisonics;                                                      \
        int32_t shadow_closest; /* 1: literal closest-hit shadow query */               \
        int32_t num_threads;    /* 0: OpenMP default */                                 \
        uint64_t seed;                                                                  \
        REALT source[3], listener[3], source_forward[3];                                \
        REALT bin_dur, c, max_path_len, direct_delay;                                   \
        REALT listener_radius, source_radius, irradiance_min_distance;                  \
        REALT specular_exponent, dipole_weight, dipole_power;                           \
        const REALT *directions;         /* [N,3]  (rng_mode 0) */                      \
        const REALT *diffuse_bank;       /* [M,3] */                                    \
        const REALT *diffuse_random;     /* [B,N]  (rng_mode 0) */                      \
        const int32_t *diffuse_indices;  /* [B,N]  (rng_mode 0) */                      \
        int32_t log_rays, log_bounces;                                                  \
        int32_t ray_begin, ray_end;      /* traced ray range; ray_end 0 = all */        \
        int32_t visual_limit, visual_stride; /* candidate rays ri % stride == 0 */      \
    } ao_params_##SUF;                                                                  \
                                                                                        \
    typedef struct {                                                                    \
        REALT *echogram;        /* [6,K] */                                             \
        REALT *ambisonic;       /* [6,4,K] or NULL */                                   \
        int64_t *hit_counts;    /* [S] */                                               \
        int64_t *contrib_counts;/* [S] */                                               \
        REALT *surface_energy;  /* [S] */                                               \
        int32_t *log_prim;      /* [log_rays, log_bounces] or NULL */                   \
        int32_t *log_surface;                                                           \
        double *log_t;                                                                  \
        double *log_point;      /* [log_rays, log_bounces, 3] hit point o + t d, or NULL */ \
        REALT *visual_points;      /* [V,B,3] offset hit points, or NULL */             \
        int32_t *visual_surfaces;  /* [V,B] */                                          \
        int32_t *visual_rays;      /* [V] */                                            \
        int32_t *visual_orders;    /* [V] bounce + 1 of the strongest contribution */   \
        REALT *visual_distances;   /* [V] its path length */                            \
        REALT *visual_gains;       /* [V] sqrt of its band energy sum */                \
        int64_t actual_bounces, active_count;                                           \
        int64_t bounce_queries, shadow_queries;                                         \
        int64_t nodes_bounce, nodes_shadow, prims_bounce, prims_shadow;                 \
        int64_t contributions, escaped;                                                 \
    } ao_out_##SUF;                                                                     \
                                                                                        \
    int ao_trace_##SUF(const ao_scene_##SUF *scene, const ao_params_##SUF *params,      \
                       ao_out_##SUF *out);                                              \
    int ao_closest_hit_##SUF(const ao_scene_##SUF *scene, const REALT origin[3],        \
                             const REALT direction[3], int32_t *prim, int32_t *surface, \
                             REALT *t, REALT normal[3]);                                \
    void ao_diffuse_direction_##SUF(const REALT sample[3], const REALT normal[3],       \
                                    REALT out[3]);

AO_DECLARE(double, f64)
AO_DECLARE(float, f32)

void ao_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
void ao_counter_direction(uint64_t seed, int64_t ray, int64_t num_rays, double out[3]);
void ao_counter_bounce(uint64_t seed, int64_t ray, int32_t bounce, int32_t num_samples,
                       uint32_t *u24, int32_t *sample_index);
int ao_build_bvh(const double *lower, const double *upper, int64_t n, int64_t leaf_size,
                 double *node_min, double *node_max, int64_t *start, int64_t *count,
                 int64_t *escape, int64_t *prims, int64_t *stats);

#endif
