/*
 * aa_tracer.h — C ABI of the MI355X (gfx950) reflection tracer.
 *
 * This is the drop-in boundary for the reference's GPU tracer backend.  The
 * reference binds its backend as a Python module with two calls:
 *   cuda_available(device_id) -> bool             acoustic_agent/steam_rt_cuda.py:33-41
 *   trace_energy_field_cuda(**kw) -> dict         acoustic_agent/steam_rt_cuda.py:44-226
 * and builds the scene arrays and BVH it consumes in
 *   _scene_kernel_arrays(scene)                   acoustic_agent/steam_rt.py:834-924
 *   build_bvh(bounds_min, bounds_max, leaf_size)  acoustic_agent/bvh.py:8-77
 * The entry points below replace those calls one for one (see each comment).
 * All entry points return 0 on success and a negative AA_ERR_* code on
 * failure; aa_last_error() then holds a message (thread-local).
 * Pointers named d_* are device pointers (e.g. PyTorch-ROCm tensor data);
 * the library never frees caller memory.  Every call is reentrant; scene
 * handles may be shared between host threads.
 */
#ifndef AA_TRACER_H
#define AA_TRACER_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define AA_NUM_BANDS 6
#define AA_MAX_SOURCES 64

enum {
    AA_OK = 0,
    AA_ERR_INVALID = -1,  /* bad argument: maps to the reference's ValueError */
    AA_ERR_DEVICE = -2,   /* HIP runtime failure / no device: RuntimeError */
    AA_ERR_MEMORY = -3,   /* allocation failed or over budget */
    AA_ERR_LIMIT = -4     /* scene or request exceeds a compiled limit */
};

/* RNG streams. */
enum {
    AA_RNG_ARRAYS = 0,  /* caller-supplied directions [N,3], U [B,N], J [B,N]
                           (the reference's NumPy streams, steam_rt.py:3958-4044) */
    AA_RNG_COUNTER = 1  /* Philox4x32-10 keyed by the seed, generated on device */
};

/* Counters accumulated by one trace (int64 slots of aa_trace_outputs.d_counters). */
enum {
    AA_CNT_BOUNCE_QUERIES = 0,
    AA_CNT_SHADOW_QUERIES = 1,
    AA_CNT_NODES_BOUNCE = 2,
    AA_CNT_NODES_SHADOW = 3,
    AA_CNT_TRIS_BOUNCE = 4,
    AA_CNT_TRIS_SHADOW = 5,
    AA_CNT_CONTRIBUTIONS = 6,
    AA_CNT_ESCAPED = 7,
    AA_CNT_RAYS_CLAIMED = 8,  /* work cursor of the persistent lanes (>= num_rays) */
    AA_CNT_COUNT = 9
};

typedef struct aa_scene_s *aa_scene_t;

/* ---- device query (replaces cuda_available, steam_rt_cuda.py:33-41) ---- */
int aa_device_count(int32_t *count);
int aa_device_info(int32_t device, char *name, int32_t name_len, char *arch, int32_t arch_len,
                   uint64_t *hbm_bytes, int32_t *compute_units);
const char *aa_last_error(void);

/* ---- canonical BVH (replaces build_bvh, bvh.py:8-77) ----
 * Host-only.  Same outputs as the reference: preorder nodes with escape
 * links; node arrays sized at least 2n-1, primitives n; stats = {node_count,
 * leaf_count, max_depth}. */
int aa_build_bvh(const double *bounds_min, const double *bounds_max, int64_t n, int64_t leaf_size,
                 double *node_min, double *node_max, int64_t *start, int64_t *count, int64_t *escape,
                 int64_t *primitives, int64_t *stats);

/* The binned-SAH tree the device traverses (no reference counterpart: the
 * reference walks its median-split tree; the hit a query returns does not
 * depend on the tree).  Same arguments and output layout as aa_build_bvh. */
int aa_build_bvh_sah(const double *bounds_min, const double *bounds_max, int64_t n, int64_t leaf_size,
                     double *node_min, double *node_max, int64_t *start, int64_t *count, int64_t *escape,
                     int64_t *primitives, int64_t *stats);

/* ---- scene (replaces _scene_kernel_arrays + the device input cache,
 *      steam_rt.py:834-924 and steam_rt_cuda.py:229-261) ----
 * Triangles carry the reference surface id they tessellate; reflection is
 * the per-surface energy reflection 1 - alpha [S,6] and scattering the
 * per-surface mean scattering [S] (steam_rt.py:861,896).  The library builds
 * the canonical BVH, collapses it to a 4-wide BVH and uploads it. */
typedef struct {
    int32_t num_triangles;
    int32_t num_surfaces;
    int32_t num_nodes;       /* 4-wide nodes on the device */
    int32_t depth;           /* 4-wide tree depth */
    int32_t canonical_nodes; /* binary nodes (reference layout) */
    int32_t canonical_depth;
    int32_t leaf_size;
    int32_t device;
    uint64_t device_bytes;   /* device memory the scene holds */
    uint64_t content_hash;   /* FNV-1a over the inputs */
} aa_scene_info_t;

int aa_scene_create(int32_t device, const float *tri_vertices /*[T,9]*/, const int32_t *tri_surface /*[T]*/,
                    const float *tri_normals /*[T,3]*/, int32_t num_triangles, const float *reflection /*[S,6]*/,
                    const float *scattering /*[S]*/, int32_t num_surfaces, int32_t leaf_size, aa_scene_t *out);
int aa_scene_info(aa_scene_t scene, aa_scene_info_t *info);

/* The content_hash aa_scene_create stores, computed on the host without a
 * device: FNV-1a 64 over the float32/int32 inputs in argument order, then
 * leaf_size.  It keys scenes in caches and in scene files.  (No reference
 * counterpart: steam_rt_cuda.py:229-261 keys its device cache by host
 * pointer, shape and dtype.) */
int aa_scene_hash(const float *tri_vertices, const int32_t *tri_surface, const float *tri_normals,
                  int32_t num_triangles, const float *reflection, const float *scattering, int32_t num_surfaces,
                  int32_t leaf_size, uint64_t *hash_out);
int aa_scene_release(aa_scene_t scene);

/* ---- one trace (replaces trace_energy_field_cuda, steam_rt_cuda.py:44-226) ---- */
typedef struct {
    int32_t num_rays;            /* N */
    int32_t num_bounces;         /* B */
    int32_t num_bins;            /* K */
    int32_t num_diffuse_samples; /* M (bank rows) */
    int32_t num_sources;         /* Ks: 1, or up to AA_MAX_SOURCES keyframe sources
                                    sharing one listener and one ray set */
    int32_t rng_mode;            /* AA_RNG_* */
    int32_t use_bvh;             /* 0: test every triangle (the reference's linear mode) */
    int32_t render_ambisonics;   /* accumulate the FOA echogram */
    uint64_t seed;               /* AA_RNG_COUNTER key */
    float listener[3];
    float source_forward[3];     /* weighted-dipole axis (directivity.py:58-66) */
    float bin_duration;
    float speed_of_sound;
    float max_path_length;
    float listener_radius;
    float source_radius;
    float irradiance_min_distance;
    float specular_exponent;
    float dipole_weight;
    float dipole_power;
    int32_t block_size;          /* 0: tuned default; else 64/128/256/512 */
    int32_t lds_histogram;       /* -1: tuned default; 0: global atomics; 1: LDS privatised */
    int32_t workgroups_per_cu;   /* persistent grid = CUs x this (0: occupancy limit) */
    int32_t sources_per_lane;    /* multi-source traces: sources sharing one lane's trajectory
                                    (0: 8); the trace runs num_rays x ceil(Ks / this) lane tasks */
    int32_t log_rays;            /* hit-log extent (0: off) */
    int32_t log_bounces;
    int32_t visual_limit;        /* visual path candidates (steam_rt.py:4718-4731, 4765-4769,
                                    4821-4824): ray r with r % visual_stride == 0 and
                                    r / visual_stride < visual_limit records its hit points and its
                                    strongest contribution; 0: off.  aa_trace with Ks = 1 only. */
    int32_t visual_stride;
} aa_trace_params;

typedef struct {
    float *d_echogram;        /* [Ks,6,K], zeroed by the call */
    float *d_ambisonic;       /* [Ks,6,4,K] or NULL */
    int32_t *d_hit_counts;    /* [Ks,S] */
    int32_t *d_contrib_counts;/* [Ks,S] */
    float *d_surface_energy;  /* [Ks,S] */
    int32_t *d_actual_bounces;/* [Ks] */
    int32_t *d_active_count;  /* [Ks] */
    int64_t *d_counters;      /* [AA_CNT_COUNT] */
    int32_t *d_log_triangle;  /* [log_rays, log_bounces] hit triangle (-1 none) or NULL */
    float *d_log_t;           /* [log_rays, log_bounces] or NULL */
    float *d_log_point;       /* [log_rays, log_bounces, 3] hit point o + t d (before the offset), or NULL */
    float *d_visual_points;   /* [V,B,3] offset hit point per bounce (0: none); V = visual_limit */
    int32_t *d_visual_surfaces;   /* [V,B] surface id per bounce (-1: none) */
    int32_t *d_visual_rays;       /* [V] ray index of the candidate (-1: none) */
    int32_t *d_visual_orders;     /* [V] bounce + 1 of the strongest contribution (0: none) */
    float *d_visual_distances;    /* [V] its path length */
    float *d_visual_gains;        /* [V] sqrt of its band energy sum */
    int64_t *d_ray_times;     /* [N,2] when each ray started and ended (source group 0), in ticks of the
                                 device's constant 100 MHz clock (s_memrealtime), or NULL.  No
                                 reference counterpart: a diagnostic of how long rays hold their lanes
                                 (the single-trace tail); a NULL pointer costs nothing. */
} aa_trace_outputs;

/* d_sources: [Ks,4] device floats (x, y, z, direct_delay_s).
 * d_directions / d_diffuse_random / d_diffuse_indices: AA_RNG_ARRAYS only.
 * d_diffuse_bank: [M,3].  stream: hipStream_t or NULL for the null stream.
 * The call is asynchronous on `stream`. */
int aa_trace(aa_scene_t scene, const aa_trace_params *params, const float *d_sources, const float *d_directions,
             const float *d_diffuse_bank, const float *d_diffuse_random, const int32_t *d_diffuse_indices,
             const aa_trace_outputs *outputs, void *stream);

/* The plain one-scene-per-launch tracer (replaces the execution model of
 * the reference GPU backend, _trace_energy_cuda_kernel steam_rt_cuda.py:
 * 491-710, launched by trace_energy_field_cuda steam_rt_cuda.py:136-160):
 * one thread per ray, blocks of 128, no persistence or refill, a stackless
 * walk of the canonical binary BVH, shadow queries answered by a full
 * closest-hit walk.  Same arguments and results as aa_trace (the bounce
 * arithmetic is shared); kept as the baseline the persistent and batched
 * paths are measured against.  Tuning fields of `params` are ignored. */
int aa_trace_plain(aa_scene_t scene, const aa_trace_params *params, const float *d_sources,
                   const float *d_directions, const float *d_diffuse_bank, const float *d_diffuse_random,
                   const int32_t *d_diffuse_indices, const aa_trace_outputs *outputs, void *stream);

/* Synchronise `stream` (device errors come back as AA_ERR_DEVICE).  The
 * traversal stack cannot overflow: aa_scene_create rejects trees deeper than
 * the stack holds (AA_ERR_LIMIT), so kernels have no overflow to report. */
int aa_stream_sync(void *stream);

/* Poll `stream` without blocking: *done = 1 once every call queued on it
 * (traces, batch submissions) has finished, else 0.  With aa_stream_sync this
 * is the poll / wait pair of a submitted batch (the reference's callers wait
 * on futures of a thread pool, batch.py:170, api.py:511). */
int aa_stream_query(void *stream, int32_t *done);

/* ---- persistent batches (replaces the thread-pool fan-out of one trace per
 *      call: simulate_batch acoustic_agent/batch.py:139-184, run_many /
 *      run_batch / run_sources api.py:477-639, and the per-channel traces of
 *      engine.py:64-73) ----
 * One submission traces `count` independent traces (any mix of scenes, ray
 * counts, bounce limits, bins and source sets) in one persistent launch that
 * pulls packets of rays (`packet_rays`, or guided sizes) from device work
 * queues, one per XCD, each given whole scenes (a wave takes its own XCD's packets first, then
 * the other queues'), so the order of `traces` does not matter.  Each trace's
 * outputs are zeroed by the submission.  The tuning fields of each trace's
 * aa_trace_params (block_size, lds_*) are ignored; the batch tuning applies.
 * A batch object owns a pinned descriptor table: a new submission on the same
 * object waits until the previous one has finished reading it. */
typedef struct aa_batch_s *aa_batch_t;

typedef struct {
    aa_scene_t scene;
    aa_trace_params params;
    const float *d_sources;          /* [Ks,4] */
    const float *d_directions;       /* AA_RNG_ARRAYS only */
    const float *d_diffuse_bank;     /* [M,3] */
    const float *d_diffuse_random;   /* AA_RNG_ARRAYS only */
    const int32_t *d_diffuse_indices;/* AA_RNG_ARRAYS only */
    aa_trace_outputs outputs;
} aa_batch_trace;

typedef struct {
    int32_t block_size;         /* 64, 128, 256 or 512 lanes per workgroup */
    int32_t packet_rays;        /* lane tasks (rays x source groups) per work item; 0: guided
                                   sizes per trace, about 1 / (2 W) of the batch's work still
                                   left at the trace's place in its XCD queue, for the launch's
                                   W resident waves, as a power-of-two multiple of the smallest,
                                   min(1024, two work items per resident wave), at most 4096 */
    int32_t workgroups_per_cu;  /* persistent grid = CUs x this (0: from LDS use) */
    int32_t use_graph;          /* 1: capture copy + zero + kernel in a hipGraph and replay it */
} aa_batch_tuning;

int aa_batch_create(int32_t device, int32_t max_traces, aa_batch_t *out);
int aa_batch_submit(aa_batch_t batch, const aa_batch_trace *traces, int32_t count, const aa_batch_tuning *tuning,
                    void *stream);
int aa_batch_release(aa_batch_t batch);

/* ---- a batch that allocates nothing (no counterpart in the reference) ----
 * For callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid it).  Same traces,
 * tuning and results as aa_batch_submit on `device`, without a batch object:
 * the descriptor table and the work-queue heads live in a static device array
 * of AA_STATIC_BATCH_TRACES descriptors (about 100 KB, held by each GPU the
 * library runs on until the process ends), filled from a static host table.
 * Takes 1..AA_STATIC_BATCH_TRACES traces (more: AA_ERR_LIMIT); use_graph must
 * be 0.  The call allocates no memory; outputs are the caller's, as for
 * aa_batch_submit.  Every call on a GPU shares that GPU's table, so calls that
 * overlap on one GPU (submitted on different streams before the previous one
 * finished) race: keep one such submission in flight per GPU (calls on one
 * stream are ordered and safe). */
#define AA_STATIC_BATCH_TRACES 256
int aa_batch_submit_static(int32_t device, const aa_batch_trace *traces, int32_t count, const aa_batch_tuning *tuning,
                           void *stream);

/* ---- host: late-reverberation FDN (replaces _fdn_impulse_kernel,
 *      steam_rt.py:2235-2328) ----
 * Impulse response of num_delays (a power of two) Hadamard-mixed delay lines
 * of delays[] samples, each with a cascade of num_bands absorptive biquads
 * absorptive[line][band][b0 b1 b2 a1 a2], output = mean of the lines through
 * num_allpass Schroeder all-passes (gain 0.5) and the tone cascade
 * tone[band][5]; binary64, out[num_samples].  Host memory only. */
int aa_fdn_impulse(const int64_t *delays, int32_t num_delays, const double *absorptive, int32_t num_bands,
                   int64_t num_samples, const int64_t *allpass_delays, int32_t num_allpass, const double *tone,
                   double *out);

#ifdef __cplusplus
}
#endif

#endif
