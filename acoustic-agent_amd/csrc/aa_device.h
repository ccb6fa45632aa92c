// Device code shared by the single-trace kernel (aa_tracer.hip) and the
// persistent batch kernel (aa_batch.hip): BVH traversal, triangle test,
// counter RNG and the bounce loop of the reference tracer
// (acoustic_agent/steam_rt.py:4724-4849), run as a per-lane state machine
// (a few traversal steps per loop round) with a per-wave queue of shadow
// connections.
//
// Floating-point contract: built with -ffp-contract=off and IEEE division /
// sqrt, so every trajectory operation rounds like the binary32 CPU
// restatement; only the BVH box tests (culling only: results do not depend on
// them) use explicit fused multiply-adds.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <numeric>
#include <type_traits>

#include "../../include/aa_tracer.h"
#include "aa_internal.h"

namespace aa {

constexpr int kBands = AA_NUM_BANDS;
constexpr int kStack = 40;          // per-lane traversal stack entries (uint16 node indices in LDS)
constexpr int kStackSlots = kStack + 1;  // LDS slots per lane: one beyond a full stack (node_step's push)
constexpr int kMaxWideNodes = 65535; // node indices must fit the 16-bit stack entries and sort keys
constexpr unsigned kNoChild = 0xFFFFFFFFu;  // sort key of a child that is not pushed
constexpr float kEps = 1e-6f;       // steam_rt_cuda.py:19
constexpr float kHitOffset = 1e-2f; // steam_rt_cuda.py:20
constexpr float kBvhEps = 1e-5f;    // steam_rt_cuda.py:21
constexpr float kBig = 1e30f;
constexpr unsigned long long kGoldenQuarter = 0x30E44323405AC1F5ull;  // round((3 - sqrt 5) / 4 * 2^64)

// Per-wave LDS queue of shadow connections waiting for a lane (see
// trace_rays_refill): kQueueCap slots of kQueueFields words.
constexpr int kQueueCap = 64;
constexpr int kQueueFields = 7 + kBands + 2 + 3;  // origin, direction, tmax | energies | bin, ks | d0

// Per-wave LDS log of histogram deposits (log_deposits): kLogCap entries of
// kLogFields words, the band-0 histogram index and the six band energies.
constexpr int kLogCap = 96;
constexpr int kLogFields = 1 + kBands;

// LDS bytes private to the lanes and waves of one workgroup: the per-lane
// stacks (rounded up to 16), the per-wave connection queues, then the
// per-wave deposit logs (16-byte aligned throughout).
__host__ __device__ constexpr size_t stack_only_bytes(int wg) {
    return ((size_t)kStackSlots * wg * sizeof(unsigned short) + 15) & ~(size_t)15;
}
__host__ __device__ constexpr size_t log_offset(int wg) {
    return stack_only_bytes(wg) + (size_t)(wg / 64) * kQueueFields * kQueueCap * sizeof(float);
}
__host__ __device__ constexpr size_t private_bytes(int wg) {
    return log_offset(wg) + (size_t)(wg / 64) * kLogFields * kLogCap * sizeof(float);
}

// A device-resident triangle scene: 4-wide BVH nodes (8 x float4 each),
// triangles in leaf order (3 x float4: v0|tri id, e1|surface id, e2|0),
// per-triangle normals (original order) and per-surface materials.
struct SceneView {
    const float4* nodes;
    const float4* tris;
    const float4* normals;
    const float* reflection;
    const float* scattering;
    int num_nodes, num_tris, num_surfaces, use_bvh;
};

// Per-trace ray set and deposit parameters.
struct RayParams {
    const float* directions;
    const float* bank;
    const float* urand;
    const int* uidx;
    const float4* sources;  // x, y, z, direct delay
    int num_rays, num_bounces, num_bins, num_samples, num_sources, rng_mode, amb;
    int src_group;  // sources handled by one lane (one task = one ray x one group)
    int groups;     // ceil(num_sources / src_group); tasks = num_rays * groups, group-major
    int ray_stride; // INTERLEAVE order: task t of a group traces ray (t * ray_stride) mod num_rays
    unsigned seed_lo, seed_hi;
    float lx, ly, lz, fx, fy, fz;
    float bin_dur, c, max_path_len, listener_radius, source_radius, dmin, spec_exp, dip_w, dip_p;
    float ray_scale;
};

struct Outputs {
    float* echogram;
    float* ambisonic;
    int* hit_counts;
    int* contrib_counts;
    float* surface_energy;
    int* actual_bounces;
    int* active_count;
    unsigned long long* counters;
    int* log_tri;
    float* log_t;
    float* log_point;  // [rays, bounces, 3] or null
    int log_rays, log_bounces;
    // visual path candidates (null / 0 when off)
    float* vis_points;  // [V,B,3]
    int* vis_surf;      // [V,B]
    int* vis_rays;      // [V]
    int* vis_orders;    // [V]
    float* vis_dist;    // [V]
    float* vis_gain;    // [V]
    int vis_limit, vis_stride, vis_bounces;
    unsigned long long* ray_times;  // [rays, 2] start / end ticks of source group 0, or null
};

// Per-lane tallies flushed once per wave.
struct LaneStats {
    unsigned q_b = 0, q_s = 0, nb = 0, ns = 0, tb = 0, ts = 0, contrib = 0, escaped = 0;
    int max_iter_full = 0;  // bounce iterations of rays that ended with every source of `grp` alive
    int active_full = 0;
    int grp = -1;           // source group of max_iter_full / active_full (multi-group traces)
};

// LDS views of one workgroup.
struct Lds {
    unsigned short* stack;  // kStackSlots x WG
    float* queue;           // per wave: kQueueFields x kQueueCap
    float* log;             // per wave: kLogFields x kLogCap (deposits waiting for their atomics)
    float* hist;         // [Ks,6,K] or null
    int* hits;           // [S] or null (single source only)
    int* contrib;
    float* senergy;
};

__device__ __forceinline__ void philox4x32_10(unsigned c0, unsigned c1, unsigned c2, unsigned c3, unsigned k0,
                                              unsigned k1, unsigned& r0, unsigned& r1) {
    // The keys are uniform: an empty asm makes them opaque here, so the round
    // keys are recomputed with scalar adds at each call instead of being
    // precomputed once and kept in (spilled) scalar registers.
    asm volatile("" : "+s"(k0), "+s"(k1));
#pragma unroll
    for (int round = 0; round < 10; ++round) {
        if (round > 0) {
            k0 += 0x9E3779B9u;
            k1 += 0xBB67AE85u;
        }
        // one 32 x 32 -> 64-bit multiply per product (v_mad_u64_u32), not
        // separate high and low multiplies
        const unsigned long long m0 = (unsigned long long)0xD2511F53u * c0;
        const unsigned long long m1 = (unsigned long long)0xCD9E8D57u * c2;
        const unsigned hi0 = (unsigned)(m0 >> 32), lo0 = (unsigned)m0;
        const unsigned hi1 = (unsigned)(m1 >> 32), lo1 = (unsigned)m1;
        const unsigned n0 = hi1 ^ c1 ^ k0;
        const unsigned n2 = hi0 ^ c3 ^ k1;
        c0 = n0;
        c1 = lo1;
        c2 = n2;
        c3 = lo0;
    }
    r0 = c0;
    r1 = c1;
}

// Address-space-typed memory operations.  The kernel argument structs hold
// generic pointers, for which the compiler emits FLAT instructions (they
// count against both the vector-memory and the LDS wait counters and check
// the aperture per lane); these casts make it emit global_* and ds_*.
typedef __attribute__((address_space(1))) float global_float;
typedef __attribute__((address_space(1))) int global_int;
typedef __attribute__((address_space(1))) float global_cfloat;
typedef __attribute__((address_space(3))) float shared_float;
typedef __attribute__((address_space(3))) int shared_int;

__device__ __forceinline__ void add_global(float* p, float v) {
    __hip_atomic_fetch_add((global_float*)p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void add_global(int* p, int v) {
    __hip_atomic_fetch_add((global_int*)p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void add_shared(float* p, float v) {
    __hip_atomic_fetch_add((shared_float*)p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
__device__ __forceinline__ void add_shared(int* p, int v) {
    __hip_atomic_fetch_add((shared_int*)p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

__device__ __forceinline__ float safe_inv(float d) {
    // Axis with |d| <= 1e-12: the slab test must only require the origin to
    // lie inside the slab (steam_rt_cuda.py:416-419); a huge signed inverse
    // gives exactly that and never produces a NaN.  The inverse only culls
    // boxes, which carry a 1e-5 m margin, so the hardware reciprocal (1 ulp)
    // serves instead of a correctly rounded division.
    return fabsf(d) > 1e-12f ? __builtin_amdgcn_rcpf(d) : copysignf(1e30f, d);
}

__device__ __forceinline__ bool tri_test(const float4 a, const float4 b, const float4 c, float ox, float oy, float oz,
                                         float dx, float dy, float dz, float& t_out) {
    // Moller-Trumbore in the operation order of the binary32 CPU restatement:
    // a*b - c*d as fma(a, b, -(c*d)) and a.b as fma(ax, bx, fma(ay, by, az*bz))
    // (measured against unfused products: batched traces +1.7-1.9 %, single
    // trace -2.7 % time).
    const float e1x = b.x, e1y = b.y, e1z = b.z;
    const float e2x = c.x, e2y = c.y, e2z = c.z;
    const float px = __builtin_fmaf(dy, e2z, -(dz * e2y));
    const float py = __builtin_fmaf(dz, e2x, -(dx * e2z));
    const float pz = __builtin_fmaf(dx, e2y, -(dy * e2x));
    const float det = __builtin_fmaf(e1x, px, __builtin_fmaf(e1y, py, e1z * pz));
    const float sx = ox - a.x, sy = oy - a.y, sz = oz - a.z;
    // the (never true) NaN test ties a's load to the first branch, so it is
    // issued with the others instead of after the determinant test
    if (!(fabsf(det) > 1e-12f) || (sx != sx)) return false;
    const float inv = 1.0f / det;
    const float u = __builtin_fmaf(sx, px, __builtin_fmaf(sy, py, sz * pz)) * inv;
    if (u < -1e-6f || u > 1.000001f) return false;
    const float qx = __builtin_fmaf(sy, e1z, -(sz * e1y));
    const float qy = __builtin_fmaf(sz, e1x, -(sx * e1z));
    const float qz = __builtin_fmaf(sx, e1y, -(sy * e1x));
    const float v = __builtin_fmaf(dx, qx, __builtin_fmaf(dy, qy, dz * qz)) * inv;
    if (v < -1e-6f || u + v > 1.000001f) return false;
    const float t = __builtin_fmaf(e2x, qx, __builtin_fmaf(e2y, qy, e2z * qz)) * inv;
    if (!(t > kEps)) return false;
    t_out = t;
    return true;
}

struct Ray {
    float ox, oy, oz, dx, dy, dz, ix, iy, iz;
};

struct Closest {
    float t;
    int surf;
    int tri;
};

__device__ __forceinline__ bool better(float t, int surf, int tri, const Closest& h) {
    // (t, surface, triangle) lexicographic: the reference's tie rule
    // steam_rt.py:4499 extended by the triangle index so that the result does
    // not depend on traversal order.
    if (t < h.t) return true;
    if (t == h.t) {
        if (h.surf < 0 || surf < h.surf) return true;
        if (surf == h.surf && tri < h.tri) return true;
    }
    return false;
}

// A 4-wide node, read from global memory (a typed pointer, so that the loads
// are global_* rather than FLAT).  The top levels are not staged in LDS: the
// hot nodes stay in L1/L2, and measured on MI355X, staging 21-341 nodes never
// paid, while the branch between the two paths cost 2 % on single traces and
// batches of keyframes.
//
// The rows are read in the ray's near/far order: for each axis the row of
// the planes the ray enters first (lo if its inverse direction is positive,
// hi if negative) goes to v[2a] and the other to v[2a + 1].  Row a of node n
// sits at byte 128 n + 32 a (lo) and 128 n + 32 a + 16 (hi), so the choice
// is the sign bit of the inverse direction moved to bit 4 of a 32-bit
// offset from the node array's base, and the far row's offset is the near
// row's with bit 4 flipped.
__device__ __forceinline__ void load_node(const SceneView& sv, int node, const Ray& r, float4* v) {
    const char* nb = reinterpret_cast<const char*>(sv.nodes);
    const unsigned base = (unsigned)node << 7;  // node < 65536: fits 23 bits
    const unsigned sx = (__float_as_uint(r.ix) >> 31) << 4;
    const unsigned sy = (__float_as_uint(r.iy) >> 31) << 4;
    const unsigned sz = (__float_as_uint(r.iz) >> 31) << 4;
    const unsigned off[7] = {base | sx, (base | sx) ^ 16u, base | sy, (base | sy) ^ 16u,
                             base | sz, (base | sz) ^ 16u, base};
#pragma unroll
    for (int k = 0; k < 7; ++k) {
        const global_cfloat* g = (const global_cfloat*)(nb + (size_t)off[k] + 32 * (k >> 1));
        v[k] = make_float4(g[0], g[1], g[2], g[3]);
    }
}

// Slab test of the 4 children against the ray's near planes (v[0], v[2],
// v[4]) and far planes (v[1], v[3], v[5]).  For a positive inverse
// direction fma(lo, i, -o i) <= fma(hi, i, -o i) (lo <= hi, and a correctly
// rounded fma is monotonic in its first operand), for a negative one the
// reverse, so the near-plane distance is exactly min(t0, t1) and the far one
// max(t0, t1): the same values as testing both planes, without the min/max
// pairs.  The entry distance is clamped below at the float just above EPS,
// which folds the test tmax > EPS into tmin <= tmax (a child with
// tmax <= EPS fails both, one with tmax > EPS >= its raw entry distance
// passes both).  The clamp moves only entry distances below EPS: children
// the ray enters that close to its origin now sort in index order, which
// changes no result (the closest hit does not depend on the visit order),
// only which of them a shadow query visits first.  Measured on MI355X
// against testing both planes: batched mic arrays +5.2 %, batched keyframes
// +5.8 %, single trace -7.5 % time.
__device__ __forceinline__ void slab4(const float4* v, const Ray& r, float tlim, float* tn, bool* hit) {
    const float oix = r.ox * r.ix, oiy = r.oy * r.iy, oiz = r.oz * r.iz;
    const float eps_up = __uint_as_float(__float_as_uint(kEps) + 1u);  // next float above EPS
    const float nxs[4] = {v[0].x, v[0].y, v[0].z, v[0].w};
    const float fxs[4] = {v[1].x, v[1].y, v[1].z, v[1].w};
    const float nys[4] = {v[2].x, v[2].y, v[2].z, v[2].w};
    const float fys[4] = {v[3].x, v[3].y, v[3].z, v[3].w};
    const float nzs[4] = {v[4].x, v[4].y, v[4].z, v[4].w};
    const float fzs[4] = {v[5].x, v[5].y, v[5].z, v[5].w};
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        const float tnx = __builtin_fmaf(nxs[c], r.ix, -oix), tfx = __builtin_fmaf(fxs[c], r.ix, -oix);
        const float tny = __builtin_fmaf(nys[c], r.iy, -oiy), tfy = __builtin_fmaf(fys[c], r.iy, -oiy);
        const float tnz = __builtin_fmaf(nzs[c], r.iz, -oiz), tfz = __builtin_fmaf(fzs[c], r.iz, -oiz);
        const float tmin = fmaxf(fmaxf(tnx, tny), fmaxf(tnz, eps_up));
        const float tmax = fminf(fminf(tfx, tfy), fminf(tfz, tlim));
        tn[c] = tmin;
        hit[c] = tmin <= tmax;
    }
}

__device__ __forceinline__ bool sphere_before(float ox, float oy, float oz, float dx, float dy, float dz, float cx,
                                              float cy, float cz, float radius, float max_distance) {
    const float px = ox - cx, py = oy - cy, pz = oz - cz;
    const float b = px * dx + py * dy + pz * dz;
    const float cc = px * px + py * py + pz * pz - radius * radius;
    const float disc = b * b - cc;
    if (disc < 0.0f) return false;
    const float root = -b - sqrtf(disc > 0.0f ? disc : 0.0f);
    return root >= 0.0f && root < max_distance;
}

__device__ __forceinline__ void normalize3(float& x, float& y, float& z) {
    const float n = sqrtf(x * x + y * y + z * z);
    if (n <= 1e-12f) {
        x = y = z = 0.0f;
        return;
    }
    x = x / n;
    y = y / n;
    z = z / n;
}

__device__ __forceinline__ void initial_direction(const RayParams& p, int ray, float& dx, float& dy, float& dz) {
    if (p.rng_mode == AA_RNG_ARRAYS) {
        dx = p.directions[(size_t)ray * 3];
        dy = p.directions[(size_t)ray * 3 + 1];
        dz = p.directions[(size_t)ray * 3 + 2];
        return;
    }
    // Fibonacci lattice with a U(-0.02, 0.02) jitter (steam_rt.py:3958-3968), in binary64.
    // The golden-angle step (i + 1/2) pi (3 - sqrt 5) is reduced modulo 2 pi
    // exactly, as 2 pi frac((2 ray + 1) K / 2^64) with K = (3 - sqrt 5) / 4 in
    // 64-bit fixed point (the product wraps modulo 2^64), so the cosine and
    // sine see arguments in [-0.02, 2 pi + 0.02] and avoid the long
    // large-argument reduction (measured: about 300 binary64 instructions of
    // each completion pass that starts a ray).
    unsigned w0, w1;
    philox4x32_10(0xFFFFFFFFu, (unsigned)ray, 0u, 0u, p.seed_lo, p.seed_hi, w0, w1);
    const double u = (double)(w0 >> 8) * (1.0 / 16777216.0);
    const double jitter = -0.02 + 0.04 * u;
    const double n = (double)(p.num_rays > 0 ? p.num_rays : 1);
    const double i = (double)ray + 0.5;
    const double z = 1.0 - 2.0 * i / n;
    double rr = 1.0 - z * z;
    rr = rr < 0.0 ? 0.0 : (rr > 1.0 ? 1.0 : rr);
    const double rad = sqrt(rr);
    const unsigned long long f = (2ull * (unsigned long long)(unsigned)ray + 1ull) * kGoldenQuarter;
    const double theta = 6.283185307179586 * ((double)(f >> 11) * 0x1p-53) + jitter;
    dx = (float)(cos(theta) * rad);
    dy = (float)(sin(theta) * rad);
    dz = (float)z;
}

// Resumable state of one listener ray between bounces, so that a lane whose
// ray has ended can start the next ray while the rest of its wave continues.
struct RayState {
    float ox, oy, oz, dx, dy, dz;
    float d0x, d0y, d0z;  // initial direction (FOA weights)
    float A[kBands];      // accumulated energy reflection per band
    float accum;          // path length so far
    unsigned long long mask;        // sources for which the ray is still alive
    unsigned long long group_mask;  // the sources of this task's group
    int ray;
    int group;
    int bounce;
};

__device__ __forceinline__ unsigned long long group_source_mask(const RayParams& p, int group) {
    const int first = group * p.src_group;
    const int last = min(p.num_sources, first + p.src_group);
    const unsigned long long upto = last >= 64 ? ~0ull : ((1ull << last) - 1ull);
    const unsigned long long below = (1ull << first) - 1ull;
    return upto & ~below;
}

// Per-source end-of-ray tallies of one lane, written out directly.
__device__ __forceinline__ void flush_group_tallies(const RayParams& p, const Outputs& out, LaneStats& st) {
    if (st.grp >= 0 && (st.max_iter_full || st.active_full)) {
        unsigned long long m = group_source_mask(p, st.grp);
        while (m) {
            const int k = __builtin_ctzll(m);
            m &= m - 1ull;
            if (st.max_iter_full) atomicMax(&out.actual_bounces[k], st.max_iter_full);
            if (st.active_full) atomicAdd(&out.active_count[k], st.active_full);
        }
    }
    st.max_iter_full = 0;
    st.active_full = 0;
}

__device__ __forceinline__ unsigned long long full_source_mask(int Ks) {
    return Ks >= 64 ? ~0ull : ((1ull << Ks) - 1ull);
}

// Start / end stamps of a ray (source group 0) on the constant-rate
// realtime clock, when the caller asked for them.
__device__ __forceinline__ void stamp_ray(const Outputs& out, int ray, int group, int which) {
    if (out.ray_times != nullptr && group == 0) out.ray_times[2 * (size_t)ray + which] = wall_clock64();
}

// End-of-ray bookkeeping (steam_rt.py:4745-4746, 4848-4849): actual bounces
// and the active count, per source.
__device__ __forceinline__ void finish_ray(const RayParams& p, const Outputs& out, const RayState& s, bool alive,
                                           int iter, LaneStats& st) {
    stamp_ray(out, s.ray, s.group, 1);
    if (s.mask == s.group_mask) {
        if (p.groups > 1 && st.grp != s.group) {
            flush_group_tallies(p, out, st);
            st.grp = s.group;
        }
        st.max_iter_full = max(st.max_iter_full, iter);
        if (alive) ++st.active_full;
    } else {
        unsigned long long m = s.mask;
        while (m) {
            const int k = __builtin_ctzll(m);
            m &= m - 1ull;
            atomicMax(&out.actual_bounces[k], iter);
            if (alive) atomicAdd(&out.active_count[k], 1);
        }
    }
}

// Visual candidate slot of `ray`, or -1 (steam_rt.py:4724-4729): rays with
// ray % stride == 0 while ray / stride is below the limit.
__device__ __forceinline__ int visual_slot(const Outputs& out, int ray) {
    if (out.vis_limit <= 0 || ray % out.vis_stride != 0) return -1;
    const int v = ray / out.vis_stride;
    return v < out.vis_limit ? v : -1;
}

// Start lane task `task` (ray x source group); returns false if the ray has
// no bounce to trace (num_bounces == 0).  VISUAL: the caller may record
// visual candidates (the batch kernel never does).  INTERLEAVE: consecutive
// tasks take rays a golden-ratio stride apart.  Consecutive rays of the
// lattice share an elevation, and the steep ones all run to the bounce
// limit (short segments between floor and ceiling), so in ray order they
// fill whole waves whose lanes then walk in lockstep to the end; spread
// over the waves, the longest rays finish sooner (measured on a 200k-ray
// single trace: 7.60 -> 7.39 ms).  Every ray is traced with the same
// arithmetic in either order.
template <bool VISUAL = true, bool INTERLEAVE = false>
__device__ __forceinline__ bool start_ray(const RayParams& p, const Outputs& out, int task, RayState& s,
                                          LaneStats& st) {
    const int group = task / p.num_rays;  // group-major: a wave's claims mostly share a group
    const int local = task - group * p.num_rays;
    const int ray = INTERLEAVE ? (int)(((unsigned long long)local * (unsigned)p.ray_stride) % (unsigned)p.num_rays)
                               : local;
    s.ray = ray;
    s.group = group;
    s.group_mask = group_source_mask(p, group);
    s.bounce = 0;
    s.ox = p.lx;
    s.oy = p.ly;
    s.oz = p.lz;
    initial_direction(p, ray, s.dx, s.dy, s.dz);
    s.d0x = s.dx;
    s.d0y = s.dy;
    s.d0z = s.dz;
#pragma unroll
    for (int b = 0; b < kBands; ++b) s.A[b] = 1.0f;
    s.accum = 0.0f;
    s.mask = s.group_mask;
    stamp_ray(out, ray, group, 0);
    if (VISUAL && out.vis_limit > 0) {
        const int v = visual_slot(out, ray);
        if (v >= 0) out.vis_rays[v] = ray;
    }
    if (p.num_bounces <= 0) {
        finish_ray(p, out, s, true, 0, st);
        return false;
    }
    return true;
}

// ---------------------------------------------------------------------------
// Per-lane state machine: one loop round = a few traversal steps
// ---------------------------------------------------------------------------
//
// A lane is IDLE (needs a ray), in a BOUNCE query (closest hit of the ray's
// current segment) or in a SHADOW query (any hit towards one source).  Every
// round advances each busy lane's query by up to kStepsPerRound steps (a
// node visit, or the wave's leaf pass for the lanes waiting on hit leaves),
// so the lanes of a wave run the same traversal code whatever bounce or
// query they are in.  When a query ends, the lane does the
// reference's work for that point of the bounce loop (steam_rt.py:4743-4848)
// and sets up its next query:
//   bounce done -> termination tests, hit counts, then one connection per
//                  live source that needs a shadow query (sources in
//                  ascending order), each parked in the wave's queue for an
//                  idle lane (the ray's own lane runs it if the queue is
//                  full), then the continuation;
//   shadow done -> deposit if unoccluded; a lane that ran its own ray's
//                  connection goes on with that ray, others become idle;
//   continuation -> energy / path update, specular or diffuse direction, then
//                  the next bounce query, or the end of the ray.
// The arithmetic and its order are those of the one-ray-at-a-time loop; only
// the order of the float atomic sums changes.

enum : int { kIdle = 0, kBounce = 1, kShadow = 2 };

struct Query {
    Ray r;
    float tmax;   // shadow: distance to the source
    int node;     // next node to visit; -1 when the walk is over
    int pend;     // leaves waiting for the wave's leaf pass: visited node | hit leaf slots << 16, or -1
    int sp;       // stack depth
    Closest h;    // bounce: best hit so far
    bool occluded;
};

// Ray-tracing context of the lane between the queries of one bounce.
struct HitContext {
    float hx, hy, hz;        // offset hit point (shadow origin, next origin)
    float nx, ny, nz;        // normal oriented against the ray
    float t;                 // segment length
    int surf;
    unsigned long long nee;  // sources still to connect at this hit
};

__device__ __forceinline__ void begin_query(Query& q, float ox, float oy, float oz, float dx, float dy, float dz,
                                            float tmax) {
    q.r.ox = ox;
    q.r.oy = oy;
    q.r.oz = oz;
    q.r.dx = dx;
    q.r.dy = dy;
    q.r.dz = dz;
    q.r.ix = safe_inv(dx);
    q.r.iy = safe_inv(dy);
    q.r.iz = safe_inv(dz);
    q.tmax = tmax;
    q.node = 0;
    q.pend = -1;
    q.sp = 0;
    q.h.t = kBig;
    q.h.surf = -1;
    q.h.tri = -1;
    q.occluded = false;
}

// One triangle against the query: bounce queries keep the closest hit,
// shadow queries flag a hit before the source (tri_test already requires
// t > EPS, the lower end of (EPS, tmax - EPS)).
__device__ __forceinline__ void test_one(const float4 ta, const float4 tb, const float4 tc, Query& q, bool shadow) {
    float t;
    if (tri_test(ta, tb, tc, q.r.ox, q.r.oy, q.r.oz, q.r.dx, q.r.dy, q.r.dz, t)) {
        if (shadow) {
            if (t < q.tmax - kEps) q.occluded = true;
        } else {
            const int tri = __float_as_int(ta.w), surf = __float_as_int(tb.w);
            if (better(t, surf, tri, q.h)) {
                q.h.t = t;
                q.h.surf = surf;
                q.h.tri = tri;
            }
        }
    }
}

// Triangles [start, start + count) against the query.  Triangles are loaded
// two at a time (the second index clamped into the range, so both loads are
// issued before the first test waits): a leaf costs one dependent round trip
// to L2 per pair instead of one per triangle.  One copy serves both query
// kinds, so lanes of either kind test together.
__device__ __forceinline__ void test_leaf(const SceneView& sv, int start, int count, Query& q, bool shadow,
                                          unsigned& tri_cnt) {
    for (int k = 0; k < count; k += 2) {
        const bool two = k + 1 < count;
        const float4* t0 = sv.tris + (size_t)(start + k) * 3;
        const float4* t1 = sv.tris + (size_t)(start + (two ? k + 1 : k)) * 3;
        const float4 a0 = t0[0], b0 = t0[1], c0 = t0[2];
        const float4 a1 = t1[0], b1 = t1[1], c1 = t1[2];
        ++tri_cnt;
        test_one(a0, b0, c0, q, shadow);
        if (q.occluded) break;
        if (two) {
            ++tri_cnt;
            test_one(a1, b1, c1, q, shadow);
            if (q.occluded) break;
        }
    }
}

// One node visit of `q`: the slab test of its 4 children, then the hit
// internal children are pushed nearest first (culled against the best hit
// so far: bounce queries beyond it, shadow queries beyond the source) and
// the nearest is visited next (or the stack is popped).  Hit leaf children
// are not tested here: they are recorded in q.pend and the lane waits for
// the wave's next leaf pass (leaf_pass), which tests them before the lane
// visits another node, so a query makes the same decisions as when it tests
// its leaves at once, except that children are culled against the best hit
// from before this node's leaves (a few node visits in 10^3 more).
template <int WG>
__device__ __forceinline__ void node_step(const SceneView& sv, unsigned short* stk, Query& q, bool shadow,
                                          unsigned& node_cnt) {
    ++node_cnt;
    float4 v[7];
    load_node(sv, q.node, q.r, v);
    const float tlim = (shadow ? q.tmax : q.h.t) + kBvhEps;
    float tn[4];
    bool hit[4];
    slab4(v, q.r, tlim, tn, hit);
    const int ch[4] = {__float_as_int(v[6].x), __float_as_int(v[6].y), __float_as_int(v[6].z), __float_as_int(v[6].w)};
    unsigned leaves = 0;
#pragma unroll
    for (int c = 0; c < 4; ++c) leaves |= (hit[c] && ch[c] <= -2) ? (1u << c) : 0u;
    if (leaves) q.pend = q.node | (int)(leaves << 16);
    // Sort key of an internal child: the top 16 bits of its entry distance
    // (a non-negative float, so its bits order as unsigned integers; 7
    // mantissa bits) above its 16-bit node index, so that one unsigned
    // min/max pair swaps both and the stack entry is the key's low half.
    // Children closer than 1 part in 128 go in index order; the closest
    // hit does not depend on the order (the tie rule above).  Measured
    // against a float-key network that swaps the indices alongside:
    // batched traces +1.0-2.0 %, single trace -1.7 % time.
    unsigned key[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        const bool ok = hit[c] && ch[c] >= 0;
        key[c] = ok ? ((__float_as_uint(tn[c]) & 0x7FFF0000u) | (unsigned)ch[c]) : kNoChild;
    }
#define AA_CSWAP(i, j)                           \
    {                                            \
        const unsigned lo = min(key[i], key[j]); \
        key[j] = max(key[i], key[j]);            \
        key[i] = lo;                             \
    }
    // nearest first for both kinds: it tightens the culling of bounce
    // queries and makes shadow queries meet their occluder sooner (13 % fewer
    // node visits per shadow query than unordered on a furnished plan)
    AA_CSWAP(0, 1)
    AA_CSWAP(2, 3)
    AA_CSWAP(0, 2)
    AA_CSWAP(1, 3)
    AA_CSWAP(1, 2)
#undef AA_CSWAP
    // Push key[3], key[2], key[1] (far to near, so the nearest is on top)
    // without branches: the keys are sorted, so the valid ones are a prefix,
    // and writing key[3] at sp, key[2] at sp + v3 and key[1] at sp + v3 + v2
    // (v = valid) puts every valid key above the old top in order, while an
    // invalid one is overwritten by the next write or left just above the
    // new top.  The stack never holds more than kStack entries
    // (aa_scene_create caps 3 (depth - 1) at kStack), and the one slot beyond
    // it takes the last case's write at a full stack.
    const int v1 = key[1] != kNoChild, v2 = key[2] != kNoChild, v3 = key[3] != kNoChild;
    stk[q.sp * WG] = (unsigned short)key[3];
    stk[(q.sp + v3) * WG] = (unsigned short)key[2];
    stk[(q.sp + v3 + v2) * WG] = (unsigned short)key[1];
    q.sp += v1 + v2 + v3;
    if (key[0] != kNoChild) {
        q.node = (int)(key[0] & 0xFFFFu);
    } else if (q.sp > 0) {
        --q.sp;
        q.node = (int)stk[q.sp * WG];
    } else {
        q.node = -1;
    }
}

// The leaf pass of a lane with pending leaves: the triangles of the first
// pending leaf child (lowest slot) of node q.pend & 0xFFFF; the other
// pending slots stay in q.pend for the next pass, so every pass is one leaf
// per lane (lanes holding several leaves would otherwise run the pass again
// with a few lanes).  A shadow query ends at its first occluder.  The leaf's
// code is read again from the node (its line was just loaded by the visit).
__device__ __forceinline__ void leaf_pass(const SceneView& sv, Query& q, bool shadow, unsigned& tri_cnt) {
    const int node = q.pend & 0xFFFF;
    const unsigned leaves = (unsigned)q.pend >> 16;
    const int c = __builtin_ctz(leaves);
    const unsigned rest = leaves & (leaves - 1u);
    const global_cfloat* g =
        (const global_cfloat*)(reinterpret_cast<const char*>(sv.nodes) + ((size_t)node << 7) + 96 + 4 * c);
    const int code = -(__float_as_int(g[0]) + 2);
    test_leaf(sv, code >> 3, (code & 7) + 1, q, shadow, tri_cnt);
    if (q.occluded) {
        q.pend = -1;
        q.node = -1;
    } else {
        q.pend = rest ? (node | (int)(rest << 16)) : -1;
    }
}

// A shadow connection from the current hit to one source, with its deposit
// (steam_rt.py:4793-4826) evaluated before the shadow query: the query itself
// is the lane's Query (origin = hit point, direction = towards the source,
// tmax = distance); if the query finds no occluder the lane adds these
// energies.  Evaluating the deposit first lets any lane of the wave resolve
// the connection while the ray that made it goes on to its next bounce.
struct Connection {
    float e[kBands];      // energy per band
    int bin;              // histogram bin
    int ks;               // source | surface << 6
    float d0x, d0y, d0z;  // initial direction of the ray (FOA weights)
};

// Connection from the hit in `c` to source k (steam_rt.py:4770-4826, same
// arithmetic and order).  Returns false if the reference casts no shadow
// query (D <= d_min or the source is behind the surface); bin is -1 when the
// arrival falls outside the histogram (the query is counted, not traced: it
// cannot change the result).
__device__ __forceinline__ bool connect(const SceneView& sv, const RayParams& p, const RayState& s,
                                        const HitContext& c, int k, Query& q, Connection& cn) {
    const float4 S = p.sources[k];
    const float tsx = S.x - c.hx, tsy = S.y - c.hy, tsz = S.z - c.hz;
    const float D = sqrtf(tsx * tsx + tsy * tsy + tsz * tsz);
    if (!(D > p.dmin)) return false;
    const float sdx = tsx / D, sdy = tsy / D, sdz = tsz / D;
    const float facing = c.nx * sdx + c.ny * sdy + c.nz * sdz;
    if (!(facing > 0.0f)) return false;
    const int K = p.num_bins;
    const float rel_delay = (s.accum + c.t + D) / p.c - S.w;
    const float qb = rel_delay / p.bin_dur;
    cn.bin = (qb >= 0.0f && qb < (float)K) ? (int)floorf(qb) : -1;
    if (cn.bin >= 0) {
        const float scat = sv.scattering[c.surf];
        const float diffuse = 0.31830988618379067f * scat * facing;
        // half vector: energy only (not the trajectory), so a hardware
        // reciprocal square root replaces the correctly rounded normalisation
        float hvx = sdx - s.dx, hvy = sdy - s.dy, hvz = sdz - s.dz;
        const float hn2 = hvx * hvx + hvy * hvy + hvz * hvz;
        const float hinv = hn2 > 1e-24f ? __builtin_amdgcn_rsqf(hn2) : 0.0f;
        float cos_half = (hvx * c.nx + hvy * c.ny + hvz * c.nz) * hinv;
        cos_half = cos_half < 0.0f ? 0.0f : cos_half;
        // cos_half^n as 2^(n log2 cos_half) with the hardware base-2 logarithm
        // and exponential (energy only): a relative error of about
        // n ln2 |log2 cos_half| 2^-22, below 4e-6 wherever the lobe is above
        // 1e-6 at n = 100; cos_half = 0 gives 2^-inf = 0, and n = 0 gives 1 as
        // powf does (measured against powf: batched keyframes +3.2 %,
        // single trace -2.6 % time)
        const float lobe = p.spec_exp == 0.0f
                               ? 1.0f
                               : __builtin_amdgcn_exp2f(p.spec_exp * __builtin_amdgcn_logf(cos_half));
        const float specular = (p.spec_exp + 2.0f) * 0.039788735772973836f * (1.0f - scat) * lobe;
        const float dm = D > p.dmin ? D : p.dmin;
        const float distance_term = 0.07957747154594767f * __builtin_amdgcn_rcpf(dm * dm);
        // omni source (w = 0): (1 - 0) + 0 * cos = 1 exactly, and 1^p = 1
        float source_gain = 1.0f;
        if (p.dip_w != 0.0f) {
            const float source_cosine = -(p.fx * sdx + p.fy * sdy + p.fz * sdz);
            source_gain = powf(fabsf((1.0f - p.dip_w) + p.dip_w * source_cosine), p.dip_p);
        }
        const float base = p.ray_scale * source_gain * distance_term * (diffuse + specular);
#pragma unroll
        for (int b = 0; b < kBands; ++b) cn.e[b] = base * sv.reflection[c.surf * kBands + b] * s.A[b];
        cn.ks = k | (c.surf << 6);
        cn.d0x = s.d0x;
        cn.d0y = s.d0y;
        cn.d0z = s.d0z;
        begin_query(q, c.hx, c.hy, c.hz, sdx, sdy, sdz, D);
    }
    return true;
}

// Next connection of the hit that needs a shadow query (sources in ascending
// order); false once every source of the hit is resolved.
__device__ __forceinline__ bool next_connection(const SceneView& sv, const RayParams& p, const RayState& s,
                                                HitContext& c, Query& q, Connection& cn, LaneStats& st) {
    while (c.nee) {
        const int k = __builtin_ctzll(c.nee);
        c.nee &= c.nee - 1ull;
        if (!connect(sv, p, s, c, k, q, cn)) continue;
        ++st.q_s;
        if (cn.bin >= 0) return true;
    }
    return false;
}

// The one connection of a hit of a single-source trace (next_connection
// without its loop over sources).
__device__ __forceinline__ bool only_connection(const SceneView& sv, const RayParams& p, const RayState& s,
                                                HitContext& c, Query& q, Connection& cn, LaneStats& st) {
    if (c.nee == 0ull) return false;
    c.nee = 0ull;
    if (!connect(sv, p, s, c, 0, q, cn)) return false;
    ++st.q_s;
    return cn.bin >= 0;
}

// Deposit of an unoccluded connection (steam_rt.py:4810-4826).  With
// LOGGED the histogram part has already gone to the wave's deposit log.
template <bool LDS_HIST, bool LOGGED = false>
__device__ __forceinline__ float deposit(const RayParams& p, const Outputs& out, const Lds& lds, int num_surfaces,
                                         const Connection& cn, LaneStats& st) {
    const int K = p.num_bins;
    const int k = cn.ks & 63, surf = cn.ks >> 6;
    float esum = 0.0f;
#pragma unroll
    for (int b = 0; b < kBands; ++b) {
        const float e = cn.e[b];
        const size_t hi = ((size_t)k * kBands + b) * K + cn.bin;
        if (LDS_HIST)
            add_shared(&lds.hist[hi], e);
        else if (!LOGGED)
            add_global(&out.echogram[hi], e);
        if (p.amb) {
            float* am = out.ambisonic + (((size_t)k * kBands + b) * 4) * K + cn.bin;
            add_global(am, e);
            add_global(am + K, e * cn.d0x);
            add_global(am + 2 * K, e * cn.d0y);
            add_global(am + 3 * K, e * cn.d0z);
        }
        esum += e;
    }
    ++st.contrib;
    if (lds.hits) {
        add_shared(&lds.contrib[surf], 1);
        add_shared(&lds.senergy[surf], esum);
    } else {
        add_global(&out.contrib_counts[(size_t)k * num_surfaces + surf], 1);
        add_global(&out.surface_energy[(size_t)k * num_surfaces + surf], esum);
    }
    return esum;
}

// A visual candidate's strongest contribution so far (steam_rt.py:4821-4824:
// replaced when the band energy sum exceeds the square of the stored root).
// Candidates resolve their own connections, so only the ray's own lane
// updates its slot, in bounce order.
__device__ __forceinline__ void visual_update(const Outputs& out, const RayState& s, const HitContext& c,
                                              const Query& q, float esum) {
    const int v = visual_slot(out, s.ray);
    if (v < 0) return;
    const float g = out.vis_gain[v];
    if (esum > g * g) {
        out.vis_orders[v] = s.bounce + 1;
        out.vis_dist[v] = s.accum + c.t + q.tmax;  // q.tmax is the distance to the source
        out.vis_gain[v] = sqrtf(esum > 0.0f ? esum : 0.0f);
    }
}

__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

// The wave's deposit log.  A device-scope float atomic completes beyond the
// XCD's L2, and on this ISA one counter (vmcnt) tracks loads and atomics
// alike, so a lane that issues its six histogram atomics in a completion pass
// makes its next node load wait for them.  Instead the deposits of a pass are
// written to the log (LDS) and the log is flushed in a burst, all lanes
// issuing atomics together, when the next pass's deposits would not fit and
// when the wave's work ends: the wait is paid once per flush, not once per
// pass.  Measured on MI355X: the histogram atomics cost the batched mic-array
// workload 12 % when issued per pass.  Entry field f of slot i is
// log[f * kLogCap + i]; the histogram is [Ks, 6, K], band b at index + b K.

__device__ __forceinline__ void flush_log(const RayParams& p, const Outputs& out, const float* lg, int count,
                                          int lane) {
    wave_sync();  // the entries written by other lanes are visible
    const int K = p.num_bins;
    for (int i = lane; i < count; i += 64) {
        const int h0 = __float_as_int(lg[i]);
#pragma unroll
        for (int b = 0; b < kBands; ++b) add_global(&out.echogram[(size_t)h0 + (size_t)b * K], lg[(1 + b) * kLogCap + i]);
    }
    wave_sync();  // entries are read before they are overwritten
}

// Append the deposits of this pass (lanes with `dep`) to the log, flushing it
// first if they would not fit.  Called by every lane of the wave; `count` is
// wave-uniform.
__device__ __forceinline__ void log_deposits(bool dep, const RayParams& p, const Outputs& out, const Connection& cn,
                                             float* lg, int& count, int lane, unsigned long long below) {
    const unsigned long long m = __ballot(dep);
    const int n = __popcll(m);
    if (n == 0) return;
    if (count + n > kLogCap) {
        flush_log(p, out, lg, count, lane);
        count = 0;
    }
    if (dep) {
        const int slot = count + __popcll(m & below);
        lg[slot] = __int_as_float((cn.ks & 63) * kBands * p.num_bins + cn.bin);
#pragma unroll
        for (int b = 0; b < kBands; ++b) lg[(1 + b) * kLogCap + slot] = cn.e[b];
    }
    count += n;
}

// Field f of queue slot i is queue[f * kQueueCap + i] (consecutive lanes
// touch consecutive words).

__device__ __forceinline__ void park(float* qu, int slot, const Query& q, const Connection& cn) {
    float* f = qu + slot;
    f[0 * kQueueCap] = q.r.ox;
    f[1 * kQueueCap] = q.r.oy;
    f[2 * kQueueCap] = q.r.oz;
    f[3 * kQueueCap] = q.r.dx;
    f[4 * kQueueCap] = q.r.dy;
    f[5 * kQueueCap] = q.r.dz;
    f[6 * kQueueCap] = q.tmax;
#pragma unroll
    for (int b = 0; b < kBands; ++b) f[(7 + b) * kQueueCap] = cn.e[b];
    f[(7 + kBands) * kQueueCap] = __int_as_float(cn.bin);
    f[(8 + kBands) * kQueueCap] = __int_as_float(cn.ks);
    f[(9 + kBands) * kQueueCap] = cn.d0x;
    f[(10 + kBands) * kQueueCap] = cn.d0y;
    f[(11 + kBands) * kQueueCap] = cn.d0z;
}

__device__ __forceinline__ void unpark(const float* qu, int slot, Query& q, Connection& cn) {
    const float* f = qu + slot;
    begin_query(q, f[0], f[1 * kQueueCap], f[2 * kQueueCap], f[3 * kQueueCap], f[4 * kQueueCap], f[5 * kQueueCap],
                f[6 * kQueueCap]);
#pragma unroll
    for (int b = 0; b < kBands; ++b) cn.e[b] = f[(7 + b) * kQueueCap];
    cn.bin = __float_as_int(f[(7 + kBands) * kQueueCap]);
    cn.ks = __float_as_int(f[(8 + kBands) * kQueueCap]);
    cn.d0x = f[(9 + kBands) * kQueueCap];
    cn.d0y = f[(10 + kBands) * kQueueCap];
    cn.d0z = f[(11 + kBands) * kQueueCap];
}

// Continuation after the hit's connections (steam_rt.py:4827-4848); returns
// the lane's next phase (a new bounce query, or idle when the ray ended).
__device__ __forceinline__ int continue_ray(const SceneView& sv, const RayParams& p, const Outputs& out,
                                            RayState& s, const HitContext& c, Query& q, LaneStats& st) {
    const int bounce = s.bounce;
#pragma unroll
    for (int b = 0; b < kBands; ++b) s.A[b] *= sv.reflection[c.surf * kBands + b];
    s.accum += c.t;
    s.ox = c.hx;
    s.oy = c.hy;
    s.oz = c.hz;
    const float scat = sv.scattering[c.surf];
    float u;
    int j;
    if (p.rng_mode == AA_RNG_ARRAYS) {
        u = p.urand[(size_t)bounce * p.num_rays + s.ray];
        j = p.uidx[(size_t)bounce * p.num_rays + s.ray];
    } else {
        unsigned w0, w1;
        philox4x32_10((unsigned)bounce, (unsigned)s.ray, 0u, 0u, p.seed_lo, p.seed_hi, w0, w1);
        u = (float)(w0 >> 8) * (1.0f / 16777216.0f);
        j = (int)(((unsigned long long)w1 * (unsigned long long)(unsigned)p.num_samples) >> 32);
    }
    const float nx = c.nx, ny = c.ny, nz = c.nz;
    if (u < scat) {
        const float sx = p.bank[j * 3], sy = p.bank[j * 3 + 1], sz = p.bank[j * 3 + 2];
        float hx2 = 1.0f, hy2 = 0.0f, hz2 = 0.0f;
        if (fabsf(nx) >= 0.9f) {
            hx2 = 0.0f;
            hy2 = 1.0f;
        }
        float tx = hy2 * nz - hz2 * ny;
        float ty = hz2 * nx - hx2 * nz;
        float tz = hx2 * ny - hy2 * nx;
        normalize3(tx, ty, tz);
        const float bx = ny * tz - nz * ty;
        const float by = nz * tx - nx * tz;
        const float bz = nx * ty - ny * tx;
        s.dx = sx * tx + sy * bx + sz * nx;
        s.dy = sx * ty + sy * by + sz * ny;
        s.dz = sx * tz + sy * bz + sz * nz;
    } else {
        const float dot = s.dx * nx + s.dy * ny + s.dz * nz;
        s.dx = s.dx - 2.0f * dot * nx;
        s.dy = s.dy - 2.0f * dot * ny;
        s.dz = s.dz - 2.0f * dot * nz;
    }
    normalize3(s.dx, s.dy, s.dz);  // both branches end in the same normalisation (one copy in the pass)
    if (s.accum > p.max_path_len) {
        finish_ray(p, out, s, false, bounce + 1, st);
        return kIdle;
    }
    s.bounce = bounce + 1;
    if (s.bounce >= p.num_bounces) {
        finish_ray(p, out, s, true, s.bounce, st);
        return kIdle;
    }
    begin_query(q, s.ox, s.oy, s.oz, s.dx, s.dy, s.dz, kBig);
    ++st.q_b;
    return kBounce;
}

// A finished bounce query (steam_rt.py:4744-4769): termination tests, hit
// counts and the hit context.  Returns false if the ray has ended.
template <bool VISUAL = true>
__device__ __forceinline__ bool close_bounce(const SceneView& sv, const RayParams& p, const Outputs& out,
                                            const Lds& lds, RayState& s, HitContext& c, const Query& q,
                                            LaneStats& st) {
    const int bounce = s.bounce;
    const Closest h = q.h;
    if (out.log_tri != nullptr && s.group == 0 && s.ray < out.log_rays && bounce < out.log_bounces) {
        const size_t li = (size_t)s.ray * out.log_bounces + bounce;
        out.log_tri[li] = h.tri;
        out.log_t[li] = h.t;
        if (out.log_point != nullptr) {
            out.log_point[li * 3] = s.ox + h.t * s.dx;
            out.log_point[li * 3 + 1] = s.oy + h.t * s.dy;
            out.log_point[li * 3 + 2] = s.oz + h.t * s.dz;
        }
    }
    if (h.surf < 0) ++st.escaped;
    if (h.surf < 0 || h.t <= p.listener_radius || s.accum > p.max_path_len) {
        finish_ray(p, out, s, false, bounce + 1, st);
        return false;
    }
    if (bounce > 0) {
        if (sphere_before(s.ox, s.oy, s.oz, s.dx, s.dy, s.dz, p.lx, p.ly, p.lz, p.listener_radius, h.t)) {
            finish_ray(p, out, s, false, bounce + 1, st);
            return false;
        }
        unsigned long long m = s.mask;
        while (m) {
            const int k = __builtin_ctzll(m);
            m &= m - 1ull;
            const float4 S = p.sources[k];
            const float sdist =
                sqrtf((p.lx - S.x) * (p.lx - S.x) + (p.ly - S.y) * (p.ly - S.y) + (p.lz - S.z) * (p.lz - S.z));
            if (sdist > p.source_radius &&
                sphere_before(s.ox, s.oy, s.oz, s.dx, s.dy, s.dz, S.x, S.y, S.z, p.source_radius, h.t)) {
                s.mask &= ~(1ull << k);
                atomicMax(&out.actual_bounces[k], bounce + 1);
            }
        }
        if (s.mask == 0ull) {  // every source has recorded this ray's end
            stamp_ray(out, s.ray, s.group, 1);
            return false;
        }
    }
    const float4 nn = sv.normals[h.tri];
    float nx = nn.x, ny = nn.y, nz = nn.z;
    if (nx * s.dx + ny * s.dy + nz * s.dz > 0.0f) {
        nx = -nx;
        ny = -ny;
        nz = -nz;
    }
    c.surf = h.surf;
    c.t = h.t;
    c.nx = nx;
    c.ny = ny;
    c.nz = nz;
    c.hx = s.ox + h.t * s.dx + kHitOffset * nx;
    c.hy = s.oy + h.t * s.dy + kHitOffset * ny;
    c.hz = s.oz + h.t * s.dz + kHitOffset * nz;
    if (VISUAL && out.vis_limit > 0) {
        const int v = visual_slot(out, s.ray);
        if (v >= 0) {
            float* vp = out.vis_points + ((size_t)v * out.vis_bounces + bounce) * 3;
            vp[0] = c.hx;
            vp[1] = c.hy;
            vp[2] = c.hz;
            out.vis_surf[(size_t)v * out.vis_bounces + bounce] = c.surf;
        }
    }
    unsigned long long m = s.mask;
    while (m) {
        const int k = __builtin_ctzll(m);
        m &= m - 1ull;
        if (lds.hits)
            add_shared(&lds.hits[c.surf], 1);
        else
            add_global(&out.hit_counts[(size_t)k * sv.num_surfaces + c.surf], 1);
    }
    c.nee = s.mask;
    return true;
}

// Multi-source traces: the candidate connections (ray, source) of the
// wave's waiting rays are spread over its free lanes (lanes not traversing: their
// query and connection registers are unused during a completion pass), each
// free lane computing and parking one connection with the same arithmetic as
// the ray's own lane would (connect() on the owner's ray and hit, read with
// lane shuffles).  Without it a ray makes its group's connections one per
// loop round (about 6 rounds of ~10 lanes per pass on 64-keyframe traces;
// measured: batched 64-keyframe traces +21 %).  Takes at most as many
// as the queue has room for; the rest stay in the owners' c.nee for the
// per-lane loop.  Every lane of the wave calls it.
__device__ __forceinline__ void spread_connections(const SceneView& sv, const RayParams& p, const RayState& s,
                                                   HitContext& c, bool owner, bool free_lane, Query& q,
                                                   Connection& cn, float* qu, int q_head, int& q_count,
                                                   int lane, unsigned long long below, LaneStats& st) {
    const int cnt = owner ? __popcll(c.nee) : 0;
    int prefix = 0, total = 0;
#pragma unroll
    for (int b = 0; b < 7; ++b) {  // cnt <= 64: 7-bit wave prefix sum from ballots
        const unsigned long long bb = __ballot((cnt >> b) & 1);
        prefix += __popcll(bb & below) << b;
        total += __popcll(bb) << b;
    }
    const unsigned long long fb = __ballot(free_lane);
    const int take = min(min(total, __popcll(fb)), kQueueCap - q_count);
    if (take <= 0) return;
    const int r = __popcll(fb & below);  // this lane's rank among the free lanes
    const bool helper = free_lane && r < take;
    // owner of candidate r: the first lane whose inclusive prefix exceeds r
    const int incl = prefix + cnt;
    int lo = 0, hi = 63;
#pragma unroll
    for (int step = 0; step < 6; ++step) {
        const int mid = (lo + hi) >> 1;
        const int v = __shfl(incl, mid);
        if (v > r)
            hi = mid;
        else
            lo = mid + 1;
    }
    const int o = helper ? lo : lane;
    const int o_prefix = __shfl(prefix, o);
    const unsigned o_lo = __shfl((unsigned)c.nee, o), o_hi = __shfl((unsigned)(c.nee >> 32), o);
    RayState ts;
    HitContext tc;
    ts.dx = __shfl(s.dx, o);
    ts.dy = __shfl(s.dy, o);
    ts.dz = __shfl(s.dz, o);
    ts.d0x = __shfl(s.d0x, o);
    ts.d0y = __shfl(s.d0y, o);
    ts.d0z = __shfl(s.d0z, o);
#pragma unroll
    for (int b = 0; b < kBands; ++b) ts.A[b] = __shfl(s.A[b], o);
    ts.accum = __shfl(s.accum, o);
    tc.hx = __shfl(c.hx, o);
    tc.hy = __shfl(c.hy, o);
    tc.hz = __shfl(c.hz, o);
    tc.nx = __shfl(c.nx, o);
    tc.ny = __shfl(c.ny, o);
    tc.nz = __shfl(c.nz, o);
    tc.t = __shfl(c.t, o);
    tc.surf = __shfl(c.surf, o);
    bool parked = false;
    if (helper) {
        unsigned long long m = ((unsigned long long)o_hi << 32) | o_lo;
        for (int j = r - o_prefix; j > 0; --j) m &= m - 1ull;  // the j-th pending source, ascending
        const int k = __builtin_ctzll(m);
        if (connect(sv, p, ts, tc, k, q, cn)) {
            ++st.q_s;
            parked = cn.bin >= 0;
        }
    }
    const unsigned long long pb = __ballot(parked);
    if (parked) park(qu, (q_head + q_count + __popcll(pb & below)) & (kQueueCap - 1), q, cn);
    q_count += __popcll(pb);
    // owners drop the candidates handed out (the lowest ones)
    if (owner) {
        int taken = min(max(take - prefix, 0), cnt);
        while (taken-- > 0) c.nee &= c.nee - 1ull;
    }
}

// Hand out tasks to the lanes of this wave that need one: one atomic per
// wave on `cursor`; lanes past `limit` get -1.
__device__ __forceinline__ int claim_rays(bool need, unsigned* cursor, int limit) {
    const unsigned long long ballot = __ballot(need);
    if (ballot == 0ull) return -1;
    const int count = __popcll(ballot);
    const int leader = __ffsll((long long)ballot) - 1;
    const int lane = (int)(threadIdx.x & 63);
    unsigned base = 0;
    if (lane == leader) base = atomicAdd(cursor, (unsigned)count);
    base = __shfl(base, leader);
    if (!need) return -1;
    const unsigned long long below = ballot & ((1ull << lane) - 1ull);
    const unsigned idx = base + (unsigned)__popcll(below);
    return idx < (unsigned)limit ? (int)idx : -1;
}

// A wave runs its lanes' query completions (shading, connections, refill) in
// one pass once at least kFinishEighths/8 of its lanes wait for one (of the
// live lanes once the work pool is drained), or when no lane traverses.  The
// completion code is long and runs with only the waiting lanes, so it pays to
// gather much of the wave first.  Measured on MI355X over 2/8..8/8 with leaf
// tests made at once, 7/8 gave the shortest traces (8/8, pure lockstep, lets
// the slowest lane hold the wave); with gated leaf passes (one leaf per lane
// and pass), 4/8 / 5/8 / 6/8 / 7/8 / 8/8 gave batched mic arrays 863 / 860 /
// 838 / 807 / 728 RIRs/s, batched keyframes 1526 / 1543 / 1540 / 1513 /
// 1354 and a 200k-ray single trace 6.40 / 6.25 / 6.25 / 6.31 / 6.60 ms.
constexpr int kFinishEighths = 5;

// Steps per loop round of a traversing lane (node visits and leaf passes).
// Measured on MI355X against one visit per round, with leaves tested at
// once: 2 / 3 / 4 / 8 / 16 / 32 visits took a 200k-ray single trace -2 / -5 /
// -7 / -7 / -7 / -8 % time and batched traces +10..14 % throughput (4..16
// within noise of each other).  Ending the round as soon as the completion
// threshold is met was no better.  With gated leaf passes, 16 steps instead
// of 8: batched mic arrays -3.7 %, batched keyframes -1.6 %.
constexpr int kStepsPerRound = 8;

// Leaf pass gate: the waiting lanes' leaves are tested once
// waiting * kLeafGateDen >= walking * kLeafGateNum (or no lane walks).
// Measured on MI355X with the completion threshold at 7/8 and all of a
// lane's leaves tested in one pass (waiting : walking = 1:2 / 1:1 / 2:1 /
// 4:1 / 8:1 / only when none walks): batched mic arrays 746 / 758 / 772 /
// 768 / 751 / 679 RIRs/s, batched keyframes 1454 / 1457 / 1472 / 1440 /
// 1414 / 1327, 200k-ray single trace 6.24 / 6.25 / 6.25 / 6.30 / 6.46 /
// 6.77 ms; leaves tested at once: 681 / 1344 / 6.78 ms.  With one leaf per
// pass and the threshold at 5/8, 3:2 against 2:1: 861 / 859 RIRs/s,
// 1543 / 1535, 6.27 / 6.28 ms (3:2 kept).
constexpr int kLeafGateNum = 3;
constexpr int kLeafGateDen = 2;

// Trace the tasks [first, first + count) handed out by `cursor` (relative to
// `first`) with the calling wave's lanes.  A lane holds a ray (bounce queries,
// and its own connections when the queue is full) or a parked connection of
// another lane's ray; idle lanes take parked connections first, then new rays.
template <int WG, bool LDS_HIST, bool VISUAL = true, bool INTERLEAVE = false>
__device__ void trace_rays_refill(const SceneView& sv, const RayParams& p, const Outputs& out, const Lds& lds,
                                  unsigned* cursor, int first, int count, LaneStats& st) {
    unsigned short* stk = lds.stack + threadIdx.x;
    // the wave's index, provably uniform (readfirstlane), so that its queue
    // address is a scalar
    const int wave = (int)__builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    float* qu = lds.queue + (size_t)wave * (kQueueFields * kQueueCap);
    float* lg = lds.log + (size_t)wave * (kLogFields * kLogCap);
    int log_count = 0;  // wave-uniform: entries in the deposit log
    const int lane = (int)(threadIdx.x & 63);
    const unsigned long long below = (1ull << lane) - 1ull;
    RayState s;
    HitContext c;
    Query q;
    Connection cn;
    q.node = -1;
    q.occluded = false;
    int phase = kIdle;
    bool own = false;        // kShadow: the connection is this lane's ray's, which waits for it
    int q_head = 0, q_count = 0;  // wave-uniform ring state of the parked connections
    bool exhausted = false;  // wave-uniform: a claim came back empty, so the pool is drained
    while (true) {
        bool walking = phase != kIdle && (q.node >= 0 || q.pend >= 0);
        const int n_act = __popcll(__ballot(phase != kIdle));
        const int n_trav = __popcll(__ballot(walking));
        const bool drained = exhausted && q_count == 0;  // nothing left for idle lanes
        const int n_wait = (n_act - n_trav) + (drained ? 0 : 64 - n_act);
        if (n_wait == 0 && n_trav == 0) break;
        if (n_wait > 0 && (n_trav == 0 || n_wait * 8 >= (drained ? n_act : 64) * kFinishEighths)) {
            bool nee = false;  // the lane's ray has connections of its current hit to resolve
            if (!LDS_HIST)
                log_deposits(phase == kShadow && !walking && !q.occluded, p, out, cn, lg, log_count, lane, below);
            if (phase != kIdle && !walking) {
                if (phase == kBounce) {
                    nee = close_bounce<VISUAL>(sv, p, out, lds, s, c, q, st);
                    if (!nee) phase = kIdle;
                } else {
                    if (!q.occluded) {
                        const float esum = deposit<LDS_HIST, !LDS_HIST>(p, out, lds, sv.num_surfaces, cn, st);
                        if (VISUAL && own && out.vis_limit > 0) visual_update(out, s, c, q, esum);
                    }
                    nee = own;
                    if (!own) phase = kIdle;
                }
            }
            {
                // visual candidates keep resolving their own connections
                const bool spread = nee && c.nee != 0ull &&
                                    !(VISUAL && out.vis_limit > 0 && visual_slot(out, s.ray) >= 0);
                // two spreading rounds cover a pass's connections in practice
                // (measured: one round -6 % on batched 64-keyframe traces,
                // three no better than two)
                // single-source rays make their one connection in their own
                // lane (measured against spreading it: batched mic arrays
                // +3.3 %, single trace -3.8 % time)
                if (p.num_sources > 1)
                    for (int round = 0; round < 2; ++round)
                        spread_connections(sv, p, s, c, spread && c.nee != 0ull, !walking, q, cn, qu, q_head,
                                           q_count, lane, below, st);
            }
            // park the new connections while the queue has room; a lane whose
            // connection does not fit resolves it itself before going on
            bool cont = false;  // the ray goes on once its hit has no connection left to make
            auto connect_round = [&](auto one_source) {
                bool has = false;
                if (nee) {
                    if constexpr (decltype(one_source)::value)
                        has = only_connection(sv, p, s, c, q, cn, st);
                    else
                        has = next_connection(sv, p, s, c, q, cn, st);
                }
                // a visual candidate resolves its own connections (its lane
                // holds the ray state its slot update needs)
                const bool self = VISUAL && has && out.vis_limit > 0 && visual_slot(out, s.ray) >= 0;
                const unsigned long long want = __ballot(has && !self);
                const int room = kQueueCap - q_count;
                bool parked = false;
                if (has) {
                    const int rank = __popcll(want & below);
                    if (!self && rank < room) {
                        park(qu, (q_head + q_count + rank) & (kQueueCap - 1), q, cn);
                        parked = true;
                    } else {
                        phase = kShadow;
                        own = true;
                        nee = false;
                    }
                }
                q_count += min(__popcll(want), room);
                // after parking its last connection the ray goes on in the
                // same pass
                if (nee && (!has || (parked && c.nee == 0ull))) {
                    nee = false;
                    cont = true;
                }
            };
            // A hit of a single-source trace has at most one connection, so
            // one round resolves every lane and needs no loop.  The loop
            // carries the query and connection registers that its rounds
            // rewrite under divergent control flow, and the compiler copies
            // them (about 30 registers) on entry and at every exit; the
            // continuation, run after the rounds, keeps the ray state out of
            // it too.  Measured on MI355X against the loop for every trace
            // with the continuation inside it (which also spilled ~10 vector
            // registers to scratch): batched mic arrays +3.0 %, single trace
            // -4.4 % time, batched keyframes +2.9 %.
            if (p.num_sources == 1) {
                if (__any(nee)) connect_round(std::true_type{});
            } else {
                while (__any(nee)) connect_round(std::false_type{});
            }
            if (cont) phase = continue_ray(sv, p, out, s, c, q, st);
            if (q_count > 0) {
                wave_sync();  // parked connections are visible to every lane of the wave
                const unsigned long long idle = __ballot(phase == kIdle);
                const int take = min(__popcll(idle), q_count);
                if (phase == kIdle) {
                    const int rank = __popcll(idle & below);
                    if (rank < take) {
                        unpark(qu, (q_head + rank) & (kQueueCap - 1), q, cn);
                        phase = kShadow;
                        own = false;
                    }
                }
                q_head = (q_head + take) & (kQueueCap - 1);
                q_count -= take;
                wave_sync();  // slots are read before later passes reuse them
            }
            if (!exhausted) {
                const bool need = phase == kIdle;
                const int got = claim_rays(need, cursor, count);
                if (need && got >= 0 && start_ray<VISUAL, INTERLEAVE>(p, out, first + got, s, st)) {
                    begin_query(q, s.ox, s.oy, s.oz, s.dx, s.dy, s.dz, kBig);
                    ++st.q_b;
                    phase = kBounce;
                }
                exhausted = __any(need && got < 0) != 0;
            }
            walking = phase != kIdle && (q.node >= 0 || q.pend >= 0);
        }
        if (walking) {
            const bool shadow = phase == kShadow;
            unsigned nodes = 0, tris = 0;
            // Up to kStepsPerRound node visits before the wave looks at its
            // completions again: the round's bookkeeping (ballots, the
            // completion test, reloads of spilled scalars) is paid once per
            // several visits.  A lane whose query ends drops out and waits.
            if (!sv.use_bvh) {
                // brute force (the reference's use_bvh=False): every triangle
                test_leaf(sv, 0, sv.num_tris, q, shadow, tris);
                q.node = -1;
            } else {
                // Leaf tests are gathered: a lane whose node visit hits leaf
                // children waits (q.pend) while the others walk on, and the
                // wave tests the waiting lanes' leaves together once the gate
                // opens (kLeafGateNum : kLeafGateDen waiting : walking lanes,
                // or none walks).  A leaf's triangle pair costs about a node
                // visit, and tested at once (about one pass per step) it ran
                // with a few lanes of the wave.
#pragma unroll 1
                for (int k = 0; k < kStepsPerRound; ++k) {
                    const int n_step = __popcll(__ballot(q.node >= 0 && q.pend < 0));
                    const int n_pend = __popcll(__ballot(q.pend >= 0));
                    if (n_pend > 0 && n_pend * kLeafGateDen >= n_step * kLeafGateNum && q.pend >= 0)
                        leaf_pass(sv, q, shadow, tris);
                    const bool step = q.node >= 0 && q.pend < 0;
                    if (!__any(step)) break;  // only when no lane waits either
                    if (step) node_step<WG>(sv, stk, q, shadow, nodes);
                }
            }
            if (shadow) {
                st.ns += nodes;
                st.ts += tris;
            } else {
                st.nb += nodes;
                st.tb += tris;
            }
        }
    }
    if (!LDS_HIST && log_count > 0) flush_log(p, out, lg, log_count, lane);
}

// Wave reduction of the lane tallies, one atomic per wave and counter.
__device__ __forceinline__ void flush_stats(LaneStats& st, const RayParams& p, const Outputs& out) {
    const int num_sources = p.num_sources;
    if (p.groups > 1) flush_group_tallies(p, out, st);
    unsigned long long vals[8] = {st.q_b, st.q_s, st.nb, st.ns, st.tb, st.ts, st.contrib, st.escaped};
#pragma unroll
    for (int c = 0; c < 8; ++c) {
        unsigned long long v = vals[c];
        for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
        vals[c] = v;
    }
    int mi = st.max_iter_full, ac = st.active_full;
    for (int off = 32; off > 0; off >>= 1) {
        mi = max(mi, __shfl_xor(mi, off));
        ac += __shfl_xor(ac, off);
    }
    if ((threadIdx.x & 63) == 0) {
#pragma unroll
        for (int c = 0; c < 8; ++c)
            if (vals[c]) atomicAdd(&out.counters[c], vals[c]);
        for (int k = 0; k < (p.groups > 1 ? 0 : num_sources); ++k) {
            if (mi) atomicMax(&out.actual_bounces[k], mi);
            if (ac) atomicAdd(&out.active_count[k], ac);
        }
    }
    st = LaneStats();
}

// ---------------------------------------------------------------------------
// host-side construction of the kernel argument structs
// ---------------------------------------------------------------------------

inline SceneView make_scene_view(const aa_scene_s* s, bool use_bvh) {
    SceneView sv;
    sv.nodes = reinterpret_cast<const float4*>(s->d_nodes);
    sv.tris = reinterpret_cast<const float4*>(s->d_tris);
    sv.normals = reinterpret_cast<const float4*>(s->d_normals);
    sv.reflection = s->d_reflection;
    sv.scattering = s->d_scattering;
    sv.num_nodes = (int)s->wide.nodes.size();
    sv.num_tris = s->num_triangles;
    sv.num_surfaces = s->num_surfaces;
    sv.use_bvh = use_bvh ? 1 : 0;
    return sv;
}

// The integer nearest to n (sqrt 5 - 1) / 2 that is coprime to n (1 for n < 3),
// so that t -> (t * stride) mod n visits every ray once.
inline int golden_stride(int n) {
    if (n < 3) return 1;
    for (long long s = std::llround(0.6180339887498949 * n); s < n; ++s)
        if (std::gcd(s, (long long)n) == 1) return (int)s;
    return 1;
}

inline RayParams make_ray_params(const aa_trace_params* p, const float* sources, const float* directions,
                                 const float* bank, const float* urand, const int* uidx) {
    RayParams r;
    r.directions = directions;
    r.bank = bank;
    r.urand = urand;
    r.uidx = uidx;
    r.sources = reinterpret_cast<const float4*>(sources);
    r.num_rays = p->num_rays;
    r.num_bounces = p->num_bounces;
    r.num_bins = p->num_bins;
    r.num_samples = p->num_diffuse_samples;
    r.num_sources = p->num_sources;
    r.src_group = p->sources_per_lane > 0 ? p->sources_per_lane : 8;
    if (r.src_group > r.num_sources) r.src_group = r.num_sources;
    r.groups = (r.num_sources + r.src_group - 1) / r.src_group;
    r.ray_stride = golden_stride(r.num_rays);
    r.rng_mode = p->rng_mode;
    r.amb = p->render_ambisonics ? 1 : 0;
    r.seed_lo = (unsigned)(p->seed & 0xffffffffull);
    r.seed_hi = (unsigned)(p->seed >> 32);
    r.lx = p->listener[0];
    r.ly = p->listener[1];
    r.lz = p->listener[2];
    r.fx = p->source_forward[0];
    r.fy = p->source_forward[1];
    r.fz = p->source_forward[2];
    r.bin_dur = p->bin_duration;
    r.c = p->speed_of_sound;
    r.max_path_len = p->max_path_length;
    r.listener_radius = p->listener_radius;
    r.source_radius = p->source_radius;
    r.dmin = p->irradiance_min_distance;
    r.spec_exp = p->specular_exponent;
    r.dip_w = p->dipole_weight;
    r.dip_p = p->dipole_power;
    // (float)(4 pi) / N, as the reference FP32 kernel (steam_rt_cuda.py:570)
    r.ray_scale = 12.566370614359172f / (float)(p->num_rays > 1 ? p->num_rays : 1);
    return r;
}

inline Outputs make_outputs(const aa_trace_outputs* o, const aa_trace_params* p) {
    Outputs out;
    out.echogram = o->d_echogram;
    out.ambisonic = o->d_ambisonic;
    out.hit_counts = o->d_hit_counts;
    out.contrib_counts = o->d_contrib_counts;
    out.surface_energy = o->d_surface_energy;
    out.actual_bounces = o->d_actual_bounces;
    out.active_count = o->d_active_count;
    out.counters = reinterpret_cast<unsigned long long*>(o->d_counters);
    const bool log = p->log_rays > 0 && p->log_bounces > 0;
    out.log_rays = log ? (p->log_rays < p->num_rays ? p->log_rays : p->num_rays) : 0;
    out.log_bounces = log ? p->log_bounces : 0;
    out.log_tri = log ? o->d_log_triangle : nullptr;
    out.log_t = log ? o->d_log_t : nullptr;
    out.log_point = log ? o->d_log_point : nullptr;
    const bool vis = p->visual_limit > 0;
    out.vis_limit = vis ? p->visual_limit : 0;
    out.vis_stride = vis ? p->visual_stride : 1;
    out.vis_bounces = p->num_bounces;
    out.vis_points = vis ? o->d_visual_points : nullptr;
    out.vis_surf = vis ? o->d_visual_surfaces : nullptr;
    out.vis_rays = vis ? o->d_visual_rays : nullptr;
    out.vis_orders = vis ? o->d_visual_orders : nullptr;
    out.vis_dist = vis ? o->d_visual_distances : nullptr;
    out.vis_gain = vis ? o->d_visual_gains : nullptr;
    out.ray_times = reinterpret_cast<unsigned long long*>(o->d_ray_times);
    return out;
}

// Byte extents of one trace's outputs (for zeroing).
struct OutputExtents {
    void* ptr[8];
    size_t bytes[8];
};

inline OutputExtents output_extents(const aa_trace_outputs* o, const aa_trace_params* p, int num_surfaces) {
    const size_t Ks = (size_t)p->num_sources, K = (size_t)p->num_bins, S = (size_t)num_surfaces;
    OutputExtents e;
    e.ptr[0] = o->d_echogram;
    e.bytes[0] = Ks * AA_NUM_BANDS * K * 4;
    e.ptr[1] = p->render_ambisonics ? o->d_ambisonic : nullptr;
    e.bytes[1] = Ks * AA_NUM_BANDS * 4 * K * 4;
    e.ptr[2] = o->d_hit_counts;
    e.bytes[2] = Ks * S * 4;
    e.ptr[3] = o->d_contrib_counts;
    e.bytes[3] = Ks * S * 4;
    e.ptr[4] = o->d_surface_energy;
    e.bytes[4] = Ks * S * 4;
    e.ptr[5] = o->d_actual_bounces;
    e.bytes[5] = Ks * 4;
    e.ptr[6] = o->d_active_count;
    e.bytes[6] = Ks * 4;
    e.ptr[7] = o->d_counters;
    e.bytes[7] = AA_CNT_COUNT * 8;
    return e;
}

}  // namespace aa
