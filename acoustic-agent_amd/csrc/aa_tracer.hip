// MI355X (gfx950) reflection tracer: the single-trace kernel and the C ABI
// of include/aa_tracer.h (device queries, canonical BVH, scenes, traces).
//
// One lane traces one listener ray through up to B bounces
// (acoustic_agent/steam_rt.py:4724-4849, device code in aa_device.h):
// closest hit through a 4-wide BVH (short per-lane stack in LDS), listener/source sphere termination, a next-event shadow
// query towards each source (any-hit with early exit, equal to the
// reference's closest-hit test `EPS < t < D - EPS`, steam_rt.py:4788-4792),
// the BRDF-weighted six-band deposit into the time histogram
// (steam_rt.py:4795-4826), then specular or diffuse continuation
// (steam_rt.py:4827-4848).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/aa_tracer.h"
#include "aa_device.h"
#include "aa_internal.h"

// ---------------------------------------------------------------------------
// error reporting
// ---------------------------------------------------------------------------

namespace {
thread_local std::string g_last_error;
}  // namespace

int aa::set_error(int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_last_error = buf;
    return code;
}

#define fail aa::set_error

#define AA_HIP(call)                                                                             \
    do {                                                                                         \
        hipError_t e_ = (call);                                                                  \
        if (e_ != hipSuccess) return fail(AA_ERR_DEVICE, "%s failed: %s", #call, hipGetErrorString(e_)); \
    } while (0)

// ---------------------------------------------------------------------------
// single-trace kernel
// ---------------------------------------------------------------------------

namespace aa {

struct TraceArgs {
    SceneView sv;
    RayParams p;
    Outputs out;
    int lds_surf;  // per-surface tallies privatised in LDS (single source)
};

// Zero a device range with a kernel (no hipMemsetAsync, so the call stays
// capturable into a HIP graph).
__global__ void fill_zero_kernel(unsigned* p, size_t words) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < words; i += (size_t)gridDim.x * blockDim.x)
        p[i] = 0u;
}

__global__ void fill_word_kernel(unsigned* p, size_t words, unsigned value) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < words; i += (size_t)gridDim.x * blockDim.x)
        p[i] = value;
}

// Persistent lanes over the rays of one trace (refilled as rays end).  The
// register budget is held to 3 waves per SIMD (168 VGPRs, no spills): the
// tracer needs waves in flight to hide its dependent loads.
template <int WG, bool LDS_HIST>
__global__ __launch_bounds__(WG) __attribute__((amdgpu_waves_per_eu(3))) void trace_kernel(const TraceArgs a) {
    extern __shared__ __align__(16) unsigned char smem[];
    const int S = a.sv.num_surfaces;
    unsigned short* s_stack = reinterpret_cast<unsigned short*>(smem);
    float* s_queue = reinterpret_cast<float*>(smem + stack_only_bytes(WG));
    float* s_hist = reinterpret_cast<float*>(smem + private_bytes(WG));
    const int hist_words = LDS_HIST ? a.p.num_sources * kBands * a.p.num_bins : 0;
    int* s_hits = reinterpret_cast<int*>(s_hist + hist_words);
    int* s_contrib = s_hits + (a.lds_surf ? S : 0);
    float* s_senergy = reinterpret_cast<float*>(s_contrib + (a.lds_surf ? S : 0));

    const int tid = threadIdx.x;
    if (LDS_HIST)
        for (int i = tid; i < hist_words; i += WG) s_hist[i] = 0.0f;
    if (a.lds_surf) {
        for (int i = tid; i < S; i += WG) {
            s_hits[i] = 0;
            s_contrib[i] = 0;
            s_senergy[i] = 0.0f;
        }
    }
    __syncthreads();

    Lds lds;
    lds.stack = s_stack;
    lds.queue = s_queue;
    lds.log = reinterpret_cast<float*>(smem + log_offset(WG));
    lds.hist = LDS_HIST ? s_hist : nullptr;
    lds.hits = a.lds_surf ? s_hits : nullptr;
    lds.contrib = a.lds_surf ? s_contrib : nullptr;
    lds.senergy = a.lds_surf ? s_senergy : nullptr;
    LaneStats st;
    // persistent lanes: each wave claims rays from the trace's cursor as its
    // lanes' rays end
    unsigned* cursor = reinterpret_cast<unsigned*>(a.out.counters + AA_CNT_RAYS_CLAIMED);
    trace_rays_refill<WG, LDS_HIST, true, true>(a.sv, a.p, a.out, lds, cursor, 0, a.p.num_rays * a.p.groups, st);
    flush_stats(st, a.p, a.out);
    __syncthreads();
    if (LDS_HIST) {
        for (int i = tid; i < hist_words; i += WG) {
            const float v = s_hist[i];
            if (v != 0.0f) atomicAdd(&a.out.echogram[i], v);
        }
    }
    if (a.lds_surf) {
        for (int i = tid; i < S; i += WG) {
            if (s_hits[i]) atomicAdd(&a.out.hit_counts[i], s_hits[i]);
            if (s_contrib[i]) atomicAdd(&a.out.contrib_counts[i], s_contrib[i]);
            if (s_senergy[i] != 0.0f) atomicAdd(&a.out.surface_energy[i], s_senergy[i]);
        }
    }
}

// ---------------------------------------------------------------------------
// plain kernel: the reference GPU backend's execution model on gfx950
// ---------------------------------------------------------------------------
//
// One thread per lane task, no persistence and no refill, blocks of 128
// (steam_rt_cuda.py:491-710): each thread walks the canonical median-split
// binary tree stacklessly with escape links (steam_rt_cuda.py:392-456) for
// its bounce queries and, like the reference, answers each shadow query with
// a full closest-hit walk tested against (EPS, D - EPS) (steam_rt_cuda.py:
// 636-661).  The bounce arithmetic is the shared one of aa_device.h, so the
// results equal the persistent kernel's; this kernel is the plain
// one-scene-per-launch baseline the persistent and batched paths are
// measured against.

struct PlainArgs {
    SceneView sv;
    const float4* nodes;  // [nodes, 2]: (lo | escape) (hi | start << 4 | count)
    const float4* tris;   // canonical primitive order
    int num_nodes;
    RayParams p;
    Outputs out;
};

// Ray-box test of the reference GPU path (steam_rt_cuda.py:392-411).
__device__ __forceinline__ bool plain_box(const Ray& r, const float4 lo, const float4 hi, float max_distance) {
    float enter = 0.0f, exit_ = max_distance + kBvhEps;
    const float o[3] = {r.ox, r.oy, r.oz}, d[3] = {r.dx, r.dy, r.dz};
    const float l[3] = {lo.x, lo.y, lo.z}, u[3] = {hi.x, hi.y, hi.z};
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        if (fabsf(d[a]) <= 1e-12f) {
            if (o[a] < l[a] || o[a] > u[a]) return false;
            continue;
        }
        float first = (l[a] - o[a]) / d[a], second = (u[a] - o[a]) / d[a];
        if (first > second) {
            const float t = first;
            first = second;
            second = t;
        }
        if (first > enter) enter = first;
        if (second < exit_) exit_ = second;
        if (enter > exit_) return false;
    }
    return exit_ > kEps;
}

// Closest hit by the stackless walk (steam_rt_cuda.py:429-449).
__device__ Closest plain_closest(const PlainArgs& a, const Ray& r, unsigned& node_cnt, unsigned& tri_cnt) {
    Closest h;
    h.t = kBig;
    h.surf = -1;
    h.tri = -1;
    // linear mode tests every triangle (steam_rt_cuda.py:415-426); the tie
    // rule makes the order irrelevant
    const int num_nodes = a.sv.use_bvh ? a.num_nodes : 1;
    int node = 0;
    while (node < num_nodes) {
        ++node_cnt;
        const float4 lo = a.nodes[2 * node], hi = a.nodes[2 * node + 1];
        if (!a.sv.use_bvh) {
            for (int k = 0; k < a.sv.num_tris; ++k) {
                ++tri_cnt;
                const float4* t4 = a.tris + (size_t)k * 3;
                const float4 ta = t4[0], tb = t4[1], tc = t4[2];
                float t;
                if (tri_test(ta, tb, tc, r.ox, r.oy, r.oz, r.dx, r.dy, r.dz, t)) {
                    const int tri = __float_as_int(ta.w), surf = __float_as_int(tb.w);
                    if (better(t, surf, tri, h)) {
                        h.t = t;
                        h.surf = surf;
                        h.tri = tri;
                    }
                }
            }
            break;
        }
        if (!plain_box(r, lo, hi, h.t)) {
            node = __float_as_int(lo.w);
            continue;
        }
        const int sc = __float_as_int(hi.w);
        const int count = sc & 15;
        for (int k = 0; k < count; ++k) {
            ++tri_cnt;
            const float4* t4 = a.tris + (size_t)((sc >> 4) + k) * 3;
            const float4 ta = t4[0], tb = t4[1], tc = t4[2];
            float t;
            if (tri_test(ta, tb, tc, r.ox, r.oy, r.oz, r.dx, r.dy, r.dz, t)) {
                const int tri = __float_as_int(ta.w), surf = __float_as_int(tb.w);
                if (better(t, surf, tri, h)) {
                    h.t = t;
                    h.surf = surf;
                    h.tri = tri;
                }
            }
        }
        ++node;
    }
    return h;
}

template <int WG>
__global__ __launch_bounds__(WG) void plain_kernel(const PlainArgs a) {
    const int task = (int)(blockIdx.x * WG + threadIdx.x);
    LaneStats st;
    Lds lds = {};  // every tally goes to global atomics
    if (task < a.p.num_rays * a.p.groups) {
        RayState s;
        if (start_ray<false>(a.p, a.out, task, s, st)) {  // no visual candidates here
            Query q;
            begin_query(q, s.ox, s.oy, s.oz, s.dx, s.dy, s.dz, kBig);
            ++st.q_b;
            while (true) {
                q.h = plain_closest(a, q.r, st.nb, st.tb);
                HitContext c;
                if (!close_bounce<false>(a.sv, a.p, a.out, lds, s, c, q, st)) break;
                unsigned long long m = c.nee;
                while (m) {  // next-event connections (steam_rt.py:4770-4826)
                    const int k = __builtin_ctzll(m);
                    m &= m - 1ull;
                    const float4 S = a.p.sources[k];
                    const float tsx = S.x - c.hx, tsy = S.y - c.hy, tsz = S.z - c.hz;
                    const float D = sqrtf(tsx * tsx + tsy * tsy + tsz * tsz);
                    if (!(D > a.p.dmin)) continue;
                    const float sdx = tsx / D, sdy = tsy / D, sdz = tsz / D;
                    if (!(c.nx * sdx + c.ny * sdy + c.nz * sdz > 0.0f)) continue;
                    ++st.q_s;
                    Ray sr;
                    sr.ox = c.hx;
                    sr.oy = c.hy;
                    sr.oz = c.hz;
                    sr.dx = sdx;
                    sr.dy = sdy;
                    sr.dz = sdz;
                    const Closest sh = plain_closest(a, sr, st.ns, st.ts);
                    if (sh.surf >= 0 && sh.t > kEps && sh.t < D - kEps) continue;
                    Query unused;
                    Connection cn;
                    if (connect(a.sv, a.p, s, c, k, unused, cn) && cn.bin >= 0)
                        deposit<false>(a.p, a.out, lds, a.sv.num_surfaces, cn, st);
                }
                // the next bounce query (set up and counted by continue_ray)
                if (continue_ray(a.sv, a.p, a.out, s, c, q, st) == kIdle) break;
            }
        }
    }
    flush_stats(st, a.p, a.out);
}

}  // namespace aa

// ---------------------------------------------------------------------------
// scene handle
// ---------------------------------------------------------------------------


namespace {

uint64_t fnv1a(uint64_t h, const void* data, size_t bytes) {
    const unsigned char* p = static_cast<const unsigned char*>(data);
    for (size_t i = 0; i < bytes; ++i) {
        h ^= p[i];
        h *= 1099511628211ull;
    }
    return h;
}

// FNV-1a 64 over the scene inputs in argument order: the key of a scene in
// the caches and on disk.
uint64_t scene_hash(const float* tri_vertices, const int32_t* tri_surface, const float* tri_normals,
                    int32_t num_triangles, const float* reflection, const float* scattering, int32_t num_surfaces,
                    int32_t leaf_size) {
    uint64_t h = 1469598103934665603ull;
    h = fnv1a(h, tri_vertices, (size_t)num_triangles * 9 * sizeof(float));
    h = fnv1a(h, tri_surface, (size_t)num_triangles * sizeof(int32_t));
    h = fnv1a(h, tri_normals, (size_t)num_triangles * 3 * sizeof(float));
    h = fnv1a(h, reflection, (size_t)num_surfaces * 6 * sizeof(float));
    h = fnv1a(h, scattering, (size_t)num_surfaces * sizeof(float));
    return fnv1a(h, &leaf_size, sizeof(leaf_size));
}

void free_scene(aa_scene_s* s) {
    if (!s) return;
    (void)hipSetDevice(s->device);
    (void)hipFree(s->d_nodes);
    (void)hipFree(s->d_tris);
    (void)hipFree(s->d_plain_nodes);
    (void)hipFree(s->d_plain_tris);
    (void)hipFree(s->d_normals);
    (void)hipFree(s->d_reflection);
    (void)hipFree(s->d_scattering);
    delete s;
}

std::mutex g_props_mutex;
std::vector<aa::DeviceProps> g_props;

}  // namespace

const aa::DeviceProps& aa::device_props(int device) {
    std::lock_guard<std::mutex> lock(g_props_mutex);
    if ((int)g_props.size() <= device) g_props.resize((size_t)device + 1);
    aa::DeviceProps& p = g_props[(size_t)device];
    if (!p.valid) {
        hipDeviceProp_t prop;
        if (hipGetDeviceProperties(&prop, device) == hipSuccess) {
            p.cu = prop.multiProcessorCount;
            p.lds_per_block = std::max(prop.sharedMemPerBlock, prop.sharedMemPerBlockOptin);
            p.lds_per_cu = std::max(prop.maxSharedMemoryPerMultiProcessor, p.lds_per_block);
            p.valid = true;
        }
    }
    return p;
}

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------

extern "C" {

const char* aa_last_error(void) { return g_last_error.c_str(); }

int aa_device_count(int32_t* count) {
    if (!count) return fail(AA_ERR_INVALID, "count is NULL");
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess) {
        *count = 0;
        return fail(AA_ERR_DEVICE, "hipGetDeviceCount failed: %s", hipGetErrorString(e));
    }
    *count = n;
    return AA_OK;
}

int aa_device_info(int32_t device, char* name, int32_t name_len, char* arch, int32_t arch_len, uint64_t* hbm_bytes,
                   int32_t* compute_units) {
    hipDeviceProp_t prop;
    AA_HIP(hipGetDeviceProperties(&prop, device));
    if (name && name_len > 0) snprintf(name, (size_t)name_len, "%s", prop.name);
    if (arch && arch_len > 0) snprintf(arch, (size_t)arch_len, "%s", prop.gcnArchName);
    if (hbm_bytes) *hbm_bytes = (uint64_t)prop.totalGlobalMem;
    if (compute_units) *compute_units = prop.multiProcessorCount;
    return AA_OK;
}

static int build_bvh_common(bool sah, const double* bounds_min, const double* bounds_max, int64_t n, int64_t leaf_size, double* node_min,
                 double* node_max, int64_t* start, int64_t* count, int64_t* escape, int64_t* primitives,
                 int64_t* stats) {
    if (!bounds_min || !bounds_max || !node_min || !node_max || !start || !count || !escape || !primitives || !stats)
        return fail(AA_ERR_INVALID, "aa_build_bvh: NULL argument");
    aa::CanonicalBVH bvh;
    std::string err = sah ? aa::build_sah_bvh(bounds_min, bounds_max, n, leaf_size, bvh)
                          : aa::build_canonical_bvh(bounds_min, bounds_max, n, leaf_size, bvh);
    if (!err.empty()) return fail(AA_ERR_INVALID, "%s", err.c_str());
    const size_t nodes = (size_t)bvh.node_count();
    std::memcpy(node_min, bvh.lo.data(), nodes * 3 * sizeof(double));
    std::memcpy(node_max, bvh.hi.data(), nodes * 3 * sizeof(double));
    std::memcpy(start, bvh.start.data(), nodes * sizeof(int64_t));
    std::memcpy(count, bvh.count.data(), nodes * sizeof(int64_t));
    std::memcpy(escape, bvh.escape.data(), nodes * sizeof(int64_t));
    std::memcpy(primitives, bvh.prims.data(), bvh.prims.size() * sizeof(int64_t));
    stats[0] = (int64_t)nodes;
    stats[1] = bvh.leaf_count;
    stats[2] = bvh.max_depth;
    return AA_OK;
}

int aa_build_bvh(const double* bounds_min, const double* bounds_max, int64_t n, int64_t leaf_size, double* node_min,
                 double* node_max, int64_t* start, int64_t* count, int64_t* escape, int64_t* primitives,
                 int64_t* stats) {
    return build_bvh_common(false, bounds_min, bounds_max, n, leaf_size, node_min, node_max, start, count, escape,
                            primitives, stats);
}

int aa_build_bvh_sah(const double* bounds_min, const double* bounds_max, int64_t n, int64_t leaf_size,
                     double* node_min, double* node_max, int64_t* start, int64_t* count, int64_t* escape,
                     int64_t* primitives, int64_t* stats) {
    return build_bvh_common(true, bounds_min, bounds_max, n, leaf_size, node_min, node_max, start, count, escape,
                            primitives, stats);
}

int aa_scene_hash(const float* tri_vertices, const int32_t* tri_surface, const float* tri_normals,
                  int32_t num_triangles, const float* reflection, const float* scattering, int32_t num_surfaces,
                  int32_t leaf_size, uint64_t* hash_out) {
    if (!hash_out || num_triangles < 0 || num_surfaces < 0 ||
        (num_triangles > 0 && (!tri_vertices || !tri_surface || !tri_normals)) ||
        (num_surfaces > 0 && (!reflection || !scattering)))
        return fail(AA_ERR_INVALID, "aa_scene_hash: NULL array or negative count");
    *hash_out = scene_hash(tri_vertices, tri_surface, tri_normals, num_triangles, reflection, scattering,
                           num_surfaces, leaf_size);
    return AA_OK;
}

int aa_scene_create(int32_t device, const float* tri_vertices, const int32_t* tri_surface, const float* tri_normals,
                    int32_t num_triangles, const float* reflection, const float* scattering, int32_t num_surfaces,
                    int32_t leaf_size, aa_scene_t* out) {
    if (!out) return fail(AA_ERR_INVALID, "out is NULL");
    *out = nullptr;
    if (!tri_vertices || !tri_surface || !tri_normals || !reflection || !scattering)
        return fail(AA_ERR_INVALID, "aa_scene_create: NULL input array");
    if (num_triangles < 1) return fail(AA_ERR_INVALID, "scene needs at least one triangle");
    if (num_surfaces < 1) return fail(AA_ERR_INVALID, "scene needs at least one surface");
    if (leaf_size < 1 || leaf_size > aa::kMaxLeafSize) return fail(AA_ERR_INVALID, "leaf_size must be 1..8");
    if (num_triangles > (1 << 27)) return fail(AA_ERR_LIMIT, "too many triangles");
    for (int32_t k = 0; k < num_triangles; ++k) {
        if (tri_surface[k] < 0 || tri_surface[k] >= num_surfaces)
            return fail(AA_ERR_INVALID, "triangle %d has surface id %d outside [0,%d)", k, tri_surface[k], num_surfaces);
        for (int c = 0; c < 9; ++c)
            if (!std::isfinite(tri_vertices[(size_t)k * 9 + c]))
                return fail(AA_ERR_INVALID, "triangle %d has a non-finite vertex", k);
    }
    int ndev = 0;
    AA_HIP(hipGetDeviceCount(&ndev));
    if (device < 0 || device >= ndev) return fail(AA_ERR_DEVICE, "device %d not available (%d devices)", device, ndev);

    // canonical BVH over triangle bounds padded like the reference FP32 path
    std::vector<double> lo((size_t)num_triangles * 3), hi((size_t)num_triangles * 3);
    for (int32_t k = 0; k < num_triangles; ++k) {
        for (int a = 0; a < 3; ++a) {
            double mn = INFINITY, mx = -INFINITY;
            for (int v = 0; v < 3; ++v) {
                const double x = (double)tri_vertices[(size_t)k * 9 + v * 3 + a];
                mn = std::min(mn, x);
                mx = std::max(mx, x);
            }
            lo[(size_t)k * 3 + a] = mn - (double)aa::kBvhEps;
            hi[(size_t)k * 3 + a] = mx + (double)aa::kBvhEps;
        }
    }
    // The reference's median-split tree (canonical; its size is reported) and
    // the SAH tree the device walks (same hits, fewer node visits).
    aa::CanonicalBVH canonical;
    std::string err = aa::build_canonical_bvh(lo.data(), hi.data(), num_triangles, leaf_size, canonical);
    if (!err.empty()) return fail(AA_ERR_INVALID, "%s", err.c_str());
    aa::CanonicalBVH bin;
    err = aa::build_sah_bvh(lo.data(), hi.data(), num_triangles, leaf_size, bin);
    if (!err.empty()) return fail(AA_ERR_INVALID, "%s", err.c_str());

    aa_scene_s* s = new aa_scene_s();
    s->device = device;
    s->num_triangles = num_triangles;
    s->num_surfaces = num_surfaces;
    s->leaf_size = leaf_size;
    s->canonical_nodes = (int32_t)canonical.node_count();
    s->canonical_depth = (int32_t)canonical.max_depth;
    aa::collapse_to_wide(bin, s->wide);
    if (3 * (s->wide.depth - 1) > aa::kStack) {
        free_scene(s);
        return fail(AA_ERR_LIMIT, "BVH depth %d exceeds the traversal stack", s->wide.depth);
    }
    if ((int64_t)s->wide.nodes.size() > aa::kMaxWideNodes) {
        free_scene(s);
        return fail(AA_ERR_LIMIT, "%zu BVH nodes exceed the 16-bit traversal stack", s->wide.nodes.size());
    }

    // triangles in a tree's primitive order: (v0, tri id) (e1, surface) (e2, 0)
    auto pack_triangles = [&](const std::vector<int64_t>& prims) {
        std::vector<float> packed((size_t)num_triangles * 12);
        for (int32_t k = 0; k < num_triangles; ++k) {
            const int64_t tri = prims[(size_t)k];
            const float* v = tri_vertices + (size_t)tri * 9;
            float* o = packed.data() + (size_t)k * 12;
            o[0] = v[0];
            o[1] = v[1];
            o[2] = v[2];
            int32_t id = (int32_t)tri;
            std::memcpy(&o[3], &id, 4);
            o[4] = v[3] - v[0];
            o[5] = v[4] - v[1];
            o[6] = v[5] - v[2];
            int32_t sid = tri_surface[tri];
            std::memcpy(&o[7], &sid, 4);
            o[8] = v[6] - v[0];
            o[9] = v[7] - v[1];
            o[10] = v[8] - v[2];
            o[11] = 0.0f;
        }
        return packed;
    };
    const std::vector<float> tris = pack_triangles(bin.prims);
    const std::vector<float> plain_tris = pack_triangles(canonical.prims);
    // canonical nodes in FP32 (the reference GPU path's AABBs, steam_rt_cuda.py:392-411)
    const int64_t ncan = canonical.node_count();
    std::vector<float> plain_nodes((size_t)ncan * 8);
    for (int64_t n = 0; n < ncan; ++n) {
        float* o = plain_nodes.data() + (size_t)n * 8;
        for (int a = 0; a < 3; ++a) {  // rounded outwards, so no box loses a hit
            const double l = canonical.lo[(size_t)n * 3 + a], u = canonical.hi[(size_t)n * 3 + a];
            float fl = (float)l, fu = (float)u;
            if ((double)fl > l) fl = std::nextafter(fl, -INFINITY);
            if ((double)fu < u) fu = std::nextafter(fu, INFINITY);
            o[a] = fl;
            o[4 + a] = fu;
        }
        const int32_t esc = (int32_t)canonical.escape[(size_t)n];
        const int32_t sc = (int32_t)((canonical.start[(size_t)n] << 4) | canonical.count[(size_t)n]);
        std::memcpy(&o[3], &esc, 4);
        std::memcpy(&o[7], &sc, 4);
    }
    s->plain_node_count = (int32_t)ncan;
    std::vector<float> normals((size_t)num_triangles * 4);
    for (int32_t k = 0; k < num_triangles; ++k) {
        normals[(size_t)k * 4 + 0] = tri_normals[(size_t)k * 3 + 0];
        normals[(size_t)k * 4 + 1] = tri_normals[(size_t)k * 3 + 1];
        normals[(size_t)k * 4 + 2] = tri_normals[(size_t)k * 3 + 2];
        normals[(size_t)k * 4 + 3] = 0.0f;
    }
    s->content_hash = scene_hash(tri_vertices, tri_surface, tri_normals, num_triangles, reflection, scattering,
                                 num_surfaces, leaf_size);

    hipError_t e = hipSetDevice(device);
    const size_t node_bytes = s->wide.nodes.size() * sizeof(aa::WideNode);
    const size_t tri_bytes = tris.size() * sizeof(float);
    const size_t nrm_bytes = normals.size() * sizeof(float);
    const size_t refl_bytes = (size_t)num_surfaces * 6 * sizeof(float);
    const size_t scat_bytes = (size_t)num_surfaces * sizeof(float);
    const size_t plain_node_bytes = plain_nodes.size() * sizeof(float);
    if (e == hipSuccess) e = hipMalloc((void**)&s->d_nodes, node_bytes);
    if (e == hipSuccess) e = hipMalloc((void**)&s->d_plain_nodes, plain_node_bytes);
    if (e == hipSuccess) e = hipMalloc((void**)&s->d_plain_tris, tri_bytes);
    if (e == hipSuccess) e = hipMalloc((void**)&s->d_tris, tri_bytes);
    if (e == hipSuccess) e = hipMalloc((void**)&s->d_normals, nrm_bytes);
    if (e == hipSuccess) e = hipMalloc((void**)&s->d_reflection, refl_bytes);
    if (e == hipSuccess) e = hipMalloc((void**)&s->d_scattering, scat_bytes);
    if (e == hipSuccess) e = hipMemcpy(s->d_nodes, s->wide.nodes.data(), node_bytes, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(s->d_tris, tris.data(), tri_bytes, hipMemcpyHostToDevice);
    if (e == hipSuccess)
        e = hipMemcpy(s->d_plain_nodes, plain_nodes.data(), plain_node_bytes, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(s->d_plain_tris, plain_tris.data(), tri_bytes, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(s->d_normals, normals.data(), nrm_bytes, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(s->d_reflection, reflection, refl_bytes, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(s->d_scattering, scattering, scat_bytes, hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        free_scene(s);
        return fail(AA_ERR_MEMORY, "scene upload failed: %s", hipGetErrorString(e));
    }
    s->device_bytes = node_bytes + 2 * tri_bytes + plain_node_bytes + nrm_bytes + refl_bytes + scat_bytes;
    *out = s;
    return AA_OK;
}

int aa_scene_info(aa_scene_t scene, aa_scene_info_t* info) {
    if (!scene || !info) return fail(AA_ERR_INVALID, "aa_scene_info: NULL argument");
    info->num_triangles = scene->num_triangles;
    info->num_surfaces = scene->num_surfaces;
    info->num_nodes = (int32_t)scene->wide.nodes.size();
    info->depth = scene->wide.depth;
    info->canonical_nodes = scene->canonical_nodes;
    info->canonical_depth = scene->canonical_depth;
    info->leaf_size = scene->leaf_size;
    info->device = scene->device;
    info->device_bytes = scene->device_bytes;
    info->content_hash = scene->content_hash;
    return AA_OK;
}

int aa_scene_release(aa_scene_t scene) {
    if (!scene) return fail(AA_ERR_INVALID, "aa_scene_release: NULL scene");
    free_scene(scene);
    return AA_OK;
}

int aa_stream_sync(void* stream) {
    AA_HIP(hipStreamSynchronize(static_cast<hipStream_t>(stream)));
    return AA_OK;
}

int aa_stream_query(void* stream, int32_t* done) {
    if (!done) return fail(AA_ERR_INVALID, "aa_stream_query: NULL done");
    const hipError_t e = hipStreamQuery(static_cast<hipStream_t>(stream));
    if (e == hipSuccess) {
        *done = 1;
        return AA_OK;
    }
    if (e == hipErrorNotReady) {
        (void)hipGetLastError();  // not an error: clear it
        *done = 0;
        return AA_OK;
    }
    return fail(AA_ERR_DEVICE, "hipStreamQuery failed: %s", hipGetErrorString(e));
}

}  // extern "C"

namespace {

template <int WG, bool LDS>
hipError_t launch_trace(const aa::TraceArgs& args, int grid_cap, int per_cu, int cus, size_t lds, hipStream_t stream) {
    auto kern = aa::trace_kernel<WG, LDS>;
    if (lds > 64 * 1024) {
        hipError_t e = hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return e;
    }
    if (per_cu <= 0) {
        int blocks = 0;
        hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, (const void*)kern, WG, lds);
        per_cu = (e == hipSuccess && blocks > 0) ? blocks : 1;
    }
    const int grid = std::max(1, std::min(grid_cap, cus * per_cu));
    hipLaunchKernelGGL(kern, dim3(grid), dim3(WG), lds, stream, args);
    return hipGetLastError();
}

}  // namespace

int aa::validate_trace(const aa_scene_s* scene, const aa_trace_params* p, const float* d_sources,
                       const float* d_directions, const float* d_diffuse_bank, const float* d_diffuse_random,
                       const int32_t* d_diffuse_indices, const aa_trace_outputs* o) {
    if (!scene || !p || !o) return fail(AA_ERR_INVALID, "aa_trace: NULL argument");
    if (p->num_rays < 1) return fail(AA_ERR_INVALID, "num_rays must be >= 1");
    if (p->num_bounces < 0) return fail(AA_ERR_INVALID, "num_bounces must be >= 0");
    if (p->num_bins < 1) return fail(AA_ERR_INVALID, "num_bins must be >= 1");
    if (p->num_diffuse_samples < 1) return fail(AA_ERR_INVALID, "num_diffuse_samples must be >= 1");
    if (p->num_sources < 1 || p->num_sources > AA_MAX_SOURCES)
        return fail(AA_ERR_INVALID, "num_sources must be in [1, %d]", AA_MAX_SOURCES);
    if (!(p->bin_duration > 0.0f) || !(p->speed_of_sound > 0.0f))
        return fail(AA_ERR_INVALID, "bin_duration and speed_of_sound must be positive");
    if (p->rng_mode != AA_RNG_ARRAYS && p->rng_mode != AA_RNG_COUNTER) return fail(AA_ERR_INVALID, "unknown rng_mode");
    if (p->rng_mode == AA_RNG_ARRAYS && (!d_directions || !d_diffuse_random || !d_diffuse_indices))
        return fail(AA_ERR_INVALID, "AA_RNG_ARRAYS needs directions, diffuse_random and diffuse_indices");
    if (!d_sources || !d_diffuse_bank) return fail(AA_ERR_INVALID, "sources and diffuse_bank are required");
    if (!o->d_echogram || !o->d_hit_counts || !o->d_contrib_counts || !o->d_surface_energy || !o->d_actual_bounces ||
        !o->d_active_count || !o->d_counters)
        return fail(AA_ERR_INVALID, "aa_trace: NULL output");
    if (p->render_ambisonics && !o->d_ambisonic) return fail(AA_ERR_INVALID, "render_ambisonics needs d_ambisonic");
    if (p->log_rays > 0 && p->log_bounces > 0 && (!o->d_log_triangle || !o->d_log_t))
        return fail(AA_ERR_INVALID, "hit log needs d_log_triangle and d_log_t");
    if ((int64_t)p->num_sources * AA_NUM_BANDS * p->num_bins > (int64_t)1 << 30)
        return fail(AA_ERR_LIMIT, "histogram too large");
    {
        // lane tasks = rays x source groups (the grouping of make_ray_params);
        // kernels index tasks with 32-bit ints
        const int64_t per_lane = p->sources_per_lane > 0 ? std::min(p->sources_per_lane, p->num_sources)
                                                         : std::min(8, p->num_sources);
        const int64_t groups = (p->num_sources + per_lane - 1) / per_lane;
        if ((int64_t)p->num_rays * groups > (int64_t)1 << 30)
            return fail(AA_ERR_LIMIT, "num_rays x source groups = %lld lane tasks exceeds 2^30",
                        (long long)p->num_rays * groups);
    }
    if (p->visual_limit < 0) return fail(AA_ERR_INVALID, "visual_limit must be >= 0");
    if (p->visual_limit > 0) {
        if (p->num_sources != 1) return fail(AA_ERR_INVALID, "visual candidates need a single-source trace");
        if (p->visual_stride < 1) return fail(AA_ERR_INVALID, "visual_stride must be >= 1");
        if (!o->d_visual_points || !o->d_visual_surfaces || !o->d_visual_rays || !o->d_visual_orders ||
            !o->d_visual_distances || !o->d_visual_gains)
            return fail(AA_ERR_INVALID, "visual_limit > 0 needs every d_visual_* output");
        if ((int64_t)p->visual_limit * (p->num_bounces > 0 ? p->num_bounces : 1) * 3 > (int64_t)1 << 30)
            return fail(AA_ERR_LIMIT, "visual candidate buffers too large");
    }
    return AA_OK;
}

namespace {

hipError_t zero_range(void* ptr, size_t bytes, hipStream_t stream) {
    if (!ptr || bytes == 0) return hipSuccess;
    const size_t words = bytes / 4;
    const int grid = (int)std::min<size_t>((words + 255) / 256, 1024);
    hipLaunchKernelGGL(aa::fill_zero_kernel, dim3(grid), dim3(256), 0, stream, static_cast<unsigned*>(ptr), words);
    return hipGetLastError();
}

hipError_t fill_words(void* ptr, size_t words, unsigned value, hipStream_t stream) {
    if (!ptr || words == 0) return hipSuccess;
    const int grid = (int)std::min<size_t>((words + 255) / 256, 1024);
    hipLaunchKernelGGL(aa::fill_word_kernel, dim3(grid), dim3(256), 0, stream, static_cast<unsigned*>(ptr), words,
                       value);
    return hipGetLastError();
}

// Initial values of the visual candidate outputs: no hit (surface / ray -1),
// no contribution (order, distance, gain 0), points at the origin.
hipError_t init_visual(const aa_trace_outputs* o, const aa_trace_params* p, hipStream_t stream) {
    if (p->visual_limit <= 0) return hipSuccess;
    const size_t V = (size_t)p->visual_limit, B = (size_t)(p->num_bounces > 0 ? p->num_bounces : 0);
    hipError_t e;
    if ((e = zero_range(o->d_visual_points, V * B * 3 * 4, stream)) != hipSuccess) return e;
    if ((e = fill_words(o->d_visual_surfaces, V * B, 0xffffffffu, stream)) != hipSuccess) return e;
    if ((e = fill_words(o->d_visual_rays, V, 0xffffffffu, stream)) != hipSuccess) return e;
    if ((e = zero_range(o->d_visual_orders, V * 4, stream)) != hipSuccess) return e;
    if ((e = zero_range(o->d_visual_distances, V * 4, stream)) != hipSuccess) return e;
    return zero_range(o->d_visual_gains, V * 4, stream);
}

}  // namespace

extern "C" int aa_trace(aa_scene_t scene, const aa_trace_params* p, const float* d_sources, const float* d_directions,
                        const float* d_diffuse_bank, const float* d_diffuse_random, const int32_t* d_diffuse_indices,
                        const aa_trace_outputs* o, void* stream_ptr) {
    int rc = aa::validate_trace(scene, p, d_sources, d_directions, d_diffuse_bank, d_diffuse_random,
                                d_diffuse_indices, o);
    if (rc != AA_OK) return rc;
    hipStream_t stream = static_cast<hipStream_t>(stream_ptr);
    AA_HIP(hipSetDevice(scene->device));
    const aa::DeviceProps& props = aa::device_props(scene->device);
    if (!props.valid) return fail(AA_ERR_DEVICE, "cannot query device %d", scene->device);
    const int Ks = p->num_sources, K = p->num_bins, S = scene->num_surfaces;

    // launch shape
    int wg = p->block_size > 0 ? p->block_size : 256;
    if (wg != 64 && wg != 128 && wg != 256 && wg != 512) return fail(AA_ERR_INVALID, "block_size must be 64..512");
    const size_t stack_bytes = aa::private_bytes(wg);  // stacks + connection queues
    const size_t hist_bytes = (size_t)Ks * AA_NUM_BANDS * K * sizeof(float);
    const size_t surf_bytes = (size_t)S * 3 * sizeof(int);
    const size_t lds_cap = props.lds_per_block > 0 ? props.lds_per_block : 65536;
    if (stack_bytes > lds_cap) return fail(AA_ERR_LIMIT, "LDS stacks of %zu bytes exceed %zu", stack_bytes, lds_cap);
    // LDS that keeps the workgroups the register budget allows resident (3
    // waves per SIMD: 12 waves per CU); optional privatisations must fit in
    // it (measured: a histogram that costs occupancy is 1.4-2.3x slower)
    const size_t lds_resident = (props.lds_per_cu > 0 ? props.lds_per_cu : lds_cap) /
                                (size_t)std::max(1, 12 * 64 / wg);
    bool lds_hist;
    if (p->lds_histogram >= 0)
        lds_hist = p->lds_histogram != 0;
    else
        lds_hist = stack_bytes + hist_bytes <= lds_resident;
    if (lds_hist && stack_bytes + hist_bytes > lds_cap) {
        if (p->lds_histogram > 0)
            return fail(AA_ERR_LIMIT, "LDS histogram of %zu bytes does not fit next to the stack", hist_bytes);
        lds_hist = false;
    }
    aa::TraceArgs a;
    a.sv = aa::make_scene_view(scene, p->use_bvh != 0);
    a.p = aa::make_ray_params(p, d_sources, d_directions, d_diffuse_bank, d_diffuse_random, d_diffuse_indices);
    a.out = aa::make_outputs(o, p);
    a.lds_surf = (Ks == 1 && stack_bytes + (lds_hist ? hist_bytes : 0) + surf_bytes <=
                  std::max(lds_resident, stack_bytes + (lds_hist ? hist_bytes : 0))) ? 1 : 0;
    const size_t lds = stack_bytes + (lds_hist ? hist_bytes : 0) + (a.lds_surf ? surf_bytes : 0);
    if (lds > lds_cap) return fail(AA_ERR_LIMIT, "LDS request %zu exceeds %zu bytes per workgroup", lds, lds_cap);
    const int grid_cap = (int)std::max<int64_t>(1, ((int64_t)p->num_rays + wg - 1) / wg);

    // zero the outputs with a kernel (capturable, unlike hipMemsetAsync)
    const aa::OutputExtents ext = aa::output_extents(o, p, S);
    for (int i = 0; i < 8; ++i) AA_HIP(zero_range(ext.ptr[i], ext.bytes[i], stream));
    AA_HIP(init_visual(o, p, stream));

    hipError_t e = hipSuccess;
#define AA_LAUNCH(W)                                                                   \
    case W:                                                                            \
        e = lds_hist ? launch_trace<W, true>(a, grid_cap, p->workgroups_per_cu, props.cu, lds, stream) \
                     : launch_trace<W, false>(a, grid_cap, p->workgroups_per_cu, props.cu, lds, stream); \
        break;
    switch (wg) {
        AA_LAUNCH(64)
        AA_LAUNCH(128)
        AA_LAUNCH(256)
        AA_LAUNCH(512)
    }
#undef AA_LAUNCH
    if (e != hipSuccess) return fail(AA_ERR_DEVICE, "trace launch failed: %s", hipGetErrorString(e));
    return AA_OK;
}

extern "C" int aa_trace_plain(aa_scene_t scene, const aa_trace_params* p, const float* d_sources,
                              const float* d_directions, const float* d_diffuse_bank, const float* d_diffuse_random,
                              const int32_t* d_diffuse_indices, const aa_trace_outputs* o, void* stream_ptr) {
    int rc = aa::validate_trace(scene, p, d_sources, d_directions, d_diffuse_bank, d_diffuse_random,
                                d_diffuse_indices, o);
    if (rc != AA_OK) return rc;
    if (p->visual_limit > 0) return fail(AA_ERR_INVALID, "visual candidates are recorded by aa_trace only");
    hipStream_t stream = static_cast<hipStream_t>(stream_ptr);
    AA_HIP(hipSetDevice(scene->device));
    aa::PlainArgs a;
    a.sv = aa::make_scene_view(scene, p->use_bvh != 0);
    a.nodes = reinterpret_cast<const float4*>(scene->d_plain_nodes);
    a.tris = reinterpret_cast<const float4*>(scene->d_plain_tris);
    a.num_nodes = scene->plain_node_count;
    a.p = aa::make_ray_params(p, d_sources, d_directions, d_diffuse_bank, d_diffuse_random, d_diffuse_indices);
    a.out = aa::make_outputs(o, p);
    const aa::OutputExtents ext = aa::output_extents(o, p, scene->num_surfaces);
    for (int i = 0; i < 8; ++i) AA_HIP(zero_range(ext.ptr[i], ext.bytes[i], stream));
    constexpr int kPlainBlock = 128;  // the reference's block size (steam_rt_cuda.py:136-137)
    const int64_t tasks = (int64_t)p->num_rays * a.p.groups;
    const int grid = (int)((tasks + kPlainBlock - 1) / kPlainBlock);
    hipLaunchKernelGGL(aa::plain_kernel<kPlainBlock>, dim3(grid), dim3(kPlainBlock), 0, stream, a);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(AA_ERR_DEVICE, "plain trace launch failed: %s", hipGetErrorString(e));
    return AA_OK;
}
