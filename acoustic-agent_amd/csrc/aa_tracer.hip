// MI355X (gfx950) reflection tracer: HIP kernels and the C ABI of
// include/aa_tracer.h.
//
// One lane traces one listener ray through up to B bounces
// (acoustic_agent/steam_rt.py:4724-4849): closest hit through a 4-wide BVH
// (top levels staged in LDS, short per-lane stack in LDS), listener/source
// sphere termination, a next-event shadow query towards each source (any-hit
// with early exit, equal to the reference's closest-hit test
// `EPS < t < D - EPS`, steam_rt.py:4788-4792), the BRDF-weighted six-band
// deposit into the time histogram (steam_rt.py:4795-4826), then specular or
// diffuse continuation (steam_rt.py:4827-4848).
//
// Floating-point contract: device code is built with -ffp-contract=off and
// IEEE division/sqrt so every trajectory operation rounds like the binary32
// CPU restatement; only the BVH box tests (pure culling, results do not depend
// on them) use explicit fused multiply-adds.
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
#include "aa_internal.h"

// ---------------------------------------------------------------------------
// error reporting
// ---------------------------------------------------------------------------

namespace {

thread_local std::string g_last_error;

int fail(int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_last_error = buf;
    return code;
}

#define AA_HIP(call)                                                                             \
    do {                                                                                         \
        hipError_t e_ = (call);                                                                  \
        if (e_ != hipSuccess) return fail(AA_ERR_DEVICE, "%s failed: %s", #call, hipGetErrorString(e_)); \
    } while (0)

}  // namespace

// ---------------------------------------------------------------------------
// device constants and helpers
// ---------------------------------------------------------------------------

namespace aa {

constexpr int kBands = AA_NUM_BANDS;
constexpr int kStack = 24;          // per-lane traversal stack entries (uint32)
constexpr float kEps = 1e-6f;       // steam_rt_cuda.py:19
constexpr float kHitOffset = 1e-2f; // steam_rt_cuda.py:20
constexpr float kBvhEps = 1e-5f;    // steam_rt_cuda.py:21
constexpr float kBig = 1e30f;

struct TraceArgs {
    // scene
    const float4* nodes;
    const float4* tris;     // leaf order, 3 float4 per triangle
    const float4* normals;  // original triangle order
    const float* reflection;
    const float* scattering;
    int num_nodes, num_tris, num_surfaces, use_bvh;
    int lds_nodes;
    // rays
    const float* directions;
    const float* bank;
    const float* urand;
    const int* uidx;
    const float4* sources;
    int num_rays, num_bounces, num_bins, num_samples, num_sources, rng_mode, amb;
    unsigned int seed_lo, seed_hi;
    float lx, ly, lz, fx, fy, fz;
    float bin_dur, c, max_path_len, listener_radius, source_radius, dmin, spec_exp, dip_w, dip_p;
    float ray_scale;
    // outputs
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
    int log_rays, log_bounces;
    int* error_flag;
    int lds_surf;  // per-surface counters privatised in LDS
};

__device__ __forceinline__ void philox4x32_10(unsigned c0, unsigned c1, unsigned c2, unsigned c3, unsigned k0,
                                              unsigned k1, unsigned& r0, unsigned& r1) {
#pragma unroll
    for (int round = 0; round < 10; ++round) {
        if (round > 0) {
            k0 += 0x9E3779B9u;
            k1 += 0xBB67AE85u;
        }
        const unsigned hi0 = __umulhi(0xD2511F53u, c0), lo0 = 0xD2511F53u * c0;
        const unsigned hi1 = __umulhi(0xCD9E8D57u, c2), lo1 = 0xCD9E8D57u * c2;
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

__device__ __forceinline__ float safe_inv(float d) {
    // Axis with |d| <= 1e-12: the slab test must only require the origin to
    // lie inside the slab (steam_rt_cuda.py:416-419); a huge signed inverse
    // gives exactly that and never produces a NaN.
    return fabsf(d) > 1e-12f ? 1.0f / d : copysignf(1e30f, d);
}

__device__ __forceinline__ bool tri_test(const float4 a, const float4 b, const float4 c, float ox, float oy, float oz,
                                         float dx, float dy, float dz, float& t_out) {
    // Moller-Trumbore in the operand order of the binary32 CPU restatement.
    const float e1x = b.x, e1y = b.y, e1z = b.z;
    const float e2x = c.x, e2y = c.y, e2z = c.z;
    const float px = dy * e2z - dz * e2y;
    const float py = dz * e2x - dx * e2z;
    const float pz = dx * e2y - dy * e2x;
    const float det = e1x * px + e1y * py + e1z * pz;
    if (!(fabsf(det) > 1e-12f)) return false;
    const float inv = 1.0f / det;
    const float sx = ox - a.x, sy = oy - a.y, sz = oz - a.z;
    const float u = (sx * px + sy * py + sz * pz) * inv;
    if (u < -1e-6f || u > 1.000001f) return false;
    const float qx = sy * e1z - sz * e1y;
    const float qy = sz * e1x - sx * e1z;
    const float qz = sx * e1y - sy * e1x;
    const float v = (dx * qx + dy * qy + dz * qz) * inv;
    if (v < -1e-6f || u + v > 1.000001f) return false;
    const float t = (e2x * qx + e2y * qy + e2z * qz) * inv;
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

__device__ __forceinline__ void load_node(const TraceArgs& a, const float4* s_nodes, int node, float4* v) {
    if (node < a.lds_nodes) {
#pragma unroll
        for (int r = 0; r < 7; ++r) v[r] = s_nodes[node * 8 + r];
    } else {
        const float4* g = a.nodes + (size_t)node * 8;
#pragma unroll
        for (int r = 0; r < 7; ++r) v[r] = g[r];
    }
}

__device__ __forceinline__ void slab4(const float4* v, const Ray& r, float tlim, float* tn, bool* hit) {
    const float oix = r.ox * r.ix, oiy = r.oy * r.iy, oiz = r.oz * r.iz;
    const float lxs[4] = {v[0].x, v[0].y, v[0].z, v[0].w};
    const float hxs[4] = {v[1].x, v[1].y, v[1].z, v[1].w};
    const float lys[4] = {v[2].x, v[2].y, v[2].z, v[2].w};
    const float hys[4] = {v[3].x, v[3].y, v[3].z, v[3].w};
    const float lzs[4] = {v[4].x, v[4].y, v[4].z, v[4].w};
    const float hzs[4] = {v[5].x, v[5].y, v[5].z, v[5].w};
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        const float tx0 = __builtin_fmaf(lxs[c], r.ix, -oix), tx1 = __builtin_fmaf(hxs[c], r.ix, -oix);
        const float ty0 = __builtin_fmaf(lys[c], r.iy, -oiy), ty1 = __builtin_fmaf(hys[c], r.iy, -oiy);
        const float tz0 = __builtin_fmaf(lzs[c], r.iz, -oiz), tz1 = __builtin_fmaf(hzs[c], r.iz, -oiz);
        const float tmin = fmaxf(fmaxf(fminf(tx0, tx1), fminf(ty0, ty1)), fmaxf(fminf(tz0, tz1), 0.0f));
        const float tmax = fminf(fminf(fmaxf(tx0, tx1), fmaxf(ty0, ty1)), fminf(fmaxf(tz0, tz1), tlim));
        tn[c] = tmin;
        hit[c] = tmin <= tmax && tmax > kEps;
    }
}

__device__ __forceinline__ void test_leaf(const TraceArgs& a, int code, const Ray& r, Closest& h, unsigned& tri_cnt) {
    const int v = -(code + 2);
    const int start = v >> 3, count = (v & 7) + 1;
    for (int k = 0; k < count; ++k) {
        const float4* t4 = a.tris + (size_t)(start + k) * 3;
        const float4 ta = t4[0], tb = t4[1], tc = t4[2];
        ++tri_cnt;
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
}

template <int WG>
__device__ void closest_hit(const TraceArgs& a, const float4* s_nodes, unsigned* stk, const Ray& r, Closest& h,
                            unsigned& node_cnt, unsigned& tri_cnt) {
    h.t = kBig;
    h.surf = -1;
    h.tri = -1;
    if (!a.use_bvh) {
        for (int k = 0; k < a.num_tris; ++k) {
            const float4* t4 = a.tris + (size_t)k * 3;
            const float4 ta = t4[0], tb = t4[1], tc = t4[2];
            ++tri_cnt;
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
        return;
    }
    int sp = 0;
    int node = 0;
    while (true) {
        ++node_cnt;
        float4 v[7];
        load_node(a, s_nodes, node, v);
        float tn[4];
        bool hit[4];
        slab4(v, r, h.t + kBvhEps, tn, hit);
        const int ch[4] = {__float_as_int(v[6].x), __float_as_int(v[6].y), __float_as_int(v[6].z),
                           __float_as_int(v[6].w)};
#pragma unroll
        for (int c = 0; c < 4; ++c)
            if (hit[c] && ch[c] <= -2) test_leaf(a, ch[c], r, h, tri_cnt);
        // internal children, nearest first
        float key[4];
        int idx[4];
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            const bool ok = hit[c] && ch[c] >= 0 && tn[c] <= h.t + kBvhEps;
            key[c] = ok ? tn[c] : INFINITY;
            idx[c] = ch[c];
        }
#define AA_CSWAP(i, j)                         \
    if (key[j] < key[i]) {                     \
        const float tk = key[i];               \
        key[i] = key[j];                       \
        key[j] = tk;                           \
        const int ti = idx[i];                 \
        idx[i] = idx[j];                       \
        idx[j] = ti;                           \
    }
        AA_CSWAP(0, 1)
        AA_CSWAP(2, 3)
        AA_CSWAP(0, 2)
        AA_CSWAP(1, 3)
        AA_CSWAP(1, 2)
#undef AA_CSWAP
#pragma unroll
        for (int c = 3; c >= 1; --c) {
            if (key[c] < INFINITY) {
                if (sp < kStack) {
                    stk[sp * WG] = (unsigned)idx[c];
                    ++sp;
                } else {
                    atomicOr(a.error_flag, 1);
                }
            }
        }
        if (key[0] < INFINITY) {
            node = idx[0];
        } else if (sp > 0) {
            --sp;
            node = (int)stk[sp * WG];
        } else {
            break;
        }
    }
}

template <int WG>
__device__ bool any_hit(const TraceArgs& a, const float4* s_nodes, unsigned* stk, const Ray& r, float max_distance,
                        unsigned& node_cnt, unsigned& tri_cnt) {
    const float lo = kEps, hi = max_distance - kEps;
    if (!a.use_bvh) {
        for (int k = 0; k < a.num_tris; ++k) {
            const float4* t4 = a.tris + (size_t)k * 3;
            ++tri_cnt;
            float t;
            if (tri_test(t4[0], t4[1], t4[2], r.ox, r.oy, r.oz, r.dx, r.dy, r.dz, t) && t > lo && t < hi) return true;
        }
        return false;
    }
    int sp = 0;
    int node = 0;
    const float tlim = max_distance + kBvhEps;
    while (true) {
        ++node_cnt;
        float4 v[7];
        load_node(a, s_nodes, node, v);
        float tn[4];
        bool hit[4];
        slab4(v, r, tlim, tn, hit);
        const int ch[4] = {__float_as_int(v[6].x), __float_as_int(v[6].y), __float_as_int(v[6].z),
                           __float_as_int(v[6].w)};
        int next = -1;
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            if (!hit[c]) continue;
            if (ch[c] <= -2) {
                const int vv = -(ch[c] + 2);
                const int start = vv >> 3, count = (vv & 7) + 1;
                for (int k = 0; k < count; ++k) {
                    const float4* t4 = a.tris + (size_t)(start + k) * 3;
                    ++tri_cnt;
                    float t;
                    if (tri_test(t4[0], t4[1], t4[2], r.ox, r.oy, r.oz, r.dx, r.dy, r.dz, t) && t > lo && t < hi)
                        return true;
                }
            } else if (ch[c] >= 0) {
                if (next < 0) {
                    next = ch[c];
                } else if (sp < kStack) {
                    stk[sp * WG] = (unsigned)ch[c];
                    ++sp;
                } else {
                    atomicOr(a.error_flag, 1);
                }
            }
        }
        if (next >= 0) {
            node = next;
        } else if (sp > 0) {
            --sp;
            node = (int)stk[sp * WG];
        } else {
            return false;
        }
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

// Zero a device range with a kernel (no hipMemsetAsync, so the call stays
// capturable into a HIP graph).
__global__ void fill_zero_kernel(unsigned* p, size_t words) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < words; i += (size_t)gridDim.x * blockDim.x)
        p[i] = 0u;
}

template <int WG, bool LDS_HIST>
__global__ __launch_bounds__(WG) void trace_kernel(const TraceArgs a) {
    extern __shared__ __align__(16) unsigned char smem[];
    unsigned* s_stack = reinterpret_cast<unsigned*>(smem);
    float4* s_nodes = reinterpret_cast<float4*>(s_stack + kStack * WG);
    float* s_hist = reinterpret_cast<float*>(s_nodes + (size_t)a.lds_nodes * 8);
    const int hist_words = LDS_HIST ? a.num_sources * kBands * a.num_bins : 0;
    int* s_hits = reinterpret_cast<int*>(s_hist + hist_words);
    int* s_contrib = s_hits + (a.lds_surf ? a.num_surfaces : 0);
    float* s_senergy = reinterpret_cast<float*>(s_contrib + (a.lds_surf ? a.num_surfaces : 0));

    const int tid = threadIdx.x;
    for (int i = tid; i < a.lds_nodes * 8; i += WG) s_nodes[i] = a.nodes[i];
    if (LDS_HIST)
        for (int i = tid; i < hist_words; i += WG) s_hist[i] = 0.0f;
    if (a.lds_surf) {
        for (int i = tid; i < a.num_surfaces; i += WG) {
            s_hits[i] = 0;
            s_contrib[i] = 0;
            s_senergy[i] = 0.0f;
        }
    }
    __syncthreads();

    unsigned* stk = s_stack + tid;
    unsigned nb = 0, ns = 0, tb = 0, ts = 0, q_b = 0, q_s = 0, contrib_total = 0, escaped = 0;
    int max_iter_full = 0;  // max bounce iterations of rays that ended with every source alive
    int active_full = 0;
    const int K = a.num_bins;
    const int Ks = a.num_sources;
    const unsigned long long full_mask = Ks >= 64 ? ~0ull : ((1ull << Ks) - 1ull);

    for (int ray = blockIdx.x * WG + tid; ray < a.num_rays; ray += gridDim.x * WG) {
        Ray r;
        r.ox = a.lx;
        r.oy = a.ly;
        r.oz = a.lz;
        if (a.rng_mode == AA_RNG_ARRAYS) {
            r.dx = a.directions[(size_t)ray * 3];
            r.dy = a.directions[(size_t)ray * 3 + 1];
            r.dz = a.directions[(size_t)ray * 3 + 2];
        } else {
            unsigned w0, w1;
            philox4x32_10(0xFFFFFFFFu, (unsigned)ray, 0u, 0u, a.seed_lo, a.seed_hi, w0, w1);
            const double u = (double)(w0 >> 8) * (1.0 / 16777216.0);
            const double jitter = -0.02 + 0.04 * u;
            const double n = (double)(a.num_rays > 0 ? a.num_rays : 1);
            const double i = (double)ray + 0.5;
            const double z = 1.0 - 2.0 * i / n;
            double rr = 1.0 - z * z;
            rr = rr < 0.0 ? 0.0 : (rr > 1.0 ? 1.0 : rr);
            const double rad = sqrt(rr);
            const double theta = i * (M_PI * (3.0 - 2.23606797749978969640917366873127623544061835961152572)) + jitter;
            r.dx = (float)(cos(theta) * rad);
            r.dy = (float)(sin(theta) * rad);
            r.dz = (float)z;
        }
        const float d0x = r.dx, d0y = r.dy, d0z = r.dz;
        float A[kBands];
#pragma unroll
        for (int b = 0; b < kBands; ++b) A[b] = 1.0f;
        float accum = 0.0f;
        unsigned long long mask = full_mask;
        bool alive = true;
        int iter = 0;
        const bool logging = a.log_tri != nullptr && ray < a.log_rays;
        for (int bounce = 0; bounce < a.num_bounces; ++bounce) {
            iter = bounce + 1;
            r.ix = safe_inv(r.dx);
            r.iy = safe_inv(r.dy);
            r.iz = safe_inv(r.dz);
            Closest h;
            ++q_b;
            closest_hit<WG>(a, s_nodes, stk, r, h, nb, tb);
            if (logging && bounce < a.log_bounces) {
                a.log_tri[(size_t)ray * a.log_bounces + bounce] = h.tri;
                a.log_t[(size_t)ray * a.log_bounces + bounce] = h.t;
            }
            if (h.surf < 0) ++escaped;
            if (h.surf < 0 || h.t <= a.listener_radius || accum > a.max_path_len) {
                alive = false;
                break;
            }
            if (bounce > 0) {
                if (sphere_before(r.ox, r.oy, r.oz, r.dx, r.dy, r.dz, a.lx, a.ly, a.lz, a.listener_radius, h.t)) {
                    alive = false;
                    break;
                }
                unsigned long long m = mask;
                while (m) {
                    const int k = __builtin_ctzll(m);
                    m &= m - 1ull;
                    const float4 S = a.sources[k];
                    const float sdist = sqrtf((a.lx - S.x) * (a.lx - S.x) + (a.ly - S.y) * (a.ly - S.y) +
                                              (a.lz - S.z) * (a.lz - S.z));
                    if (sdist > a.source_radius &&
                        sphere_before(r.ox, r.oy, r.oz, r.dx, r.dy, r.dz, S.x, S.y, S.z, a.source_radius, h.t)) {
                        mask &= ~(1ull << k);
                        atomicMax(&a.actual_bounces[k], bounce + 1);
                    }
                }
                if (mask == 0ull) {
                    alive = false;
                    break;
                }
            }
            float4 nn = a.normals[h.tri];
            float nx = nn.x, ny = nn.y, nz = nn.z;
            if (nx * r.dx + ny * r.dy + nz * r.dz > 0.0f) {
                nx = -nx;
                ny = -ny;
                nz = -nz;
            }
            const int surf = h.surf;
            const float t = h.t;
            const float hx = r.ox + t * r.dx + kHitOffset * nx;
            const float hy = r.oy + t * r.dy + kHitOffset * ny;
            const float hz = r.oz + t * r.dz + kHitOffset * nz;
            const float scat = a.scattering[surf];
            float refl[kBands];
#pragma unroll
            for (int b = 0; b < kBands; ++b) refl[b] = a.reflection[surf * kBands + b];
            {
                unsigned long long m = mask;
                while (m) {
                    const int k = __builtin_ctzll(m);
                    m &= m - 1ull;
                    if (a.lds_surf && Ks == 1)
                        atomicAdd(&s_hits[surf], 1);
                    else
                        atomicAdd(&a.hit_counts[(size_t)k * a.num_surfaces + surf], 1);
                    const float4 S = a.sources[k];
                    const float tsx = S.x - hx, tsy = S.y - hy, tsz = S.z - hz;
                    const float D = sqrtf(tsx * tsx + tsy * tsy + tsz * tsz);
                    if (!(D > a.dmin)) continue;
                    const float sdx = tsx / D, sdy = tsy / D, sdz = tsz / D;
                    const float facing = nx * sdx + ny * sdy + nz * sdz;
                    if (!(facing > 0.0f)) continue;
                    Ray sr;
                    sr.ox = hx;
                    sr.oy = hy;
                    sr.oz = hz;
                    sr.dx = sdx;
                    sr.dy = sdy;
                    sr.dz = sdz;
                    sr.ix = safe_inv(sdx);
                    sr.iy = safe_inv(sdy);
                    sr.iz = safe_inv(sdz);
                    ++q_s;
                    if (any_hit<WG>(a, s_nodes, stk, sr, D, ns, ts)) continue;
                    const float diffuse = 0.31830988618379067f * scat * facing;
                    float hvx = sdx - r.dx, hvy = sdy - r.dy, hvz = sdz - r.dz;
                    normalize3(hvx, hvy, hvz);
                    float cos_half = hvx * nx + hvy * ny + hvz * nz;
                    cos_half = cos_half < 0.0f ? 0.0f : cos_half;
                    const float specular =
                        (a.spec_exp + 2.0f) * 0.039788735772973836f * (1.0f - scat) * powf(cos_half, a.spec_exp);
                    const float dm = D > a.dmin ? D : a.dmin;
                    const float distance_term = 0.07957747154594767f / (dm * dm);
                    const float source_cosine = -(a.fx * sdx + a.fy * sdy + a.fz * sdz);
                    const float source_gain = powf(fabsf((1.0f - a.dip_w) + a.dip_w * source_cosine), a.dip_p);
                    const float rel_delay = (accum + t + D) / a.c - S.w;
                    const float q = rel_delay / a.bin_dur;
                    if (!(q >= 0.0f && q < (float)K)) continue;
                    const int bin = (int)floorf(q);
                    const float base = a.ray_scale * source_gain * distance_term * (diffuse + specular);
                    float esum = 0.0f;
#pragma unroll
                    for (int b = 0; b < kBands; ++b) {
                        const float e = base * refl[b] * A[b];
                        const size_t hi = ((size_t)k * kBands + b) * K + bin;
                        if (LDS_HIST)
                            atomicAdd(&s_hist[hi], e);
                        else
                            atomicAdd(&a.echogram[hi], e);
                        if (a.amb) {
                            float* am = a.ambisonic + (((size_t)k * kBands + b) * 4) * K + bin;
                            atomicAdd(am, e);
                            atomicAdd(am + K, e * d0x);
                            atomicAdd(am + 2 * K, e * d0y);
                            atomicAdd(am + 3 * K, e * d0z);
                        }
                        esum += e;
                    }
                    ++contrib_total;
                    if (a.lds_surf && Ks == 1) {
                        atomicAdd(&s_contrib[surf], 1);
                        atomicAdd(&s_senergy[surf], esum);
                    } else {
                        atomicAdd(&a.contrib_counts[(size_t)k * a.num_surfaces + surf], 1);
                        atomicAdd(&a.surface_energy[(size_t)k * a.num_surfaces + surf], esum);
                    }
                }
            }
#pragma unroll
            for (int b = 0; b < kBands; ++b) A[b] *= refl[b];
            accum += t;
            r.ox = hx;
            r.oy = hy;
            r.oz = hz;
            float u;
            int j;
            if (a.rng_mode == AA_RNG_ARRAYS) {
                u = a.urand[(size_t)bounce * a.num_rays + ray];
                j = a.uidx[(size_t)bounce * a.num_rays + ray];
            } else {
                unsigned w0, w1;
                philox4x32_10((unsigned)bounce, (unsigned)ray, 0u, 0u, a.seed_lo, a.seed_hi, w0, w1);
                u = (float)(w0 >> 8) * (1.0f / 16777216.0f);
                j = (int)(((unsigned long long)w1 * (unsigned long long)(unsigned)a.num_samples) >> 32);
            }
            if (u < scat) {
                const float sx = a.bank[j * 3], sy = a.bank[j * 3 + 1], sz = a.bank[j * 3 + 2];
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
                float ndx = sx * tx + sy * bx + sz * nx;
                float ndy = sx * ty + sy * by + sz * ny;
                float ndz = sx * tz + sy * bz + sz * nz;
                normalize3(ndx, ndy, ndz);
                r.dx = ndx;
                r.dy = ndy;
                r.dz = ndz;
            } else {
                const float dot = r.dx * nx + r.dy * ny + r.dz * nz;
                float ndx = r.dx - 2.0f * dot * nx;
                float ndy = r.dy - 2.0f * dot * ny;
                float ndz = r.dz - 2.0f * dot * nz;
                normalize3(ndx, ndy, ndz);
                r.dx = ndx;
                r.dy = ndy;
                r.dz = ndz;
            }
            if (accum > a.max_path_len) {
                alive = false;
                break;
            }
        }
        // per-source bookkeeping of rays that ended with some sources still alive
        if (mask == full_mask) {
            max_iter_full = max(max_iter_full, iter);
            if (alive) ++active_full;
        } else {
            unsigned long long m = mask;
            while (m) {
                const int k = __builtin_ctzll(m);
                m &= m - 1ull;
                atomicMax(&a.actual_bounces[k], iter);
                if (alive) atomicAdd(&a.active_count[k], 1);
            }
        }
    }

    // wavefront reductions of the per-lane counters, then one atomic per wave
    unsigned long long vals[8] = {q_b, q_s, nb, ns, tb, ts, contrib_total, escaped};
#pragma unroll
    for (int c = 0; c < 8; ++c) {
        unsigned long long v = vals[c];
        for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
        vals[c] = v;
    }
    int mi = max_iter_full, ac = active_full;
    for (int off = 32; off > 0; off >>= 1) {
        mi = max(mi, __shfl_xor(mi, off));
        ac += __shfl_xor(ac, off);
    }
    if ((tid & 63) == 0) {
#pragma unroll
        for (int c = 0; c < 8; ++c)
            if (vals[c]) atomicAdd(&a.counters[c], vals[c]);
        for (int k = 0; k < Ks; ++k) {
            if (mi) atomicMax(&a.actual_bounces[k], mi);
            if (ac) atomicAdd(&a.active_count[k], ac);
        }
    }
    __syncthreads();
    if (LDS_HIST) {
        for (int i = tid; i < hist_words; i += WG) {
            const float v = s_hist[i];
            if (v != 0.0f) atomicAdd(&a.echogram[i], v);
        }
    }
    if (a.lds_surf) {
        for (int i = tid; i < a.num_surfaces; i += WG) {
            if (s_hits[i]) atomicAdd(&a.hit_counts[i], s_hits[i]);
            if (s_contrib[i]) atomicAdd(&a.contrib_counts[i], s_contrib[i]);
            if (s_senergy[i] != 0.0f) atomicAdd(&a.surface_energy[i], s_senergy[i]);
        }
    }
}

}  // namespace aa

// ---------------------------------------------------------------------------
// scene handle
// ---------------------------------------------------------------------------

struct aa_scene_s {
    int32_t device = 0;
    int32_t num_triangles = 0, num_surfaces = 0;
    int32_t leaf_size = 4;
    int32_t canonical_nodes = 0, canonical_depth = 0;
    aa::WideBVH wide;
    float4* d_nodes = nullptr;
    float4* d_tris = nullptr;
    float4* d_normals = nullptr;
    float* d_reflection = nullptr;
    float* d_scattering = nullptr;
    int* d_error = nullptr;
    uint64_t device_bytes = 0;
    uint64_t content_hash = 0;
};

namespace {

uint64_t fnv1a(uint64_t h, const void* data, size_t bytes) {
    const unsigned char* p = static_cast<const unsigned char*>(data);
    for (size_t i = 0; i < bytes; ++i) {
        h ^= p[i];
        h *= 1099511628211ull;
    }
    return h;
}

void free_scene(aa_scene_s* s) {
    if (!s) return;
    (void)hipSetDevice(s->device);
    (void)hipFree(s->d_nodes);
    (void)hipFree(s->d_tris);
    (void)hipFree(s->d_normals);
    (void)hipFree(s->d_reflection);
    (void)hipFree(s->d_scattering);
    (void)hipFree(s->d_error);
    delete s;
}

struct DeviceProps {
    int cu = 0;
    size_t lds_per_block = 0;
    bool valid = false;
};

std::mutex g_props_mutex;
std::vector<DeviceProps> g_props;

const DeviceProps& device_props(int device) {
    std::lock_guard<std::mutex> lock(g_props_mutex);
    if ((int)g_props.size() <= device) g_props.resize((size_t)device + 1);
    DeviceProps& p = g_props[(size_t)device];
    if (!p.valid) {
        hipDeviceProp_t prop;
        if (hipGetDeviceProperties(&prop, device) == hipSuccess) {
            p.cu = prop.multiProcessorCount;
            p.lds_per_block = std::max(prop.sharedMemPerBlock, prop.sharedMemPerBlockOptin);
            p.valid = true;
        }
    }
    return p;
}

}  // namespace

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

int aa_build_bvh(const double* bounds_min, const double* bounds_max, int64_t n, int64_t leaf_size, double* node_min,
                 double* node_max, int64_t* start, int64_t* count, int64_t* escape, int64_t* primitives,
                 int64_t* stats) {
    if (!bounds_min || !bounds_max || !node_min || !node_max || !start || !count || !escape || !primitives || !stats)
        return fail(AA_ERR_INVALID, "aa_build_bvh: NULL argument");
    aa::CanonicalBVH bvh;
    std::string err = aa::build_canonical_bvh(bounds_min, bounds_max, n, leaf_size, bvh);
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
    aa::CanonicalBVH bin;
    std::string err = aa::build_canonical_bvh(lo.data(), hi.data(), num_triangles, leaf_size, bin);
    if (!err.empty()) return fail(AA_ERR_INVALID, "%s", err.c_str());

    aa_scene_s* s = new aa_scene_s();
    s->device = device;
    s->num_triangles = num_triangles;
    s->num_surfaces = num_surfaces;
    s->leaf_size = leaf_size;
    s->canonical_nodes = (int32_t)bin.node_count();
    s->canonical_depth = (int32_t)bin.max_depth;
    aa::collapse_to_wide(bin, s->wide);
    if (3 * (s->wide.depth - 1) + 1 > aa::kStack) {
        free_scene(s);
        return fail(AA_ERR_LIMIT, "BVH depth %d exceeds the traversal stack", s->wide.depth);
    }

    // triangles in leaf order: (v0, tri id) (e1, surface) (e2, 0)
    std::vector<float> tris((size_t)num_triangles * 12);
    for (int32_t k = 0; k < num_triangles; ++k) {
        const int64_t tri = bin.prims[(size_t)k];
        const float* v = tri_vertices + (size_t)tri * 9;
        float* o = tris.data() + (size_t)k * 12;
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
    std::vector<float> normals((size_t)num_triangles * 4);
    for (int32_t k = 0; k < num_triangles; ++k) {
        normals[(size_t)k * 4 + 0] = tri_normals[(size_t)k * 3 + 0];
        normals[(size_t)k * 4 + 1] = tri_normals[(size_t)k * 3 + 1];
        normals[(size_t)k * 4 + 2] = tri_normals[(size_t)k * 3 + 2];
        normals[(size_t)k * 4 + 3] = 0.0f;
    }
    uint64_t h = 1469598103934665603ull;
    h = fnv1a(h, tri_vertices, (size_t)num_triangles * 9 * sizeof(float));
    h = fnv1a(h, tri_surface, (size_t)num_triangles * sizeof(int32_t));
    h = fnv1a(h, tri_normals, (size_t)num_triangles * 3 * sizeof(float));
    h = fnv1a(h, reflection, (size_t)num_surfaces * 6 * sizeof(float));
    h = fnv1a(h, scattering, (size_t)num_surfaces * sizeof(float));
    h = fnv1a(h, &leaf_size, sizeof(leaf_size));
    s->content_hash = h;

    hipError_t e = hipSetDevice(device);
    const size_t node_bytes = s->wide.nodes.size() * sizeof(aa::WideNode);
    const size_t tri_bytes = tris.size() * sizeof(float);
    const size_t nrm_bytes = normals.size() * sizeof(float);
    const size_t refl_bytes = (size_t)num_surfaces * 6 * sizeof(float);
    const size_t scat_bytes = (size_t)num_surfaces * sizeof(float);
    if (e == hipSuccess) e = hipMalloc((void**)&s->d_nodes, node_bytes);
    if (e == hipSuccess) e = hipMalloc((void**)&s->d_tris, tri_bytes);
    if (e == hipSuccess) e = hipMalloc((void**)&s->d_normals, nrm_bytes);
    if (e == hipSuccess) e = hipMalloc((void**)&s->d_reflection, refl_bytes);
    if (e == hipSuccess) e = hipMalloc((void**)&s->d_scattering, scat_bytes);
    if (e == hipSuccess) e = hipMalloc((void**)&s->d_error, sizeof(int));
    if (e == hipSuccess) e = hipMemcpy(s->d_nodes, s->wide.nodes.data(), node_bytes, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(s->d_tris, tris.data(), tri_bytes, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(s->d_normals, normals.data(), nrm_bytes, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(s->d_reflection, reflection, refl_bytes, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(s->d_scattering, scattering, scat_bytes, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemset(s->d_error, 0, sizeof(int));
    if (e != hipSuccess) {
        free_scene(s);
        return fail(AA_ERR_MEMORY, "scene upload failed: %s", hipGetErrorString(e));
    }
    s->device_bytes = node_bytes + tri_bytes + nrm_bytes + refl_bytes + scat_bytes + sizeof(int);
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

}  // extern "C"

namespace {

template <int WG, bool LDS>
hipError_t launch_trace(const aa::TraceArgs& args, int grid, size_t lds, hipStream_t stream) {
    auto kern = aa::trace_kernel<WG, LDS>;
    if (lds > 64 * 1024) {
        hipError_t e = hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL(kern, dim3(grid), dim3(WG), lds, stream, args);
    return hipGetLastError();
}

}  // namespace

extern "C" int aa_trace(aa_scene_t scene, const aa_trace_params* p, const float* d_sources, const float* d_directions,
                        const float* d_diffuse_bank, const float* d_diffuse_random, const int32_t* d_diffuse_indices,
                        const aa_trace_outputs* o, void* stream_ptr) {
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

    hipStream_t stream = static_cast<hipStream_t>(stream_ptr);
    AA_HIP(hipSetDevice(scene->device));
    const DeviceProps& props = device_props(scene->device);
    if (!props.valid) return fail(AA_ERR_DEVICE, "cannot query device %d", scene->device);

    const int Ks = p->num_sources, K = p->num_bins, S = scene->num_surfaces;
    // zero the outputs with a kernel (capturable, unlike hipMemsetAsync)
    auto zero = [&](void* ptr, size_t bytes) {
        if (!ptr || bytes == 0) return;
        const size_t words = bytes / 4;
        const int grid = (int)std::min<size_t>((words + 255) / 256, 1024);
        hipLaunchKernelGGL(aa::fill_zero_kernel, dim3(grid), dim3(256), 0, stream, static_cast<unsigned*>(ptr), words);
    };
    zero(o->d_echogram, (size_t)Ks * AA_NUM_BANDS * K * sizeof(float));
    if (p->render_ambisonics) zero(o->d_ambisonic, (size_t)Ks * AA_NUM_BANDS * 4 * K * sizeof(float));
    zero(o->d_hit_counts, (size_t)Ks * S * sizeof(int32_t));
    zero(o->d_contrib_counts, (size_t)Ks * S * sizeof(int32_t));
    zero(o->d_surface_energy, (size_t)Ks * S * sizeof(float));
    zero(o->d_actual_bounces, (size_t)Ks * sizeof(int32_t));
    zero(o->d_active_count, (size_t)Ks * sizeof(int32_t));
    zero(o->d_counters, AA_CNT_COUNT * sizeof(int64_t));
    AA_HIP(hipGetLastError());

    aa::TraceArgs a;
    std::memset(&a, 0, sizeof(a));
    a.nodes = scene->d_nodes;
    a.tris = scene->d_tris;
    a.normals = scene->d_normals;
    a.reflection = scene->d_reflection;
    a.scattering = scene->d_scattering;
    a.num_nodes = (int)scene->wide.nodes.size();
    a.num_tris = scene->num_triangles;
    a.num_surfaces = S;
    a.use_bvh = p->use_bvh ? 1 : 0;
    a.directions = d_directions;
    a.bank = d_diffuse_bank;
    a.urand = d_diffuse_random;
    a.uidx = d_diffuse_indices;
    a.sources = reinterpret_cast<const float4*>(d_sources);
    a.num_rays = p->num_rays;
    a.num_bounces = p->num_bounces;
    a.num_bins = K;
    a.num_samples = p->num_diffuse_samples;
    a.num_sources = Ks;
    a.rng_mode = p->rng_mode;
    a.amb = p->render_ambisonics ? 1 : 0;
    a.seed_lo = (unsigned)(p->seed & 0xffffffffull);
    a.seed_hi = (unsigned)(p->seed >> 32);
    a.lx = p->listener[0];
    a.ly = p->listener[1];
    a.lz = p->listener[2];
    a.fx = p->source_forward[0];
    a.fy = p->source_forward[1];
    a.fz = p->source_forward[2];
    a.bin_dur = p->bin_duration;
    a.c = p->speed_of_sound;
    a.max_path_len = p->max_path_length;
    a.listener_radius = p->listener_radius;
    a.source_radius = p->source_radius;
    a.dmin = p->irradiance_min_distance;
    a.spec_exp = p->specular_exponent;
    a.dip_w = p->dipole_weight;
    a.dip_p = p->dipole_power;
    a.ray_scale = 12.566370614359172f / (float)p->num_rays;
    a.echogram = o->d_echogram;
    a.ambisonic = o->d_ambisonic;
    a.hit_counts = o->d_hit_counts;
    a.contrib_counts = o->d_contrib_counts;
    a.surface_energy = o->d_surface_energy;
    a.actual_bounces = o->d_actual_bounces;
    a.active_count = o->d_active_count;
    a.counters = reinterpret_cast<unsigned long long*>(o->d_counters);
    a.log_rays = (p->log_rays > 0 && p->log_bounces > 0) ? p->log_rays : 0;
    a.log_bounces = a.log_rays ? p->log_bounces : 0;
    a.log_tri = a.log_rays ? o->d_log_triangle : nullptr;
    a.log_t = a.log_rays ? o->d_log_t : nullptr;
    a.error_flag = scene->d_error;

    // launch shape
    int wg = p->block_size > 0 ? p->block_size : 256;
    if (wg != 64 && wg != 128 && wg != 256 && wg != 512) return fail(AA_ERR_INVALID, "block_size must be 64..512");
    int lds_nodes = p->lds_nodes >= 0 ? p->lds_nodes : 0;
    if (p->lds_nodes < 0) {
        // default: the levels that fit in 4 KB
        for (size_t l = 0; l < scene->wide.level_end.size(); ++l)
            if ((size_t)scene->wide.level_end[l] * sizeof(aa::WideNode) <= 4096) lds_nodes = scene->wide.level_end[l];
    }
    lds_nodes = std::min(lds_nodes, a.num_nodes);
    a.lds_nodes = p->use_bvh ? lds_nodes : 0;
    const size_t stack_bytes = (size_t)aa::kStack * wg * sizeof(unsigned);
    const size_t node_bytes = (size_t)a.lds_nodes * sizeof(aa::WideNode);
    const size_t hist_bytes = (size_t)Ks * AA_NUM_BANDS * K * sizeof(float);
    const size_t surf_bytes = (size_t)S * 3 * sizeof(int);
    const size_t lds_cap = props.lds_per_block > 0 ? props.lds_per_block : 65536;
    bool lds_hist;
    if (p->lds_histogram >= 0)
        lds_hist = p->lds_histogram != 0;
    else
        lds_hist = stack_bytes + node_bytes + hist_bytes <= 80 * 1024;
    if (lds_hist && stack_bytes + node_bytes + hist_bytes > lds_cap) {
        if (p->lds_histogram > 0)
            return fail(AA_ERR_LIMIT, "LDS histogram of %zu bytes does not fit next to the stack", hist_bytes);
        lds_hist = false;
    }
    a.lds_surf = (Ks == 1 && stack_bytes + node_bytes + (lds_hist ? hist_bytes : 0) + surf_bytes <= lds_cap) ? 1 : 0;
    const size_t lds = stack_bytes + node_bytes + (lds_hist ? hist_bytes : 0) + (a.lds_surf ? surf_bytes : 0);
    if (lds > lds_cap) return fail(AA_ERR_LIMIT, "LDS request %zu exceeds %zu bytes per workgroup", lds, lds_cap);
    const int grid = (int)std::max<int64_t>(1, ((int64_t)p->num_rays + wg - 1) / wg);
    if (a.log_rays > p->num_rays) a.log_rays = p->num_rays;

    hipError_t e = hipSuccess;
#define AA_LAUNCH(W)                                                           \
    case W:                                                                    \
        e = lds_hist ? launch_trace<W, true>(a, grid, lds, stream)             \
                     : launch_trace<W, false>(a, grid, lds, stream);           \
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
