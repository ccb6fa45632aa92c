// Persistent batched tracer: many traces (scenes x listener channels x
// keyframes, uneven triangle and ray counts) in one launch.
//
// The reference runs one trace per call and fans calls out over a thread pool
// (acoustic_agent/batch.py:139-184, api.py:477-639); here a batch is a table
// of trace descriptors in device memory and a queue of work items, each item
// a packet of rays of one trace.  A grid sized to fill the GPU pulls items
// with one atomic per item until the queue is empty (every workgroup reaches
// that exit), re-stages the top BVH nodes in LDS when it moves to another
// scene, and traces the packet with the per-ray routine of aa_device.h (so a
// trace gives the same result alone, in a batch, or through aa_trace, up to
// the order of float atomic sums).
//
// A submission is: copy the descriptor table from pinned host memory, zero
// every trace's outputs with a kernel, run the persistent kernel.  With
// use_graph the three steps are captured once into a hipGraph (no memset
// nodes) and replayed; later submissions only rewrite the pinned table.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <map>
#include <mutex>
#include <tuple>
#include <vector>

#include "../../include/aa_tracer.h"
#include "aa_device.h"
#include "aa_internal.h"

#define fail aa::set_error

#define AA_HIP(call)                                                                                          \
    do {                                                                                                      \
        hipError_t e_ = (call);                                                                               \
        if (e_ != hipSuccess) return fail(AA_ERR_DEVICE, "%s failed: %s", #call, hipGetErrorString(e_));      \
    } while (0)

namespace aa {

struct BatchTrace {
    SceneView sv;
    RayParams p;
    Outputs out;
    int* error_flag;
    int first_packet;  // prefix sum of packets over the table
    int num_packets;
    int lds_surf;      // single source and S small enough for LDS tallies
    int hist_words;    // Ks * 6 * K
    unsigned* zero_ptr[8];
    unsigned zero_words[8];
};

struct BatchHeader {
    int num_traces;
    int total_packets;
    int packet_rays;
    unsigned queue_head;
};

__global__ void batch_zero_kernel(const BatchHeader* hdr, const BatchTrace* traces) {
    const int t = blockIdx.x;
    if (t >= hdr->num_traces) return;
    const BatchTrace& bt = traces[t];
    for (int r = 0; r < 8; ++r) {
        unsigned* p = bt.zero_ptr[r];
        const unsigned n = bt.zero_words[r];
        if (!p) continue;
        for (unsigned i = threadIdx.x; i < n; i += blockDim.x) p[i] = 0u;
    }
}

template <int WG, bool LDS_HIST>
__global__ __launch_bounds__(WG) void batch_kernel(BatchHeader* hdr, const BatchTrace* traces, int lds_nodes_cap,
                                                   int hist_cap, int surf_cap) {
    extern __shared__ __align__(16) unsigned char smem[];
    unsigned short* s_stack = reinterpret_cast<unsigned short*>(smem);
    float4* s_nodes = reinterpret_cast<float4*>(smem + kStackBytes<WG>());
    float* s_hist = reinterpret_cast<float*>(s_nodes + (size_t)lds_nodes_cap * 8);
    int* s_hits = reinterpret_cast<int*>(s_hist + (LDS_HIST ? hist_cap : 0));
    int* s_contrib = s_hits + surf_cap;
    float* s_senergy = reinterpret_cast<float*>(s_contrib + surf_cap);
    __shared__ int s_item;
    __shared__ int s_next_ray;

    const int tid = threadIdx.x;
    const int num_traces = hdr->num_traces;
    const int total = hdr->total_packets;
    const int packet_rays = hdr->packet_rays;
    int cur = -1;
    while (true) {
        if (tid == 0) s_item = (int)atomicAdd(&hdr->queue_head, 1u);
        __syncthreads();
        const int item = s_item;
        if (item >= total) break;  // uniform: every workgroup exits once the queue is drained
        // trace of this item: last t with first_packet <= item
        int lo = 0, hi = num_traces - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (traces[mid].first_packet <= item)
                lo = mid;
            else
                hi = mid - 1;
        }
        const int t = lo;
        const BatchTrace& bt = traces[t];
        SceneView sv = bt.sv;
        sv.lds_nodes = min(lds_nodes_cap, sv.use_bvh ? sv.num_nodes : 0);
        if (t != cur) {
            for (int i = tid; i < sv.lds_nodes * 8; i += WG) s_nodes[i] = sv.nodes[i];
            cur = t;
        }
        const bool lds_hist = LDS_HIST && bt.hist_words <= hist_cap;
        if (lds_hist)
            for (int i = tid; i < bt.hist_words; i += WG) s_hist[i] = 0.0f;
        if (bt.lds_surf) {
            for (int i = tid; i < sv.num_surfaces; i += WG) {
                s_hits[i] = 0;
                s_contrib[i] = 0;
                s_senergy[i] = 0.0f;
            }
        }
        const int begin = (item - bt.first_packet) * packet_rays;
        const int end = min(begin + packet_rays, bt.p.num_rays);
        if (tid == 0) s_next_ray = begin;
        __syncthreads();
        Lds lds;
        lds.stack = s_stack;
        lds.nodes = s_nodes;
        lds.hist = s_hist;
        lds.hits = bt.lds_surf ? s_hits : nullptr;
        lds.contrib = bt.lds_surf ? s_contrib : nullptr;
        lds.senergy = bt.lds_surf ? s_senergy : nullptr;
        LaneStats st;
        while (true) {
            const int ray = atomicAdd(&s_next_ray, 1);
            if (ray >= end) break;
            if (lds_hist)
                trace_one_ray<WG, true>(sv, bt.p, bt.out, lds, ray, st, bt.error_flag);
            else
                trace_one_ray<WG, false>(sv, bt.p, bt.out, lds, ray, st, bt.error_flag);
        }
        flush_stats(st, bt.out, bt.p.num_sources);
        __syncthreads();
        if (lds_hist) {
            for (int i = tid; i < bt.hist_words; i += WG) {
                const float v = s_hist[i];
                if (v != 0.0f) atomicAdd(&bt.out.echogram[i], v);
            }
        }
        if (bt.lds_surf) {
            for (int i = tid; i < sv.num_surfaces; i += WG) {
                if (s_hits[i]) atomicAdd(&bt.out.hit_counts[i], s_hits[i]);
                if (s_contrib[i]) atomicAdd(&bt.out.contrib_counts[i], s_contrib[i]);
                if (s_senergy[i] != 0.0f) atomicAdd(&bt.out.surface_energy[i], s_senergy[i]);
            }
        }
        __syncthreads();
    }
}

}  // namespace aa

struct aa_batch_s {
    int32_t device = 0;
    int32_t max_traces = 0;
    void* h_table = nullptr;  // pinned: header followed by traces
    void* d_table = nullptr;
    size_t table_bytes = 0;
    hipEvent_t copied = nullptr;  // the last submission's table copy has finished
    bool pending = false;
    std::map<std::tuple<int, int, int, int, int, int, hipStream_t>, hipGraphExec_t> graphs;
    std::vector<hipGraph_t> graph_objs;
    std::mutex mutex;
};

namespace {

size_t table_size(int max_traces) {
    const size_t hdr = (sizeof(aa::BatchHeader) + 255) & ~size_t(255);
    return hdr + (size_t)max_traces * sizeof(aa::BatchTrace);
}

template <int WG, bool LDS>
hipError_t launch_batch(aa::BatchHeader* hdr, const aa::BatchTrace* traces, int grid, size_t lds, int lds_nodes,
                        int hist_cap, int surf_cap, hipStream_t stream) {
    auto kern = aa::batch_kernel<WG, LDS>;
    if (lds > 64 * 1024) {
        hipError_t e = hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL(kern, dim3(grid), dim3(WG), lds, stream, hdr, traces, lds_nodes, hist_cap, surf_cap);
    return hipGetLastError();
}

}  // namespace

extern "C" {

int aa_batch_create(int32_t device, int32_t max_traces, aa_batch_t* out) {
    if (!out) return fail(AA_ERR_INVALID, "out is NULL");
    *out = nullptr;
    if (max_traces < 1 || max_traces > (1 << 20)) return fail(AA_ERR_INVALID, "max_traces must be in [1, 2^20]");
    int ndev = 0;
    AA_HIP(hipGetDeviceCount(&ndev));
    if (device < 0 || device >= ndev) return fail(AA_ERR_DEVICE, "device %d not available", device);
    AA_HIP(hipSetDevice(device));
    aa_batch_s* b = new aa_batch_s();
    b->device = device;
    b->max_traces = max_traces;
    b->table_bytes = table_size(max_traces);
    hipError_t e = hipHostMalloc(&b->h_table, b->table_bytes, hipHostMallocDefault);
    if (e == hipSuccess) e = hipMalloc(&b->d_table, b->table_bytes);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&b->copied, hipEventDisableTiming);
    if (e != hipSuccess) {
        if (b->h_table) (void)hipHostFree(b->h_table);
        if (b->d_table) (void)hipFree(b->d_table);
        delete b;
        return fail(AA_ERR_MEMORY, "batch allocation failed: %s", hipGetErrorString(e));
    }
    std::memset(b->h_table, 0, b->table_bytes);
    *out = b;
    return AA_OK;
}

int aa_batch_release(aa_batch_t b) {
    if (!b) return fail(AA_ERR_INVALID, "aa_batch_release: NULL batch");
    (void)hipSetDevice(b->device);
    if (b->pending) (void)hipEventSynchronize(b->copied);
    for (auto& kv : b->graphs) (void)hipGraphExecDestroy(kv.second);
    for (hipGraph_t g : b->graph_objs) (void)hipGraphDestroy(g);
    (void)hipEventDestroy(b->copied);
    (void)hipHostFree(b->h_table);
    (void)hipFree(b->d_table);
    delete b;
    return AA_OK;
}

int aa_batch_submit(aa_batch_t b, const aa_batch_trace* traces, int32_t count, const aa_batch_tuning* tuning,
                    void* stream_ptr) {
    if (!b || !traces || !tuning) return fail(AA_ERR_INVALID, "aa_batch_submit: NULL argument");
    if (count < 1 || count > b->max_traces) return fail(AA_ERR_INVALID, "count must be in [1, %d]", b->max_traces);
    const int wg = tuning->block_size > 0 ? tuning->block_size : 256;
    if (wg != 64 && wg != 128 && wg != 256 && wg != 512) return fail(AA_ERR_INVALID, "block_size must be 64..512");
    const int packet_rays = tuning->packet_rays > 0 ? tuning->packet_rays : 1024;
    if (packet_rays < 1 || packet_rays > (1 << 24)) return fail(AA_ERR_INVALID, "packet_rays out of range");
    for (int32_t i = 0; i < count; ++i) {
        const aa_batch_trace& t = traces[i];
        int rc = aa::validate_trace(t.scene, &t.params, t.d_sources, t.d_directions, t.d_diffuse_bank,
                                    t.d_diffuse_random, t.d_diffuse_indices, &t.outputs);
        if (rc != AA_OK) return rc;
        if (t.scene->device != b->device) return fail(AA_ERR_INVALID, "trace %d: scene on another device", i);
    }
    std::lock_guard<std::mutex> lock(b->mutex);
    hipStream_t stream = static_cast<hipStream_t>(stream_ptr);
    AA_HIP(hipSetDevice(b->device));
    const aa::DeviceProps& props = aa::device_props(b->device);
    if (!props.valid) return fail(AA_ERR_DEVICE, "cannot query device %d", b->device);

    // LDS plan shared by every trace of the launch
    int max_nodes = 0, max_hist = 0, max_surf = 0;
    for (int32_t i = 0; i < count; ++i) {
        const aa_batch_trace& t = traces[i];
        if (t.params.use_bvh) max_nodes = std::max(max_nodes, (int)t.scene->wide.nodes.size());
        max_hist = std::max(max_hist, t.params.num_sources * AA_NUM_BANDS * t.params.num_bins);
        if (t.params.num_sources == 1) max_surf = std::max(max_surf, t.scene->num_surfaces);
    }
    const int lds_nodes = std::min(std::max(tuning->lds_nodes, 0), max_nodes);
    const size_t lds_cap = props.lds_per_block > 0 ? props.lds_per_block : 65536;
    const size_t base = aa::stack_bytes(wg) + (size_t)lds_nodes * sizeof(aa::WideNode);
    const int surf_cap = (max_surf > 0 && base + (size_t)max_surf * 12 <= std::min<size_t>(lds_cap, 96 * 1024))
                             ? max_surf : 0;
    int hist_cap = 0;
    if (tuning->lds_histogram > 0) {
        const size_t room = lds_cap > base + (size_t)surf_cap * 12 ? lds_cap - base - (size_t)surf_cap * 12 : 0;
        hist_cap = (int)std::min<size_t>((size_t)max_hist, room / 4);
    }
    const bool lds_hist = hist_cap > 0;
    const size_t lds = base + (size_t)(lds_hist ? hist_cap : 0) * 4 + (size_t)surf_cap * 12;
    if (lds > lds_cap) return fail(AA_ERR_LIMIT, "LDS request %zu exceeds %zu bytes", lds, lds_cap);

    // the pinned table may still be read by the previous submission's copy
    if (b->pending) AA_HIP(hipEventSynchronize(b->copied));
    const size_t hdr_bytes = (sizeof(aa::BatchHeader) + 255) & ~size_t(255);
    aa::BatchHeader* h_hdr = static_cast<aa::BatchHeader*>(b->h_table);
    aa::BatchTrace* h_tr = reinterpret_cast<aa::BatchTrace*>(static_cast<char*>(b->h_table) + hdr_bytes);
    int64_t packets = 0;
    for (int32_t i = 0; i < count; ++i) {
        const aa_batch_trace& t = traces[i];
        aa::BatchTrace& d = h_tr[i];
        std::memset(&d, 0, sizeof(d));
        d.sv = aa::make_scene_view(t.scene, t.params.use_bvh != 0, 0);
        d.p = aa::make_ray_params(&t.params, t.d_sources, t.d_directions, t.d_diffuse_bank, t.d_diffuse_random,
                                  t.d_diffuse_indices);
        d.out = aa::make_outputs(&t.outputs, &t.params);
        d.error_flag = t.scene->d_error;
        d.num_packets = (t.params.num_rays + packet_rays - 1) / packet_rays;
        d.first_packet = (int)packets;
        packets += d.num_packets;
        d.hist_words = t.params.num_sources * AA_NUM_BANDS * t.params.num_bins;
        d.lds_surf = (t.params.num_sources == 1 && t.scene->num_surfaces <= surf_cap) ? 1 : 0;
        const aa::OutputExtents ext = aa::output_extents(&t.outputs, &t.params, t.scene->num_surfaces);
        for (int r = 0; r < 8; ++r) {
            d.zero_ptr[r] = static_cast<unsigned*>(ext.ptr[r]);
            d.zero_words[r] = (unsigned)(ext.bytes[r] / 4);
        }
    }
    if (packets > (int64_t)1 << 30) return fail(AA_ERR_LIMIT, "too many work items (%lld)", (long long)packets);
    h_hdr->num_traces = count;
    h_hdr->total_packets = (int)packets;
    h_hdr->packet_rays = packet_rays;
    h_hdr->queue_head = 0u;

    aa::BatchHeader* d_hdr = static_cast<aa::BatchHeader*>(b->d_table);
    const aa::BatchTrace* d_tr = reinterpret_cast<const aa::BatchTrace*>(static_cast<char*>(b->d_table) + hdr_bytes);
    const int per_cu = tuning->workgroups_per_cu > 0 ? tuning->workgroups_per_cu
                                                     : std::max(1, (int)(lds_cap / std::max<size_t>(lds, 1)));
    const int grid = std::max(1, std::min<int>(props.cu * per_cu, (int)packets));
    const int persistent_grid = tuning->use_graph ? std::max(1, props.cu * per_cu) : grid;

    auto enqueue = [&](hipStream_t s) -> hipError_t {
        hipError_t e = hipMemcpyAsync(b->d_table, b->h_table,
                                      tuning->use_graph ? b->table_bytes : hdr_bytes + (size_t)count * sizeof(aa::BatchTrace),
                                      hipMemcpyHostToDevice, s);
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL(aa::batch_zero_kernel, dim3(tuning->use_graph ? b->max_traces : count), dim3(256), 0, s,
                           d_hdr, d_tr);
        e = hipGetLastError();
        if (e != hipSuccess) return e;
        const int g = tuning->use_graph ? persistent_grid : grid;
#define AA_LAUNCH(W)                                                                                       \
    case W:                                                                                                \
        e = lds_hist ? launch_batch<W, true>(d_hdr, d_tr, g, lds, lds_nodes, hist_cap, surf_cap, s)       \
                     : launch_batch<W, false>(d_hdr, d_tr, g, lds, lds_nodes, hist_cap, surf_cap, s);     \
        break;
        switch (wg) {
            AA_LAUNCH(64)
            AA_LAUNCH(128)
            AA_LAUNCH(256)
            AA_LAUNCH(512)
        }
#undef AA_LAUNCH
        return e;
    };

    if (tuning->use_graph) {
        const auto key = std::make_tuple(wg, lds_nodes, hist_cap, surf_cap, persistent_grid, (int)lds_hist, stream);
        auto it = b->graphs.find(key);
        if (it == b->graphs.end()) {
            hipStream_t cap;
            AA_HIP(hipStreamCreateWithFlags(&cap, hipStreamNonBlocking));
            hipError_t e = hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal);
            if (e == hipSuccess) e = enqueue(cap);
            hipGraph_t graph = nullptr;
            hipError_t e2 = hipStreamEndCapture(cap, &graph);
            (void)hipStreamDestroy(cap);
            if (e != hipSuccess || e2 != hipSuccess)
                return fail(AA_ERR_DEVICE, "batch graph capture failed: %s",
                            hipGetErrorString(e != hipSuccess ? e : e2));
            hipGraphExec_t exec = nullptr;
            AA_HIP(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
            b->graph_objs.push_back(graph);
            it = b->graphs.emplace(key, exec).first;
        }
        AA_HIP(hipGraphLaunch(it->second, stream));
    } else {
        hipError_t e = enqueue(stream);
        if (e != hipSuccess) return fail(AA_ERR_DEVICE, "batch launch failed: %s", hipGetErrorString(e));
    }
    AA_HIP(hipEventRecord(b->copied, stream));
    b->pending = true;
    return AA_OK;
}

}  // extern "C"
