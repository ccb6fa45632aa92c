// Persistent batched tracer: many traces (scenes x listener channels x
// keyframes, uneven triangle and ray counts) in one launch.
//
// The reference runs one trace per call and fans calls out over a thread pool
// (acoustic_agent/batch.py:139-184, api.py:477-639); here a batch is a table
// of trace descriptors in device memory and a queue of work items, each item
// a packet of rays of one trace.  A grid sized to fill the GPU pulls items,
// one wave at a time with one atomic per item, until the queue is empty
// (every wave reaches that exit), and traces the packet with the per-ray
// routine of aa_device.h (so a trace gives the same result alone, in a batch,
// or through aa_trace, up to the order of float atomic sums).
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
#include <new>
#include <numeric>
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
    int first_packet;  // prefix sum of packets over the table
    int num_packets;
    int packet_rays;   // lane tasks per packet of this trace (the last packet may hold fewer)
    int lds_surf;      // single source and S small enough for LDS tallies
    unsigned* zero_ptr[8];
    unsigned zero_words[8];
};

constexpr int kAutoPacketRays = 1024;  // smallest packet of the automatic (guided) sizes on a full GPU
constexpr int kGuidedMaxRays = 4096;   // largest packet of the guided sizes

// Work queues, one per XCD of the MI355X (8 XCDs of 32 CUs, each with its own
// L2).  Workgroups are dispatched to the XCDs round-robin, so workgroup b
// runs on XCD b % kQueues; the host gives each queue whole scenes (balanced
// by packets), and a wave takes packets from its own XCD's queue first, so
// an XCD's L2 holds the nodes and triangles of one or two scenes at a time
// instead of every scene in flight on the chip.  A wave whose queue is empty
// takes packets from the next queues in turn, so the XCDs also finish
// together.  Items are global packet indices: queue q holds
// [queue_begin[q], queue_begin[q + 1]).
constexpr int kQueues = 8;

struct BatchHeader {
    int num_traces;
    int total_packets;
    int packet_rays;
    int queue_begin[kQueues + 1];
    unsigned queue_head[kQueues];  // packets claimed from each queue (may run past its end)
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

// A wave-uniform copy of *src, word by word through readfirstlane.  The batch
// table lives in ordinary global memory that the kernel also writes (through
// other pointers), so the compiler cannot prove it unchanged and would read
// each field with vector loads, again at every use; the copy puts the
// fields in scalar registers once per work item.
template <typename T>
__device__ __forceinline__ T uniform_copy(const T* src) {
    static_assert(sizeof(T) % 4 == 0, "word-sized struct");
    T dst;
    const int* s = reinterpret_cast<const int*>(src);
    int* d = reinterpret_cast<int*>(&dst);
#pragma unroll
    for (int i = 0; i < (int)(sizeof(T) / 4); ++i) d[i] = __builtin_amdgcn_readfirstlane(s[i]);
    return dst;
}

// Work items are claimed per wave, so that a wave whose rays run long does
// not hold up the other waves of its workgroup; each wave has its own surface
// tallies and its own ray cursor in LDS.
// Registers are held to 3 waves per SIMD, as for the single-trace kernel.
template <int WG>
__global__ __launch_bounds__(WG) __attribute__((amdgpu_waves_per_eu(3))) void batch_kernel(BatchHeader* hdr, const BatchTrace* traces, int surf_cap) {
    constexpr int kWaves = WG / 64;
    extern __shared__ __align__(16) unsigned char smem[];
    unsigned short* s_stack = reinterpret_cast<unsigned short*>(smem);
    float* s_queue = reinterpret_cast<float*>(smem + stack_only_bytes(WG));
    int* s_surf = reinterpret_cast<int*>(smem + private_bytes(WG));
    __shared__ unsigned s_cursor[kWaves];

    const int wave = (int)__builtin_amdgcn_readfirstlane(threadIdx.x >> 6);  // uniform: scalar LDS addresses
    const int lane = (int)(threadIdx.x & 63);
    int* my_hits = s_surf + (size_t)wave * surf_cap * 3;
    int* my_contrib = my_hits + surf_cap;
    float* my_energy = reinterpret_cast<float*>(my_contrib + surf_cap);
    const int num_traces = hdr->num_traces;
    const int xcd = (int)(blockIdx.x % kQueues);
    int drained = 0;  // wave-uniform: queues (from this XCD's on) found empty
    while (true) {
        int item = -1;
        while (drained < kQueues) {
            const int qx = (xcd + drained) % kQueues;
            unsigned claimed = 0u;
            if (lane == 0) claimed = atomicAdd(&hdr->queue_head[qx], 1u);
            // readfirstlane (not a shuffle) makes the item provably
            // wave-uniform, so the trace's argument struct below is read with
            // scalar loads into SGPRs instead of occupying some 60 VGPRs
            const int k = (int)__builtin_amdgcn_readfirstlane(claimed);
            const int q_begin = __builtin_amdgcn_readfirstlane(hdr->queue_begin[qx]);
            const int q_end = __builtin_amdgcn_readfirstlane(hdr->queue_begin[qx + 1]);
            if (k < q_end - q_begin) {
                item = q_begin + k;
                break;
            }
            ++drained;
        }
        if (item < 0) break;  // wave-uniform: every wave exits once every queue is drained
        int lo = 0, hi = num_traces - 1;  // last trace with first_packet <= item
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (traces[mid].first_packet <= item)
                lo = mid;
            else
                hi = mid - 1;
        }
        const BatchTrace& bt = traces[lo];
        const SceneView sv = uniform_copy(&bt.sv);
        const RayParams p = uniform_copy(&bt.p);
        const Outputs out = uniform_copy(&bt.out);
        const int first_packet = __builtin_amdgcn_readfirstlane(bt.first_packet);
        const int packet_rays = __builtin_amdgcn_readfirstlane(bt.packet_rays);
        const bool lds_surf = __builtin_amdgcn_readfirstlane(bt.lds_surf) != 0;
        if (lds_surf) {
            for (int i = lane; i < sv.num_surfaces; i += 64) {
                my_hits[i] = 0;
                my_contrib[i] = 0;
                my_energy[i] = 0.0f;
            }
        }
        if (lane == 0) s_cursor[wave] = 0u;
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        __builtin_amdgcn_wave_barrier();
        const int begin = (item - first_packet) * packet_rays;
        const int count = min(packet_rays, p.num_rays * p.groups - begin);
        Lds lds;
        lds.stack = s_stack;
        lds.queue = s_queue;
        lds.log = reinterpret_cast<float*>(smem + log_offset(WG));
        lds.hist = nullptr;
        lds.hits = lds_surf ? my_hits : nullptr;
        lds.contrib = lds_surf ? my_contrib : nullptr;
        lds.senergy = lds_surf ? my_energy : nullptr;
        LaneStats st;
        trace_rays_refill<WG, false, false, true>(sv, p, out, lds, &s_cursor[wave], begin, count, st);
        flush_stats(st, p, out);
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        __builtin_amdgcn_wave_barrier();
        if (lds_surf) {
            for (int i = lane; i < sv.num_surfaces; i += 64) {
                if (my_hits[i]) atomicAdd(&out.hit_counts[i], my_hits[i]);
                if (my_contrib[i]) atomicAdd(&out.contrib_counts[i], my_contrib[i]);
                if (my_energy[i] != 0.0f) atomicAdd(&out.surface_energy[i], my_energy[i]);
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        __builtin_amdgcn_wave_barrier();
    }
}

}  // namespace aa

namespace aa {
struct FillVectors;
}

struct aa_batch_s {
    int32_t device = 0;
    int32_t max_traces = 0;
    void* h_table = nullptr;  // pinned: header followed by traces
    void* d_table = nullptr;
    size_t table_bytes = 0;
    hipEvent_t copied = nullptr;  // the last submission's table copy has finished
    bool pending = false;
    aa::FillVectors* fill_work = nullptr;  // host work arrays of the table fill (max_traces entries)
    std::map<std::tuple<int, int, int, int, int, int, hipStream_t>, hipGraphExec_t> graphs;
    std::vector<hipGraph_t> graph_objs;
    std::mutex mutex;
};

namespace {

template <int WG>
hipError_t launch_batch(aa::BatchHeader* hdr, const aa::BatchTrace* traces, int grid, size_t lds, int surf_cap,
                        hipStream_t stream) {
    auto kern = aa::batch_kernel<WG>;
    if (lds > 64 * 1024) {
        hipError_t e = hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL(kern, dim3(grid), dim3(WG), lds, stream, hdr, traces, surf_cap);
    return hipGetLastError();
}

template <int WG>
int occupancy_blocks(size_t lds) {
    int blocks = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, (const void*)aa::batch_kernel<WG>, WG, lds) != hipSuccess)
        return 1;
    return blocks > 0 ? blocks : 1;
}

}  // namespace

namespace aa {

// Launch plan of one batch: validated traces, LDS layout, persistent grid
// and packet size (shared by aa_batch_submit and aa_batch_submit_static).
struct BatchPlan {
    int device = 0, wg = 256, waves = 4, surf_cap = 0, per_cu = 1, packet_rays = 0, grid = 1, persistent_grid = 1;
    size_t lds = 0;
    bool guided = false;     // packet sizes per trace (guided_packet), packet_rays the smallest
    int64_t resident = 1;    // resident waves of the launch
};

// Guided packet size of a trace whose queue (of `queue_tasks` lane tasks)
// still holds `remaining` when the trace's first packet is taken, in a batch
// of `batch_tasks`: the queues advance together (waves steal from the other
// queues once their own is empty), so about remaining / queue_tasks of the
// batch is left, and a packet takes 1 / (2 W) of that for the W resident
// waves; as a power-of-two multiple of the smallest size, at most
// kGuidedMaxRays.  A wave traces a packet to its last ray before it takes
// the next one, and the last rays leave most of its lanes idle (on batched
// mic arrays ~10 % of the waves' time was spent in these drains at 1024-ray
// packets); large packets early in a batch drain less often, small ones at
// its end keep the waves finishing together, and a small or lopsided batch
// (a served batch of a few scenes, which fill one or two queues) keeps the
// smallest size.  Measured on MI355X, batched mic arrays (256 scenes x 4
// mics): fixed 1024 / 2048 / 4096 / 8192-ray packets 942 / 922 / 929 / 1008 ms
// per batch, guided 887 ms.
int guided_packet(int64_t remaining, int64_t queue_tasks, int64_t batch_tasks, const BatchPlan& plan) {
    if (!plan.guided || queue_tasks <= 0) return plan.packet_rays;
    const double left = (double)remaining / (double)queue_tasks * (double)batch_tasks;
    const double want = left / (2.0 * (double)plan.resident);
    int64_t size = plan.packet_rays;
    while (size * 2 <= kGuidedMaxRays && (double)(size * 2) <= want) size *= 2;
    return (int)size;
}

size_t table_bytes(int traces) {
    const size_t hdr = (sizeof(BatchHeader) + 255) & ~size_t(255);
    return hdr + (size_t)traces * sizeof(BatchTrace);
}

int plan_batch(int device, const aa_batch_trace* traces, int32_t count, const aa_batch_tuning* tuning,
               BatchPlan& plan) {
    plan.device = device;
    plan.wg = tuning->block_size > 0 ? tuning->block_size : 256;
    const int wg = plan.wg;
    if (wg != 64 && wg != 128 && wg != 256 && wg != 512) return fail(AA_ERR_INVALID, "block_size must be 64..512");
    if (tuning->packet_rays > (1 << 24)) return fail(AA_ERR_INVALID, "packet_rays out of range");
    for (int32_t i = 0; i < count; ++i) {
        const aa_batch_trace& t = traces[i];
        int rc = validate_trace(t.scene, &t.params, t.d_sources, t.d_directions, t.d_diffuse_bank,
                                t.d_diffuse_random, t.d_diffuse_indices, &t.outputs);
        if (rc != AA_OK) return rc;
        if (t.scene->device != device) return fail(AA_ERR_INVALID, "trace %d: scene on another device", i);
        if (t.params.visual_limit > 0)
            return fail(AA_ERR_INVALID, "trace %d: visual candidates are recorded by aa_trace only", i);
    }
    if (hipSetDevice(device) != hipSuccess) return fail(AA_ERR_DEVICE, "cannot select device %d", device);
    const DeviceProps& props = device_props(device);
    if (!props.valid) return fail(AA_ERR_DEVICE, "cannot query device %d", device);

    // LDS plan shared by every trace of the launch: per-wave surface tallies
    // (histograms go to global atomics)
    int max_surf = 0;
    for (int32_t i = 0; i < count; ++i)
        if (traces[i].params.num_sources == 1) max_surf = std::max(max_surf, traces[i].scene->num_surfaces);
    plan.waves = wg / 64;
    const size_t lds_cap = props.lds_per_block > 0 ? props.lds_per_block : 65536;
    // surface tallies only while the workgroups the register budget allows
    // (12 waves per CU) stay resident
    const size_t lds_resident = (props.lds_per_cu > 0 ? props.lds_per_cu : lds_cap) /
                                (size_t)std::max(1, 12 * 64 / wg);
    plan.surf_cap =
        (max_surf > 0 && private_bytes(wg) + (size_t)plan.waves * max_surf * 12 <= lds_resident) ? max_surf : 0;
    plan.lds = private_bytes(wg) + (size_t)plan.waves * plan.surf_cap * 12;
    // the kernel's static LDS (the per-wave cursors) comes on top of the
    // dynamic request
    const size_t static_lds = (size_t)plan.waves * sizeof(unsigned);
    if (plan.lds + static_lds > lds_cap)
        return fail(AA_ERR_LIMIT, "LDS request %zu exceeds %zu bytes", plan.lds, lds_cap);

    int per_cu = tuning->workgroups_per_cu;
    if (per_cu <= 0) {
        switch (wg) {
            case 64: per_cu = occupancy_blocks<64>(plan.lds); break;
            case 128: per_cu = occupancy_blocks<128>(plan.lds); break;
            case 256: per_cu = occupancy_blocks<256>(plan.lds); break;
            default: per_cu = occupancy_blocks<512>(plan.lds); break;
        }
    }
    plan.per_cu = per_cu;
    // packet size: the tuning's, or (0) guided sizes per trace (guided_packet)
    // whose smallest is about two work items per resident wave, so that small
    // batches still spread over the whole GPU, at most kAutoPacketRays, so
    // that the end of a large batch keeps enough work items to balance uneven
    // traces (measured on 256 scenes x 4-mic arrays, 1024 traces, with one
    // size for the whole batch: 1024-ray packets 1122 ms per batch, the 16384
    // of two items per wave 1300 ms)
    int64_t tasks = 0;
    for (int32_t i = 0; i < count; ++i) {
        const aa_batch_trace& t = traces[i];
        const int g = t.params.sources_per_lane > 0 ? std::min(t.params.sources_per_lane, t.params.num_sources)
                                                    : std::min(8, t.params.num_sources);
        tasks += (int64_t)t.params.num_rays * ((t.params.num_sources + g - 1) / g);
    }
    int packet_rays = tuning->packet_rays;
    const int64_t resident = (int64_t)props.cu * per_cu * plan.waves;
    plan.resident = std::max<int64_t>(1, resident);
    if (packet_rays <= 0) {
        // the smallest guided size: about two work items per resident wave
        // (small batches still spread over the whole GPU), at most
        // kAutoPacketRays
        const int64_t want = (tasks + 2 * resident - 1) / (2 * resident);
        packet_rays = (int)std::min<int64_t>(kAutoPacketRays, std::max<int64_t>(64, (want + 63) / 64 * 64));
        plan.guided = true;
    }
    plan.packet_rays = packet_rays;
    // packets at the smallest size: an upper bound under guided sizes
    int64_t packets = 0;
    for (int32_t i = 0; i < count; ++i) {
        const aa_batch_trace& t = traces[i];
        const int g = t.params.sources_per_lane > 0 ? std::min(t.params.sources_per_lane, t.params.num_sources)
                                                    : std::min(8, t.params.num_sources);
        const int64_t trace_tasks = (int64_t)t.params.num_rays * ((t.params.num_sources + g - 1) / g);
        packets += (trace_tasks + packet_rays - 1) / packet_rays;
    }
    if (packets > (int64_t)1 << 30) return fail(AA_ERR_LIMIT, "too many work items (%lld)", (long long)packets);
    plan.persistent_grid = std::max(1, props.cu * per_cu);
    plan.grid = (int)std::max<int64_t>(1, std::min<int64_t>(plan.persistent_grid, (packets + plan.waves - 1) / plan.waves));
    return AA_OK;
}

// Host work arrays of fill_table, one entry per trace (owned by the batch
// object; static for aa_batch_submit_static).
struct FillWork {
    int64_t* tasks;          // lane tasks per trace
    int32_t* by_scene;       // trace indices sorted by scene, then index
    int32_t* group_start;    // per scene: first position in by_scene
    int32_t* group_len;
    int64_t* group_tasks;
    int32_t* group_order;    // scenes by packets, largest first
    int32_t* group_queue;
    int32_t* order;          // table slot -> trace index
};

template <int N>
struct FillArrays {
    int64_t tasks[N], group_tasks[N];
    int32_t by_scene[N], group_start[N], group_len[N], group_order[N], group_queue[N], order[N];
    FillWork view() {
        return FillWork{tasks, by_scene, group_start, group_len, group_tasks, group_order, group_queue, order};
    }
};

struct FillVectors {
    std::vector<int64_t> tasks, group_tasks;
    std::vector<int32_t> by_scene, group_start, group_len, group_order, group_queue, order;
    explicit FillVectors(size_t n)
        : tasks(n), group_tasks(n), by_scene(n), group_start(n), group_len(n), group_order(n), group_queue(n),
          order(n) {}
    FillWork view() {
        return FillWork{tasks.data(),     by_scene.data(),    group_start.data(), group_len.data(),
                        group_tasks.data(), group_order.data(), group_queue.data(), order.data()};
    }
};

// The descriptor table (header, then one BatchTrace per trace) of a planned
// batch, written to host memory `table` of at least table_bytes(count).
int fill_table(const aa_batch_trace* traces, int32_t count, const BatchPlan& plan, FillWork& work, void* table) {
    const size_t hdr_bytes = (sizeof(BatchHeader) + 255) & ~size_t(255);
    BatchHeader* h_hdr = static_cast<BatchHeader*>(table);
    BatchTrace* h_tr = reinterpret_cast<BatchTrace*>(static_cast<char*>(table) + hdr_bytes);
    // Lane tasks per trace, then the queue of every scene: scenes in order of
    // their tasks (largest first, ties in order of first appearance), each
    // to the queue with the fewest tasks so far; a queue's traces are
    // contiguous in the table, in the caller's order within a scene, and
    // each trace's packet size follows the work its queue still holds at the
    // trace (guided_packet).  Only the caller's work arrays are used
    // (in-place sorts), so the fill allocates nothing.
    int64_t* task_count = work.tasks;
    for (int32_t i = 0; i < count; ++i) {
        const aa_batch_trace& t = traces[i];
        const int g = t.params.sources_per_lane > 0 ? std::min(t.params.sources_per_lane, t.params.num_sources)
                                                    : std::min(8, t.params.num_sources);
        task_count[i] = (int64_t)t.params.num_rays * ((t.params.num_sources + g - 1) / g);
        work.by_scene[i] = i;
    }
    std::sort(work.by_scene, work.by_scene + count, [&](int32_t a, int32_t b) {
        const auto sa = reinterpret_cast<uintptr_t>(traces[a].scene), sb = reinterpret_cast<uintptr_t>(traces[b].scene);
        return sa != sb ? sa < sb : a < b;
    });
    int32_t groups = 0;
    for (int32_t k = 0; k < count; ++k) {
        const int32_t i = work.by_scene[k];
        if (k == 0 || traces[i].scene != traces[work.by_scene[k - 1]].scene) {
            work.group_start[groups] = k;
            work.group_len[groups] = 0;
            work.group_tasks[groups] = 0;
            work.group_order[groups] = groups;
            ++groups;
        }
        work.group_len[groups - 1] += 1;
        work.group_tasks[groups - 1] += task_count[i];
    }
    // (a group's first trace is its first appearance: by_scene is sorted by
    // index within a scene)
    std::sort(work.group_order, work.group_order + groups, [&](int32_t a, int32_t b) {
        if (work.group_tasks[a] != work.group_tasks[b]) return work.group_tasks[a] > work.group_tasks[b];
        return work.by_scene[work.group_start[a]] < work.by_scene[work.group_start[b]];
    });
    int64_t load[kQueues] = {0};
    int64_t batch_tasks = 0;
    for (int32_t i = 0; i < count; ++i) batch_tasks += task_count[i];
    for (int32_t k = 0; k < groups; ++k) {
        const int32_t g = work.group_order[k];
        const int q = (int)(std::min_element(load, load + kQueues) - load);
        work.group_queue[g] = q;
        load[q] += work.group_tasks[g];
    }
    int32_t* order = work.order;
    int32_t placed = 0;
    int32_t queue_slot[kQueues + 1];  // first table slot of each queue
    for (int q = 0; q < kQueues; ++q) {
        queue_slot[q] = placed;
        for (int32_t k = 0; k < groups; ++k) {
            const int32_t g = work.group_order[k];
            if (work.group_queue[g] != q) continue;
            for (int32_t m = 0; m < work.group_len[g]; ++m) order[placed++] = work.by_scene[work.group_start[g] + m];
        }
    }
    queue_slot[kQueues] = placed;
    int queue_begin[kQueues + 1];
    int64_t packets = 0;
    int q = 0;
    int64_t queued = 0;  // tasks of queue q in the slots before this one
    queue_begin[0] = 0;
    for (int32_t slot = 0; slot < count; ++slot) {
        while (slot >= queue_slot[q + 1]) {  // the next non-empty queue starts here
            ++q;
            queue_begin[q] = (int)packets;
            queued = 0;
        }
        const int32_t i = order[slot];
        const aa_batch_trace& t = traces[i];
        BatchTrace& d = h_tr[slot];
        std::memset(&d, 0, sizeof(d));
        d.sv = make_scene_view(t.scene, t.params.use_bvh != 0);
        d.p = make_ray_params(&t.params, t.d_sources, t.d_directions, t.d_diffuse_bank, t.d_diffuse_random,
                              t.d_diffuse_indices);
        // interleaved ray order for multi-source traces only (measured:
        // 64-keyframe traces +2.9 %; single-source mic arrays -0.3 %, whose
        // packets lose the coherence of neighbouring first-bounce rays)
        if (d.p.num_sources <= 1) d.p.ray_stride = 1;
        d.out = make_outputs(&t.outputs, &t.params);
        const int64_t trace_tasks = (int64_t)d.p.num_rays * d.p.groups;
        if (trace_tasks > (int64_t)1 << 30) return fail(AA_ERR_LIMIT, "trace %d has too many lane tasks", i);
        d.packet_rays = guided_packet(load[q] - queued, load[q], batch_tasks, plan);
        d.num_packets = (int)((trace_tasks + d.packet_rays - 1) / d.packet_rays);
        d.first_packet = (int)packets;
        packets += d.num_packets;
        queued += trace_tasks;
        d.lds_surf = (t.params.num_sources == 1 && t.scene->num_surfaces <= plan.surf_cap) ? 1 : 0;
        const OutputExtents ext = output_extents(&t.outputs, &t.params, t.scene->num_surfaces);
        for (int r = 0; r < 8; ++r) {
            d.zero_ptr[r] = static_cast<unsigned*>(ext.ptr[r]);
            d.zero_words[r] = (unsigned)(ext.bytes[r] / 4);
        }
    }
    while (q < kQueues) queue_begin[++q] = (int)packets;
    h_hdr->num_traces = count;
    h_hdr->total_packets = (int)packets;
    h_hdr->packet_rays = plan.packet_rays;
    for (int q = 0; q <= kQueues; ++q) h_hdr->queue_begin[q] = queue_begin[q];
    for (int q = 0; q < kQueues; ++q) h_hdr->queue_head[q] = 0u;
    return AA_OK;
}

// Copy `copy_bytes` of the host table to `d_table`, zero the outputs of
// `zero_traces` entries, run the persistent kernel on `grid` workgroups.
hipError_t enqueue_batch(void* d_table, const void* h_table, size_t copy_bytes, int zero_traces,
                         const BatchPlan& plan, int grid, hipStream_t s) {
    const size_t hdr_bytes = (sizeof(BatchHeader) + 255) & ~size_t(255);
    BatchHeader* d_hdr = static_cast<BatchHeader*>(d_table);
    const BatchTrace* d_tr = reinterpret_cast<const BatchTrace*>(static_cast<char*>(d_table) + hdr_bytes);
    hipError_t e = hipMemcpyAsync(d_table, h_table, copy_bytes, hipMemcpyHostToDevice, s);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(batch_zero_kernel, dim3(zero_traces), dim3(256), 0, s, d_hdr, d_tr);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
    switch (plan.wg) {
        case 64: return launch_batch<64>(d_hdr, d_tr, grid, plan.lds, plan.surf_cap, s);
        case 128: return launch_batch<128>(d_hdr, d_tr, grid, plan.lds, plan.surf_cap, s);
        case 256: return launch_batch<256>(d_hdr, d_tr, grid, plan.lds, plan.surf_cap, s);
        default: return launch_batch<512>(d_hdr, d_tr, grid, plan.lds, plan.surf_cap, s);
    }
}

}  // namespace aa

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
    b->table_bytes = aa::table_bytes(max_traces);
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
    try {
        b->fill_work = new aa::FillVectors((size_t)max_traces);
    } catch (const std::bad_alloc&) {
        (void)hipEventDestroy(b->copied);
        (void)hipHostFree(b->h_table);
        (void)hipFree(b->d_table);
        delete b;
        return fail(AA_ERR_MEMORY, "batch allocation failed (host work arrays)");
    }
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
    delete b->fill_work;
    delete b;
    return AA_OK;
}

int aa_batch_submit(aa_batch_t b, const aa_batch_trace* traces, int32_t count, const aa_batch_tuning* tuning,
                    void* stream_ptr) {
    if (!b || !traces || !tuning) return fail(AA_ERR_INVALID, "aa_batch_submit: NULL argument");
    if (count < 1 || count > b->max_traces) return fail(AA_ERR_INVALID, "count must be in [1, %d]", b->max_traces);
    aa::BatchPlan plan;
    int rc = aa::plan_batch(b->device, traces, count, tuning, plan);
    if (rc != AA_OK) return rc;
    std::lock_guard<std::mutex> lock(b->mutex);
    hipStream_t stream = static_cast<hipStream_t>(stream_ptr);
    AA_HIP(hipSetDevice(b->device));

    // the pinned table may still be read by the previous submission's copy
    if (b->pending) AA_HIP(hipEventSynchronize(b->copied));
    aa::FillWork work = b->fill_work->view();
    rc = aa::fill_table(traces, count, plan, work, b->h_table);
    if (rc != AA_OK) return rc;

    auto enqueue = [&](hipStream_t s) -> hipError_t {
        // a graph replays a copy of the whole table and zeroes up to max_traces
        // entries (the zero kernel skips entries past num_traces)
        const size_t bytes = tuning->use_graph ? b->table_bytes : aa::table_bytes(count);
        return aa::enqueue_batch(b->d_table, b->h_table, bytes, tuning->use_graph ? b->max_traces : count, plan,
                                 tuning->use_graph ? plan.persistent_grid : plan.grid, s);
    };

    if (tuning->use_graph) {
        const auto key = std::make_tuple(plan.wg, 0, 0, plan.surf_cap, plan.persistent_grid, 0, stream);
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

// ---------------------------------------------------------------------------
// batches without allocation: the descriptor table in a static device array
// ---------------------------------------------------------------------------

namespace {

// header (rounded to 256 bytes) + AA_STATIC_BATCH_TRACES descriptors
constexpr size_t kStaticTableBytes =
    ((sizeof(aa::BatchHeader) + 255) & ~size_t(255)) + (size_t)AA_STATIC_BATCH_TRACES * sizeof(aa::BatchTrace);

// The table of aa_batch_submit_static.  Each GPU holds its own copy (at its
// own address) for the life of the process.  In host code this name is a
// host-side stand-in, never a device address: host code uses the address
// hipGetSymbolAddress returns for the current device.
__device__ __align__(256) unsigned char scratch[kStaticTableBytes];

constexpr int kMaxDevices = 64;
std::mutex g_static_mutex;              // guards the pointers and the host staging table
void* g_static_table[kMaxDevices] = {};  // device address of `scratch` per device
alignas(256) unsigned char g_static_staging[kStaticTableBytes];  // host table, .bss (no allocation)
aa::FillArrays<AA_STATIC_BATCH_TRACES> g_static_fill;  // host work arrays of the table fill, .bss

}  // namespace

extern "C" int aa_batch_submit_static(int32_t device, const aa_batch_trace* traces, int32_t count,
                                      const aa_batch_tuning* tuning, void* stream_ptr) {
    if (!traces || !tuning) return fail(AA_ERR_INVALID, "aa_batch_submit_static: NULL argument");
    if (device < 0 || device >= kMaxDevices) return fail(AA_ERR_INVALID, "device %d out of range", device);
    if (count < 1 || count > AA_STATIC_BATCH_TRACES)
        return fail(AA_ERR_LIMIT, "aa_batch_submit_static takes 1..%d traces (got %d)", AA_STATIC_BATCH_TRACES,
                    count);
    if (tuning->use_graph) return fail(AA_ERR_INVALID, "aa_batch_submit_static does not capture graphs");
    aa::BatchPlan plan;
    int rc = aa::plan_batch(device, traces, count, tuning, plan);  // selects `device`
    if (rc != AA_OK) return rc;
    std::lock_guard<std::mutex> lock(g_static_mutex);
    if (!g_static_table[device]) {
        void* ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr)
            return fail(AA_ERR_DEVICE, "hipGetSymbolAddress(scratch) failed on device %d: %s", device,
                        hipGetErrorString(e));
        g_static_table[device] = ptr;
    }
    aa::FillWork work = g_static_fill.view();
    rc = aa::fill_table(traces, count, plan, work, g_static_staging);
    if (rc != AA_OK) return rc;
    // from pageable host memory: the staging table has been read when the
    // copy call returns, so the next call may refill it
    const hipError_t e = aa::enqueue_batch(g_static_table[device], g_static_staging, aa::table_bytes(count), count,
                                           plan, plan.grid, static_cast<hipStream_t>(stream_ptr));
    if (e != hipSuccess) return fail(AA_ERR_DEVICE, "static batch launch failed: %s", hipGetErrorString(e));
    return AA_OK;
}
