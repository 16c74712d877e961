// Plain wavefront engine (src/gpu/wavefront.cu:48-601): one pass renders one sample of every
// pixel through per-bounce stage kernels -- extend, classify (material queues), shade per
// material, shadow, mark + stable compaction -- and films every path at the end. The slot of a
// path is its pixel. Stage kernels read their item counts from device memory and are launched
// over an upper bound (the active count), so the host synchronises once per bounce, to learn
// the compacted active count. The order in which L is accumulated is the megakernel's:
// emission in classify, then the NEE contribution in the shadow stage before the next classify
// (src/gpu/wavefront.cu:118,246-250), so both arms film the same bits.
#include <algorithm>

#include "wf_wavefront.h"

namespace wf::gpu {

namespace {

constexpr int kThreads = 256;

// counters[]: 0..2 material queue counts, 3 shadow count, 4 overflow, 5 next active count.
enum { kCntLambert = 0, kCntMetal = 1, kCntDiel = 2, kCntShadow = 3, kCntOverflow = 4, kCntActive = 5, kCntWords = 8 };

__global__ __launch_bounds__(kThreads) void wf_init_kernel(SceneView s, PathPool P, uint32_t* active, uint32_t sample,
                                                           uint32_t seed) {
    const int i = blockIdx.x * kThreads + threadIdx.x;
    const int pixels = s.width * s.height;
    if (i >= pixels) return;
    const int px = i % s.width, py = i / s.width;
    const PathState ps = init_path(s, px, py, sample, seed);
    store_new_path(P, (uint32_t)i, ps);
    active[i] = (uint32_t)i;
}

__device__ __forceinline__ uint32_t clamp_count(uint32_t c, uint32_t cap) { return c < cap ? c : cap; }

// Traversal kernels run on a persistent grid (occupancy x CUs) with the wave schedule of
// trace_items; each workgroup stages the BVH top (and, when the whole tree is staged and they fit,
// the primitive records) into LDS once.
constexpr int kTraceThreads = 512;
constexpr int kTraceRefill = 32;
constexpr int kTraceLeafBatch = 16;
constexpr uint32_t kTraceSlice = 64;

__device__ __forceinline__ LdsNodes stage_top(const SceneView& s, Node* lds, int n_lds, int n_lds_prims) {
    for (int k = threadIdx.x; k < n_lds; k += blockDim.x) lds[k] = s.nodes[k];
    float4* lprim = reinterpret_cast<float4*>(lds + n_lds);
    for (int k = threadIdx.x; k < n_lds_prims; k += blockDim.x) lprim[k] = lds_prim_record(s, k);
    __syncthreads();
    return lds_nodes_of(s, lds, n_lds, n_lds_prims > 0 ? lprim : nullptr);
}

__global__ __launch_bounds__(kTraceThreads) void wf_extend_kernel(SceneView s, PathPool P, const uint32_t* active,
                                                                  uint32_t n_active, int n_lds, int n_lds_prims) {
    extern __shared__ Node lds_nodes[];
    if (block_idle(n_active, kTraceSlice)) return;
    const LdsNodes src = stage_top(s, lds_nodes, n_lds, n_lds_prims);
    trace_items(
        n_active, kTraceSlice, s.n_nodes, src, false, kTraceRefill, kTraceLeafBatch,
        [&](uint32_t idx, uint32_t* slot) {
            const uint32_t i = active[idx];
            *slot = i;
            const float4 o = P.ro[i], d = P.rd[i];
            Ray r;
            r.o = Vec3{o.x, o.y, o.z};
            r.d = Vec3{d.x, d.y, d.z};
            r.tmin = kRayEps;
            r.tmax = kInf;
            return r;
        },
        [&](uint32_t, uint32_t i, const Traversal& tr) { P.hit[i] = make_float2(tr.r.tmax, __int_as_float(tr.prim)); });
}

// Classify and shade_lambert push with one atomic per queue per 1024-thread workgroup.
constexpr int kPushThreads = 1024;

__global__ __launch_bounds__(kPushThreads) void wf_classify_kernel(SceneView s, PathPool P, const uint32_t* active,
                                                                   uint32_t n_active, uint32_t* alive, uint32_t* q0,
                                                                   uint32_t* q1, uint32_t* q2, uint32_t* cnt,
                                                                   uint32_t cap, int max_depth, DevCounters* ctr) {
    __shared__ uint32_t lds_cnt[3][16], lds_base[3];
    const uint32_t idx = blockIdx.x * kPushThreads + threadIdx.x;
    const bool valid = idx < n_active;
    uint32_t i = 0, inv = 0;
    int mat = -1;
    if (valid) {
        i = active[idx];
        const float2 hv = P.hit[i];
        Hit h{hv.x, __float_as_int(hv.y)};
        if (!hit_is_valid(s, h)) {
            inv = 1;
            h.prim = -1;
        }
        PathState ps = load_path(P, i, true);
        Surface f;
        mat = classify(s, ps, h, max_depth, &f);
        P.L[i] = make_float4(ps.L.x, ps.L.y, ps.L.z, 0.0f);
        if (mat >= 0) store_surface(P, i, f);
        alive[i] = mat >= 0 ? 1u : 0u;
    }
    uint32_t* qs[3] = {q0, q1, q2};
    const int q = valid ? mat : -1;
    const uint32_t slot = block_push_multi<3>(qs, cnt, cap, i, q, &cnt[kCntOverflow], lds_cnt, lds_base);
    if (q >= 0 && slot == 0xFFFFFFFFu) alive[i] = 0u;   // dropped on overflow: the path ends
    wave_add_u32(&ctr->invalid_hits, inv);
}


__global__ __launch_bounds__(kPushThreads) void wf_shade_lambert_kernel(SceneView s, PathPool P, ShadowQueue SQ,
                                                                        const uint32_t* queue, uint32_t* alive,
                                                                        uint32_t* cnt, uint32_t cap) {
    __shared__ uint32_t lds_cnt[1][16], lds_base[1];
    const uint32_t n = clamp_count(cnt[kCntLambert], cap);
    if (blockIdx.x * kPushThreads >= n) return;   // whole workgroup
    const uint32_t idx = blockIdx.x * kPushThreads + threadIdx.x;
    uint32_t i = 0;
    int has = 0;
    Ray sh{};
    Vec3 c{0.0f, 0.0f, 0.0f};
    if (idx < n) {
        i = queue[idx];
        const Surface f = load_surface(P, i);
        PathState ps = load_path(P, i, false);
        const bool live = shade_lambert(s, ps, f, &sh, &c, &has);
        store_scatter(P, i, ps);
        alive[i] = live ? 1u : 0u;
    }
    // Shadow-queue slots; a shadow ray that does not fit is dropped and counted (this bounce loses
    // its NEE contribution).
    const uint32_t slot = block_reserve(&cnt[kCntShadow], cap, has != 0, &cnt[kCntOverflow], lds_cnt, lds_base);
    if (has && slot != 0xFFFFFFFFu) {
        SQ.o[slot] = make_float4(sh.o.x, sh.o.y, sh.o.z, sh.tmax);
        SQ.d[slot] = make_float4(sh.d.x, sh.d.y, sh.d.z, __int_as_float((int)i));
        SQ.c[slot] = make_float4(c.x, c.y, c.z, 0.0f);
    }
}

template <int MAT>
__global__ __launch_bounds__(kThreads) void wf_shade_specular_kernel(SceneView s, PathPool P, const uint32_t* queue,
                                                                     uint32_t* alive, const uint32_t* cnt,
                                                                     uint32_t cap) {
    const uint32_t idx = blockIdx.x * kThreads + threadIdx.x;
    if (idx >= clamp_count(cnt[MAT], cap)) return;
    const uint32_t i = queue[idx];
    const Surface f = load_surface(P, i);
    PathState ps = load_path(P, i, false);
    const bool live = MAT == kMetal ? shade_metal(s, ps, f) : shade_dielectric(s, ps, f);
    store_scatter(P, i, ps);
    alive[i] = live ? 1u : 0u;
}

__global__ __launch_bounds__(kTraceThreads) void wf_shadow_kernel(SceneView s, PathPool P, ShadowQueue SQ,
                                                                  const uint32_t* cnt, uint32_t cap, int n_lds,
                                                                  int n_lds_prims) {
    extern __shared__ Node lds_nodes[];
    const uint32_t n = clamp_count(cnt[kCntShadow], cap);
    if (block_idle(n, kTraceSlice)) return;
    const LdsNodes src = stage_top(s, lds_nodes, n_lds, n_lds_prims);
    trace_items(
        n, kTraceSlice, s.n_nodes, src, true, kTraceRefill, kTraceLeafBatch,
        [&](uint32_t idx, uint32_t* slot) {
            const float4 o = SQ.o[idx], d = SQ.d[idx];
            *slot = (uint32_t)__float_as_int(d.w);
            Ray r;
            r.o = Vec3{o.x, o.y, o.z};
            r.d = Vec3{d.x, d.y, d.z};
            r.tmin = kRayEps;
            r.tmax = o.w;
            return r;
        },
        [&](uint32_t idx, uint32_t i, const Traversal& tr) {
            if (tr.prim < 0) {   // unoccluded
                const float4 c = SQ.c[idx];
                float4 l = P.L[i];
                l.x += c.x;
                l.y += c.y;
                l.z += c.z;
                P.L[i] = l;
            }
        });
}

__global__ __launch_bounds__(kThreads) void wf_mark_kernel(const uint32_t* active, uint32_t n_active,
                                                           const uint32_t* alive, uint32_t* flags) {
    const uint32_t idx = blockIdx.x * kThreads + threadIdx.x;
    if (idx < n_active) flags[idx] = alive[active[idx]];
}

// Bookkeeping between bounces (one thread): high-water marks, overflow, ray counts, trace.
__global__ void wf_update_kernel(const uint32_t* cnt, DevCounters* ctr, uint32_t n_active, uint32_t cap,
                                 uint32_t* trace_rec) {
    uint32_t c[4];
    for (int k = 0; k < 4; ++k) {
        c[k] = cnt[k] < cap ? cnt[k] : cap;
        if (c[k] > ctr->high_water[k]) ctr->high_water[k] = c[k];
    }
    if (n_active > ctr->high_water[4]) ctr->high_water[4] = n_active;
    ctr->overflows += cnt[kCntOverflow];
    ctr->extension_rays += n_active;
    ctr->shadow_rays += c[3];
    if (trace_rec) {
        trace_rec[0] = n_active;
        for (int k = 0; k < 4; ++k) trace_rec[1 + k] = c[k];
        trace_rec[5] = cnt[kCntActive];
    }
}

__global__ __launch_bounds__(kThreads) void wf_finish_kernel(DeviceFilm film, PathPool P, int pixels) {
    const int i = blockIdx.x * kThreads + threadIdx.x;
    if (i >= pixels) return;
    const float4 l = P.L[i];
    film_add3(&film.rgb[3 * i], &film.lum[i], &film.lum2[i], Vec3{l.x, l.y, l.z});
}

__global__ void zero_words_kernel(uint32_t* p, int n) {
    const int i = threadIdx.x;
    if (i < n) p[i] = 0u;
}

inline int blocks_for(uint32_t n) { return (int)((n + kThreads - 1) / kThreads); }

}  // namespace

WavefrontEngine::WavefrontEngine(const SceneView& scene, int width, int height) : scene_(scene), pixels_(width * height) {
    const size_t P = (size_t)pixels_;
    try {
        alloc(&pool_.ro, P * 16);
        alloc(&pool_.rd, P * 16);
        alloc(&pool_.T, P * 16);
        alloc(&pool_.L, P * 16);
        alloc(&pool_.key, P * 8);
        alloc(&pool_.hit, P * 8);
        alloc(&pool_.sp, P * 16);
        alloc(&pool_.sn, P * 16);
        alloc(&sq_.o, P * 16);
        alloc(&sq_.d, P * 16);
        alloc(&sq_.c, P * 16);
        alloc(&alive_, P * 4);
        for (int q = 0; q < 3; ++q) alloc(&queues_[q], P * 4);
        alloc(&active_[0], P * 4);
        alloc(&active_[1], P * 4);
        alloc(&flags_, P * 4);
        alloc(&counters_, kCntWords * 4);
        alloc(&trace_dev_, (size_t)kMaxTraceBounces * 8 * 4);
        check(scan_workspace_create(&ws_, pixels_), "scan_workspace_create");
        bytes_ += (size_t)((pixels_ + kMinScanTile - 1) / kMinScanTile + 1) * 4;
        for (auto& e : events_) check(hipEventCreate(&e), "hipEventCreate");
    } catch (...) {
        release();
        throw;
    }
}

WavefrontEngine::~WavefrontEngine() { release(); }

void WavefrontEngine::release() {
    for (void* p : allocs_) (void)hipFree(p);
    allocs_.clear();
    scan_workspace_destroy(&ws_);
    for (auto& e : events_)
        if (e) {
            (void)hipEventDestroy(e);
            e = nullptr;
        }
}

template <class T>
void WavefrontEngine::alloc(T** p, size_t bytes) {
    void* q = nullptr;
    check(hipMalloc(&q, bytes), "hipMalloc wavefront pool");
    allocs_.push_back(q);
    *p = (T*)q;
    bytes_ += bytes;
}

void WavefrontEngine::run_pass(const DeviceFilm& film, DevCounters* ctr, uint32_t sample, uint32_t seed, int max_depth,
                               const PassOptions& opt, StageTimes* stages, std::vector<BounceTrace>* trace,
                               int* bounces, hipStream_t stream) {
    const uint32_t P = (uint32_t)pixels_;
    const uint32_t cap = (opt.queue_capacity > 0 && (uint32_t)opt.queue_capacity < P) ? (uint32_t)opt.queue_capacity : P;
    const int n_lds = opt.lds_nodes < 0 ? std::min(scene_.n_nodes, kWavefrontLdsNodes) : std::min(scene_.n_nodes, opt.lds_nodes);
    // Primitive records join the nodes in LDS when the whole tree is staged and both fit in 64 KiB.
    const int n_lds_prims =
        (n_lds == scene_.n_nodes && (size_t)n_lds * sizeof(Node) + (size_t)scene_.n_prims * 16 <= 64 * 1024)
            ? scene_.n_prims : 0;
    const size_t lds_bytes = (size_t)n_lds * sizeof(Node) + (size_t)n_lds_prims * 16;
    if (lds_bytes != trace_lds_) {
        if (cus_ == 0) {
            int dev = 0;
            check(hipGetDevice(&dev), "hipGetDevice");
            check(hipDeviceGetAttribute(&cus_, hipDeviceAttributeMultiprocessorCount, dev), "hipDeviceGetAttribute");
        }
        int per = 0;
        check(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, wf_extend_kernel, kTraceThreads, lds_bytes),
              "hipOccupancyMaxActiveBlocksPerMultiprocessor");
        extend_grid_ = std::max(1, per) * cus_;
        check(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, wf_shadow_kernel, kTraceThreads, lds_bytes),
              "hipOccupancyMaxActiveBlocksPerMultiprocessor");
        shadow_grid_ = std::max(1, per) * cus_;
        trace_lds_ = lds_bytes;
    }
    const int scan_tile = opt.scan_tile > 0 ? opt.scan_tile : kDefaultScanTile;
    const bool timing = stages != nullptr && opt.stage_timing;
    int ev = 0;
    auto mark = [&](double StageTimes::*field) {
        if (!timing) return;
        hipEvent_t a = events_[ev & 1], b = events_[(ev + 1) & 1];
        check(hipEventRecord(b, stream), "hipEventRecord");
        check(hipEventSynchronize(b), "hipEventSynchronize");
        float ms = 0.0f;
        check(hipEventElapsedTime(&ms, a, b), "hipEventElapsedTime");
        stages->*field += (double)ms;
        ++ev;
    };
    if (timing) check(hipEventRecord(events_[0], stream), "hipEventRecord");

    wf_init_kernel<<<blocks_for(P), kThreads, 0, stream>>>(scene_, pool_, active_[0], sample, seed);
    check(hipGetLastError(), "wavefront init");
    mark(&StageTimes::other);
    // Algorithmic device-memory bytes per item of each stage (PathPool / ShadowQueue layouts):
    //   init      76 per pixel: ray, throughput, radiance (64) + key 8 + active index 4
    //   extend    44 per active path: index 4, ray 32, hit 8
    //   classify  104 per active path (index 4, hit 8, state 72, radiance 16, alive flag 4)
    //             + 36 per path queued for shading (surface 32, queue entry 4)
    //   shade     144 per queue entry (entry 4, surface 32, state 56 in, 48 out, alive flag 4)
    //             + 48 per shadow ray pushed (Lambert)
    //   shadow    48 per shadow ray (origin, direction, contribution); the radiance
    //             read-modify-write of unoccluded rays is not counted
    //   compact   24 per active path (mark: index 4, alive 4, flag 4; reduce: flag 4; scatter:
    //             flag 4, index 4) + 4 per surviving path
    //   finish    56 per pixel: radiance 16, film read-modify-write 40
    if (timing) {
        stages->items[7] += P;
        stages->launches[7] += 1;
        stages->bytes[7] += 76LL * P;
    }

    uint32_t n_active = P;
    int cur = 0, bounce = 0;
    const int max_bounces = max_depth + 2;
    while (n_active > 0 && bounce < max_bounces) {
        zero_words_kernel<<<1, 64, 0, stream>>>(counters_, kCntWords);
        wf_extend_kernel<<<extend_grid_, kTraceThreads, lds_bytes, stream>>>(scene_, pool_, active_[cur], n_active, n_lds,
                                                                               n_lds_prims);
        check(hipGetLastError(), "wavefront extend");
        mark(&StageTimes::extend);
        const int push_blocks = (int)((n_active + kPushThreads - 1) / kPushThreads);
        wf_classify_kernel<<<push_blocks, kPushThreads, 0, stream>>>(scene_, pool_, active_[cur], n_active, alive_,
                                                                     queues_[0], queues_[1], queues_[2], counters_, cap,
                                                                     max_depth, ctr);
        check(hipGetLastError(), "wavefront classify");
        mark(&StageTimes::classify);
        wf_shade_lambert_kernel<<<push_blocks, kPushThreads, 0, stream>>>(scene_, pool_, sq_, queues_[0], alive_,
                                                                          counters_, cap);
        check(hipGetLastError(), "wavefront shade lambert");
        mark(&StageTimes::shade_lambert);
        wf_shade_specular_kernel<kMetal><<<blocks_for(n_active), kThreads, 0, stream>>>(scene_, pool_, queues_[1],
                                                                                         alive_, counters_, cap);
        check(hipGetLastError(), "wavefront shade metal");
        mark(&StageTimes::shade_metal);
        wf_shade_specular_kernel<kDielectric><<<blocks_for(n_active), kThreads, 0, stream>>>(scene_, pool_, queues_[2],
                                                                                              alive_, counters_, cap);
        check(hipGetLastError(), "wavefront shade dielectric");
        mark(&StageTimes::shade_dielectric);
        wf_shadow_kernel<<<shadow_grid_, kTraceThreads, lds_bytes, stream>>>(scene_, pool_, sq_, counters_, cap, n_lds,
                                                                              n_lds_prims);
        check(hipGetLastError(), "wavefront shadow");
        mark(&StageTimes::shadow);
        wf_mark_kernel<<<blocks_for(n_active), kThreads, 0, stream>>>(active_[cur], n_active, alive_, flags_);
        check(compact_u32(active_[cur], flags_, active_[cur ^ 1], counters_ + kCntActive, (int)n_active, ws_, stream,
                          scan_tile),
              "wavefront compact");
        uint32_t* rec = (trace != nullptr && bounce < kMaxTraceBounces) ? trace_dev_ + bounce * 8 : nullptr;
        wf_update_kernel<<<1, 1, 0, stream>>>(counters_, ctr, n_active, cap, rec);
        check(hipGetLastError(), "wavefront update");
        uint32_t next = 0;
        check(hipMemcpyAsync(&next, counters_ + kCntActive, 4, hipMemcpyDeviceToHost, stream), "copy active count");
        check(hipStreamSynchronize(stream), "wavefront bounce sync");
        mark(&StageTimes::compact);
        if (timing) {
            uint32_t q[4];
            check(hipMemcpy(q, counters_, sizeof(q), hipMemcpyDeviceToHost), "copy queue counts");
            const long long qL = std::min(q[0], cap), qM = std::min(q[1], cap), qD = std::min(q[2], cap),
                            qS = std::min(q[3], cap), na = n_active;
            const long long items[8] = {na, na, qL, qM, qD, qS, na, 0};
            const long long launches[8] = {2, 1, 1, 1, 1, 1, 5, 0};
            const long long bytes[8] = {44 * na, 104 * na + 36 * (qL + qM + qD), 144 * qL + 48 * qS, 144 * qM,
                                        144 * qD, 48 * qS, 24 * na + 4LL * next, 0};
            for (int k = 0; k < 8; ++k) {
                stages->items[k] += items[k];
                stages->launches[k] += launches[k];
                stages->bytes[k] += bytes[k];
            }
        }
        if (trace != nullptr) {
            BounceTrace bt{};
            uint32_t r[8] = {0, 0, 0, 0, 0, 0, 0, 0};
            check(hipMemcpy(r, trace_dev_ + bounce * 8, 32, hipMemcpyDeviceToHost), "copy trace");
            bt.active = r[0];
            bt.lambert = r[1];
            bt.metal = r[2];
            bt.dielectric = r[3];
            bt.shadow = r[4];
            bt.next_active = r[5];
            std::vector<uint32_t> list(next);
            if (next) check(hipMemcpy(list.data(), active_[cur ^ 1], (size_t)next * 4, hipMemcpyDeviceToHost), "copy list");
            uint64_t h = 14695981039346656037ULL;
            for (uint32_t v : list)
                for (int b = 0; b < 4; ++b) {
                    h ^= (v >> (8 * b)) & 0xffu;
                    h *= 1099511628211ULL;
                }
            bt.list_hash = h;
            trace->push_back(bt);
        }
        n_active = next;
        cur ^= 1;
        ++bounce;
    }
    if (bounces) *bounces += bounce;
    wf_finish_kernel<<<blocks_for(P), kThreads, 0, stream>>>(film, pool_, pixels_);
    check(hipGetLastError(), "wavefront finish");
    mark(&StageTimes::other);
    if (timing) {
        stages->items[7] += P;
        stages->launches[7] += 1;
        stages->bytes[7] += 56LL * P;
    }
}

}  // namespace wf::gpu
