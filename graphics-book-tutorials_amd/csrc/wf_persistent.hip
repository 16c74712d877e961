// Persistent batched wavefront engine; the design is described in wf_persistent.h.
// Per iteration (all counts live in the device control block PState):
//   plan            1 thread: fold the previous iteration's results, size this one
//   extend_classify regenerate freed slots, traverse every live path, classify, push queues
//   shade           the three material queues as one index space (a wave sees one material
//                   except at the two boundaries); Lambert pushes shadow rays
//   shadow          any-hit traversal of the shadow queue, add unoccluded contributions
//   retire          reduce / scan / scatter: survivors -> next active list (stable),
//                   finished -> free-slot stack and their radiance -> the job's sample buffer
// The stage semantics and the order of L accumulation are those of the plain wavefront
// (src/gpu/wavefront.cu:88-270), so every path computes the same bits as in the megakernel.
#include <algorithm>
#include <chrono>
#include <cstring>

#include "wf_persistent.h"

namespace wf::gpu {

namespace {

constexpr int kMaxJobs = 4096;
constexpr int kRetireThreads = 256;
constexpr int kDefaultPool = 1 << 22;
constexpr int kDefaultBlock = 256;
constexpr int kDefaultGraphIters = 16;
constexpr int kDefaultLdsNodes = 1024;

struct PCtx {
    PState* st;
    JobDev* jobs;
    SceneView* scenes;
    LdsPlan* plan;
    PathPool P;
    ShadowQueue SQ;
    uint2* wk;          // (job, local work index) per slot
    uint32_t* alive;
    uint32_t* A[2];     // active lists (ping-pong by iteration parity)
    uint32_t* F;        // free-slot stack
    uint32_t* Q[3];     // material queues
    uint32_t* partial;  // retire tile partials
    DevCounters* ctr;
    uint32_t cap;
    int n_lds;          // single mode: nodes staged in LDS
    SceneView scene0;   // single mode: the scene
    JobDev job0;        // single mode: the job
};

__device__ __forceinline__ int find_job(const JobDev* jobs, uint32_t n_jobs, unsigned long long w) {
    int lo = 0, hi = (int)n_jobs - 1;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (jobs[mid].work_end > w) hi = mid;
        else lo = mid + 1;
    }
    return lo;
}

template <bool SINGLE>
__device__ __forceinline__ void stage_lds(const PCtx& c, Node* lds, LdsPlan* sp) {
    if (SINGLE) {
        stage_nodes(lds, c.scene0.nodes, c.n_lds);
    } else {
        if (threadIdx.x == 0) *sp = *c.plan;
        __syncthreads();
        for (int sc = 0; sc < sp->n; ++sc) {
            const Node* src = c.scenes[sc].nodes;
            for (int i = threadIdx.x; i < sp->cnt[sc]; i += blockDim.x) lds[sp->off[sc] + i] = src[i];
        }
        __syncthreads();
    }
}

template <bool SINGLE>
__device__ __forceinline__ LdsNodes node_src(const PCtx& c, const Node* lds, const LdsPlan* sp, int sc) {
    if (SINGLE) return LdsNodes{lds, c.scene0.nodes, c.n_lds};
    return LdsNodes{lds + sp->off[sc], c.scenes[sc].nodes, sp->cnt[sc]};
}

__global__ void pw_init_kernel(PState* st, uint32_t* F, uint32_t cap) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < cap) F[i] = i;
    if (i == 0) {
        PState z{};
        z.n_free = cap;
        z.min_live_work = ~0ull;
        *st = z;
    }
}

__global__ void pw_admit_kernel(PState* st, JobDev* jobs, JobDev job, uint32_t index) {
    jobs[index] = job;
    st->n_jobs = index + 1;
    st->total_work = job.work_end;
    st->done = 0;
}

__global__ void pw_plan_kernel(PState* st, DevCounters* ctr) {
    if (st->have_prev) {
        const uint32_t act = st->it_act, a = st->n_alive;
        ctr->extension_rays += act;
        ctr->shadow_rays += st->q[3];
        for (int k = 0; k < 4; ++k)
            if (st->q[k] > ctr->high_water[k]) ctr->high_water[k] = st->q[k];
        if (act > ctr->high_water[4]) ctr->high_water[4] = act;
        ctr->overflows += st->overflow;
        st->n_live = a;
        st->n_free = st->it_free_after + (act - a);
    }
    st->have_prev = 1;
    const unsigned long long rem = st->total_work - st->next_work;
    const uint32_t n_new = (uint32_t)((unsigned long long)st->n_free < rem ? (unsigned long long)st->n_free : rem);
    st->it_live_old = st->n_live;
    st->it_new = n_new;
    st->it_act = st->n_live + n_new;
    st->it_free_after = st->n_free - n_new;
    st->it_work_base = st->next_work;
    st->next_work += n_new;
    st->q[0] = st->q[1] = st->q[2] = st->q[3] = 0;
    st->n_alive = 0;
    st->overflow = 0;
    st->min_live_work = ~0ull;
    if (st->it_act > 0) st->iterations += 1;
    st->done = st->it_act == 0 ? 1u : 0u;
}

template <bool SINGLE>
__global__ __launch_bounds__(512) void pw_extend_classify_kernel(PCtx c, int parity) {
    extern __shared__ Node lds[];
    __shared__ LdsPlan sp;
    stage_lds<SINGLE>(c, lds, &sp);
    PState* st = c.st;
    const uint32_t n_act = st->it_act, n_old = st->it_live_old, free_after = st->it_free_after;
    const unsigned long long wbase = st->it_work_base;
    const uint32_t n_jobs = st->n_jobs;
    uint32_t* A = parity ? c.A[1] : c.A[0];
    uint32_t inv = 0;
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t b0 = blockIdx.x * blockDim.x; b0 < n_act; b0 += stride) {
        const uint32_t idx = b0 + threadIdx.x;
        const bool valid = idx < n_act;
        uint32_t slot = 0;
        int mat = -1;
        if (valid) {
            int jid = 0;
            PathState ps;
            if (idx >= n_old) {
                const uint32_t k = idx - n_old;
                slot = c.F[free_after + k];
                A[idx] = slot;
                const unsigned long long w = wbase + k;
                if (!SINGLE) jid = find_job(c.jobs, n_jobs, w);
                const JobDev& job = SINGLE ? c.job0 : c.jobs[jid];
                const SceneView& s = SINGLE ? c.scene0 : c.scenes[job.scene];
                const uint32_t local = (uint32_t)(w - job.work_begin);
                const uint32_t WH = (uint32_t)(job.width * job.height);
                const uint32_t sample = job.first_sample + local / WH;
                const uint32_t pix = local % WH;
                ps = init_path(s, (int)(pix % (uint32_t)job.width), (int)(pix / (uint32_t)job.width), sample, job.seed);
                store_scatter(c.P, slot, ps);
                c.P.key[slot] = make_uint2(ps.key.k0, ps.key.k1);
                c.wk[slot] = make_uint2((uint32_t)jid, local);
            } else {
                slot = A[idx];
                ps = load_path(c.P, slot, true);
                if (!SINGLE) jid = (int)c.wk[slot].x;
            }
            const JobDev& job = SINGLE ? c.job0 : c.jobs[jid];
            const SceneView& s = SINGLE ? c.scene0 : c.scenes[job.scene];
            Hit h;
            intersect_with(s, node_src<SINGLE>(c, lds, &sp, job.scene), ps.ray, &h, false);
            if (!hit_is_valid(s, h)) {
                inv += 1;
                h.prim = -1;
            }
            Surface f;
            mat = classify(s, ps, h, job.max_depth, &f);
            c.P.L[slot] = make_float4(ps.L.x, ps.L.y, ps.L.z, 0.0f);
            if (mat >= 0) store_surface(c.P, slot, f);
            c.alive[slot] = mat >= 0 ? 1u : 0u;
        }
#pragma unroll
        for (int q = 0; q < 3; ++q) {
            const bool mine = valid && mat == q;
            const uint32_t e = wave_push(c.Q[q], &st->q[q], c.cap, slot, mine, &st->overflow);
            if (mine && e == 0xFFFFFFFFu) c.alive[slot] = 0u;
        }
    }
    wave_add_u32(&c.ctr->invalid_hits, inv);
}

template <bool SINGLE>
__global__ __launch_bounds__(512) void pw_shade_kernel(PCtx c) {
    PState* st = c.st;
    const uint32_t nL = st->q[0], nM = st->q[1], nD = st->q[2];
    const uint32_t total = nL + nM + nD;
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t b0 = blockIdx.x * blockDim.x; b0 < total; b0 += stride) {
        const uint32_t idx = b0 + threadIdx.x;
        int has = 0;
        Ray sh;
        Vec3 contrib;
        uint32_t slot = 0;
        int sc = 0;
        if (idx < total) {
            int q;
            uint32_t qi;
            if (idx < nL) { q = 0; qi = idx; }
            else if (idx < nL + nM) { q = 1; qi = idx - nL; }
            else { q = 2; qi = idx - nL - nM; }
            const uint32_t* Qp = q == 0 ? c.Q[0] : (q == 1 ? c.Q[1] : c.Q[2]);
            slot = Qp[qi];
            const int jid = SINGLE ? 0 : (int)c.wk[slot].x;
            const JobDev& job = SINGLE ? c.job0 : c.jobs[jid];
            const SceneView& s = SINGLE ? c.scene0 : c.scenes[job.scene];
            sc = job.scene;
            const Surface f = load_surface(c.P, slot);
            PathState ps = load_path(c.P, slot, false);
            bool live;
            if (q == 0) live = shade_lambert(s, ps, f, &sh, &contrib, &has);
            else if (q == 1) live = shade_metal(s, ps, f);
            else live = shade_dielectric(s, ps, f);
            store_scatter(c.P, slot, ps);
            c.alive[slot] = live ? 1u : 0u;
        }
        const uint32_t e = wave_reserve(&st->q[3], has != 0);
        if (has) {
            c.SQ.o[e] = make_float4(sh.o.x, sh.o.y, sh.o.z, sh.tmax);
            c.SQ.d[e] = make_float4(sh.d.x, sh.d.y, sh.d.z, __int_as_float((int)slot));
            c.SQ.c[e] = make_float4(contrib.x, contrib.y, contrib.z, __int_as_float(sc));
        }
    }
}

template <bool SINGLE>
__global__ __launch_bounds__(512) void pw_shadow_kernel(PCtx c) {
    extern __shared__ Node lds[];
    __shared__ LdsPlan sp;
    stage_lds<SINGLE>(c, lds, &sp);
    const uint32_t n = c.st->q[3];
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t idx = blockIdx.x * blockDim.x + threadIdx.x; idx < n; idx += stride) {
        const float4 o = c.SQ.o[idx], d = c.SQ.d[idx], cc = c.SQ.c[idx];
        const int sc = SINGLE ? 0 : __float_as_int(cc.w);
        const SceneView& s = SINGLE ? c.scene0 : c.scenes[sc];
        Ray r;
        r.o = Vec3{o.x, o.y, o.z};
        r.d = Vec3{d.x, d.y, d.z};
        r.tmin = kRayEps;
        r.tmax = o.w;
        Hit h;
        if (!intersect_with(s, node_src<SINGLE>(c, lds, &sp, sc), r, &h, true)) {
            const uint32_t slot = (uint32_t)__float_as_int(d.w);
            float4 l = c.P.L[slot];
            l.x += cc.x;
            l.y += cc.y;
            l.z += cc.z;
            c.P.L[slot] = l;
        }
    }
}

// Block-wide exclusive scan of per-thread counts (256 threads).
__device__ __forceinline__ uint32_t retire_block_scan(uint32_t t, uint32_t* total) {
    __shared__ uint32_t wt[kRetireThreads / 64];
    const int lane = (int)lane_id(), w = threadIdx.x >> 6;
    uint32_t x = t;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t y = __shfl_up(x, o);
        if (lane >= o) x += y;
    }
    if (lane == 63) wt[w] = x;
    __syncthreads();
    uint32_t off = 0, tot = 0;
#pragma unroll
    for (int k = 0; k < kRetireThreads / 64; ++k) {
        const uint32_t v = wt[k];
        if (k < w) off += v;
        tot += v;
    }
    __syncthreads();
    *total = tot;
    return off + x - t;
}

template <int ITEMS>
__global__ __launch_bounds__(kRetireThreads) void pw_retire_reduce_kernel(PCtx c, int parity) {
    const uint32_t n = c.st->it_act;
    const uint32_t tile = kRetireThreads * ITEMS;
    const uint32_t n_tiles = (n + tile - 1) / tile;
    const uint32_t* A = parity ? c.A[1] : c.A[0];
    for (uint32_t t = blockIdx.x; t < n_tiles; t += gridDim.x) {
        const uint32_t i0 = t * tile + threadIdx.x * ITEMS;
        uint32_t cnt = 0;
#pragma unroll
        for (int k = 0; k < ITEMS; ++k)
            if (i0 + k < n) cnt += c.alive[A[i0 + k]];
        uint32_t total;
        retire_block_scan(cnt, &total);
        if (threadIdx.x == 0) c.partial[t] = total;
    }
}

template <int ITEMS>
__global__ __launch_bounds__(1024) void pw_retire_scan_kernel(PCtx c) {
    __shared__ uint32_t wtot[16];
    __shared__ uint32_t carry;
    const uint32_t n = c.st->it_act;
    const uint32_t tile = kRetireThreads * ITEMS;
    const uint32_t nb = (n + tile - 1) / tile;
    if (threadIdx.x == 0) carry = 0u;
    __syncthreads();
    const int w = threadIdx.x >> 6, lane = (int)lane_id();
    for (uint32_t base = 0; base < nb; base += 1024) {
        const uint32_t i = base + threadIdx.x;
        const uint32_t v = i < nb ? c.partial[i] : 0u;
        uint32_t x = v;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const uint32_t y = __shfl_up(x, o);
            if (lane >= o) x += y;
        }
        if (lane == 63) wtot[w] = x;
        __syncthreads();
        uint32_t off = 0, tot = 0;
        for (int k = 0; k < 16; ++k) {
            const uint32_t y = wtot[k];
            if (k < w) off += y;
            tot += y;
        }
        const uint32_t cv = carry;
        if (i < nb) c.partial[i] = cv + off + x - v;
        __syncthreads();
        if (threadIdx.x == 0) carry = cv + tot;
        __syncthreads();
    }
    if (threadIdx.x == 0) c.st->n_alive = carry;
}

template <bool SINGLE, int ITEMS>
__global__ __launch_bounds__(kRetireThreads) void pw_retire_scatter_kernel(PCtx c, int parity) {
    PState* st = c.st;
    const uint32_t n = st->it_act, free_after = st->it_free_after;
    const uint32_t tile = kRetireThreads * ITEMS;
    const uint32_t n_tiles = (n + tile - 1) / tile;
    const uint32_t* A = parity ? c.A[1] : c.A[0];
    uint32_t* An = parity ? c.A[0] : c.A[1];
    for (uint32_t t = blockIdx.x; t < n_tiles; t += gridDim.x) {
        const uint32_t i0 = t * tile + threadIdx.x * ITEMS;
        uint32_t slot[ITEMS], fl[ITEMS];
        uint32_t cnt = 0;
#pragma unroll
        for (int k = 0; k < ITEMS; ++k) {
            slot[k] = i0 + k < n ? A[i0 + k] : 0u;
            fl[k] = i0 + k < n ? c.alive[slot[k]] : 0u;
            cnt += fl[k];
        }
        uint32_t total;
        uint32_t pos = retire_block_scan(cnt, &total) + c.partial[t];
        unsigned long long min_work = ~0ull;
#pragma unroll
        for (int k = 0; k < ITEMS; ++k) {
            const uint32_t idx = i0 + k;
            if (idx >= n) continue;
            const uint2 wk = c.wk[slot[k]];
            const JobDev& job = SINGLE ? c.job0 : c.jobs[wk.x];
            if (fl[k]) {
                An[pos++] = slot[k];
                const unsigned long long gw = job.work_begin + wk.y;
                min_work = gw < min_work ? gw : min_work;
            } else {
                c.F[free_after + (idx - pos)] = slot[k];
                job.sbuf[wk.y] = c.P.L[slot[k]];
            }
        }
        // Wave minimum, then one atomic per wave.
        for (int o = 32; o > 0; o >>= 1) {
            const unsigned long long y = __shfl_xor(min_work, o);
            min_work = y < min_work ? y : min_work;
        }
        if (lane_id() == 0 && min_work != ~0ull) atomicMin(&st->min_live_work, min_work);
    }
}

// Adds a finished job's samples to its film, per pixel in sample order.
__global__ void pw_resolve_kernel(JobDev job) {
    const int WH = job.width * job.height;
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= WH) return;
    float rgb[3] = {job.film.rgb[3 * p], job.film.rgb[3 * p + 1], job.film.rgb[3 * p + 2]};
    float lum = job.film.lum[p], lum2 = job.film.lum2[p];
    for (uint32_t s = 0; s < job.spp; ++s) {
        const float4 l = job.sbuf[(size_t)s * WH + p];
        film_add3(rgb, &lum, &lum2, Vec3{l.x, l.y, l.z});
    }
    job.film.rgb[3 * p] = rgb[0];
    job.film.rgb[3 * p + 1] = rgb[1];
    job.film.rgb[3 * p + 2] = rgb[2];
    job.film.lum[p] = lum;
    job.film.lum2[p] = lum2;
}

int grid_for(const void* fn, int block, size_t lds, int cus) {
    int per = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, fn, block, lds) != hipSuccess || per < 1) per = 1;
    return per * cus;
}

}  // namespace

struct PersistentEngine::Impl {
    hipStream_t stream = nullptr;
    DevCounters* ctr = nullptr;
    PersistentTuning tune;
    int block = kDefaultBlock, items = 8, graph_iters = kDefaultGraphIters, lds_nodes = kDefaultLdsNodes, cus = 0;
    std::vector<void*> pool_allocs;
    PState* st = nullptr;
    PState* st_host = nullptr;   // pinned mirror
    JobDev* jobs = nullptr;
    SceneView* scenes = nullptr;
    LdsPlan* plan = nullptr;
    PCtx ctx{};
    struct HostJob {
        JobDev dev;
        bool finished = false;
    };
    std::vector<HostJob> hjobs;
    std::vector<SceneView> hscenes;
    std::vector<const SceneView*> scene_keys;
    unsigned long long total_work = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    bool graph_single = false;
    int graph_scenes = -1;
    uint64_t sbuf_bytes = 0;

    void free_graph() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        exec = nullptr;
        graph = nullptr;
        graph_scenes = -1;
    }
    void free_pool() {
        for (void* p : pool_allocs) (void)hipFree(p);
        pool_allocs.clear();
    }
    ~Impl() {
        free_graph();
        for (HostJob& j : hjobs)
            if (j.dev.sbuf) (void)hipFree(j.dev.sbuf);
        free_pool();
        if (st_host) (void)hipHostFree(st_host);
    }
};

PersistentEngine::PersistentEngine() : m_(std::make_unique<Impl>()) {}
PersistentEngine::~PersistentEngine() = default;

uint64_t PersistentEngine::bytes() const { return pool_bytes_ + m_->sbuf_bytes; }

uint64_t PersistentEngine::job_buffer_bytes(const PersistentJob& job) const {
    return (uint64_t)job.scene->width * job.scene->height * (uint64_t)job.spp * sizeof(float4);
}

void PersistentEngine::ensure_pool(const PersistentTuning& t) {
    Impl& m = *m_;
    if (m.cus == 0) {
        int dev = 0;
        check(hipGetDevice(&dev), "hipGetDevice");
        check(hipDeviceGetAttribute(&m.cus, hipDeviceAttributeMultiprocessorCount, dev), "hipDeviceGetAttribute");
    }
    int cap = t.pool_paths > 0 ? t.pool_paths : kDefaultPool;
    cap = (cap + 1023) & ~1023;
    m.block = t.block_size > 0 ? t.block_size : kDefaultBlock;
    if (m.block != 128 && m.block != 256 && m.block != 512) throw std::runtime_error("persistent: block size must be 128, 256 or 512");
    const int tile = t.scan_tile > 0 ? t.scan_tile : 2048;
    m.items = tile / kRetireThreads;
    if (m.items != 4 && m.items != 8 && m.items != 16) throw std::runtime_error("persistent: scan tile must be 1024, 2048 or 4096");
    m.graph_iters = t.graph_iters > 0 ? ((t.graph_iters + 1) & ~1) : kDefaultGraphIters;
    m.lds_nodes = t.lds_nodes >= 0 ? t.lds_nodes : kDefaultLdsNodes;
    m.tune = t;
    if (cap == cap_ && !m.pool_allocs.empty()) return;
    m.free_graph();
    m.free_pool();
    pool_bytes_ = 0;
    cap_ = cap;
    auto alloc = [&](auto** p, size_t bytes) {
        void* q = nullptr;
        check(hipMalloc(&q, bytes), "hipMalloc persistent pool");
        m.pool_allocs.push_back(q);
        *p = (std::remove_reference_t<decltype(**p)>*)q;
        pool_bytes_ += bytes;
    };
    const size_t P = (size_t)cap;
    PCtx& c = m.ctx;
    c = PCtx{};
    alloc(&c.P.ro, P * 16);
    alloc(&c.P.rd, P * 16);
    alloc(&c.P.T, P * 16);
    alloc(&c.P.L, P * 16);
    alloc(&c.P.key, P * 8);
    alloc(&c.P.sp, P * 16);
    alloc(&c.P.sn, P * 16);
    alloc(&c.SQ.o, P * 16);
    alloc(&c.SQ.d, P * 16);
    alloc(&c.SQ.c, P * 16);
    alloc(&c.wk, P * 8);
    alloc(&c.alive, P * 4);
    alloc(&c.A[0], P * 4);
    alloc(&c.A[1], P * 4);
    alloc(&c.F, P * 4);
    for (int q = 0; q < 3; ++q) alloc(&c.Q[q], P * 4);
    alloc(&c.partial, (P / 1024 + 2) * 4);
    alloc(&m.st, sizeof(PState));
    alloc(&m.jobs, sizeof(JobDev) * kMaxJobs);
    alloc(&m.scenes, sizeof(SceneView) * kMaxScenes);
    alloc(&m.plan, sizeof(LdsPlan));
    if (!m.st_host) check(hipHostMalloc(&m.st_host, sizeof(PState), hipHostMallocDefault), "hipHostMalloc");
    c.st = m.st;
    c.jobs = m.jobs;
    c.scenes = m.scenes;
    c.plan = m.plan;
    c.cap = (uint32_t)cap;
}

void PersistentEngine::open(const PersistentTuning& t, DevCounters* ctr, hipStream_t stream) {
    Impl& m = *m_;
    if (stream == nullptr) throw std::runtime_error("persistent: a non-null stream is required (graph capture)");
    m.stream = stream;
    m.ctr = ctr;
    ensure_pool(t);
    m.ctx.ctr = ctr;
    reset_jobs();
}

void PersistentEngine::reset_jobs() {
    Impl& m = *m_;
    for (Impl::HostJob& j : m.hjobs)
        if (j.dev.sbuf) (void)hipFree(j.dev.sbuf);
    m.hjobs.clear();
    m.hscenes.clear();
    m.scene_keys.clear();
    m.sbuf_bytes = 0;
    m.total_work = 0;
    m.free_graph();
    pw_init_kernel<<<(cap_ + 255) / 256, 256, 0, m.stream>>>(m.st, m.ctx.F, (uint32_t)cap_);
    check(hipGetLastError(), "persistent init");
    LdsPlan empty{};
    check(hipMemcpyAsync(m.plan, &empty, sizeof(empty), hipMemcpyHostToDevice, m.stream), "copy lds plan");
}

int PersistentEngine::admit(const PersistentJob& job) {
    Impl& m = *m_;
    if ((int)m.hjobs.size() >= kMaxJobs) throw std::runtime_error("persistent: job table full");
    if (job.spp < 1 || job.scene == nullptr) throw std::runtime_error("persistent: invalid job");
    if (job.max_depth < 0 || job.max_depth > 0xffff) throw std::runtime_error("persistent: max_depth out of range");
    const uint64_t WH = (uint64_t)job.scene->width * job.scene->height;
    const uint64_t work = WH * (uint64_t)job.spp;
    if (work >= (1ull << 32)) throw std::runtime_error("persistent: job has 2^32 or more paths");
    int sc = -1;
    for (size_t k = 0; k < m.scene_keys.size(); ++k)
        if (m.scene_keys[k] == job.scene) sc = (int)k;
    if (sc < 0) {
        if ((int)m.scene_keys.size() >= kMaxScenes) throw std::runtime_error("persistent: too many distinct scenes");
        sc = (int)m.scene_keys.size();
        m.scene_keys.push_back(job.scene);
        m.hscenes.push_back(*job.scene);
        check(hipMemcpyAsync(m.scenes + sc, job.scene, sizeof(SceneView), hipMemcpyHostToDevice, m.stream), "copy scene");
        LdsPlan plan{};
        plan.n = (int)m.hscenes.size();
        const int share = m.lds_nodes / plan.n;
        int off = 0;
        for (int k = 0; k < plan.n; ++k) {
            plan.off[k] = off;
            plan.cnt[k] = std::min(share, m.hscenes[k].n_nodes);
            off += plan.cnt[k];
        }
        check(hipMemcpyAsync(m.plan, &plan, sizeof(plan), hipMemcpyHostToDevice, m.stream), "copy lds plan");
        m.free_graph();   // the single/multi choice depends on the scene count
    }
    JobDev d{};
    d.scene = sc;
    d.width = job.scene->width;
    d.height = job.scene->height;
    d.max_depth = job.max_depth;
    d.first_sample = (uint32_t)job.first_sample;
    d.spp = (uint32_t)job.spp;
    d.seed = job.seed;
    d.work_begin = m.total_work;
    d.work_end = m.total_work + work;
    d.film = job.film;
    check(hipMalloc(&d.sbuf, (size_t)work * sizeof(float4)), "hipMalloc sample buffer");
    m.sbuf_bytes += (uint64_t)work * sizeof(float4);
    m.total_work = d.work_end;
    const uint32_t index = (uint32_t)m.hjobs.size();
    pw_admit_kernel<<<1, 1, 0, m.stream>>>(m.st, m.jobs, d, index);
    check(hipGetLastError(), "persistent admit");
    m.hjobs.push_back(Impl::HostJob{d, false});
    if (m.graph && m.graph_single) m.free_graph();   // single-job graph bakes the job in
    return (int)index;
}

struct PersistentEngine::LaunchPlan {
    PCtx c;
    bool single;
    size_t lds_bytes;
    int gx, gs, gw, gr;
};

PersistentEngine::LaunchPlan PersistentEngine::make_plan(bool single) {
    Impl& m = *m_;
    LaunchPlan L{};
    L.single = single;
    L.c = m.ctx;
    L.c.n_lds = 0;
    if (single) {
        L.c.scene0 = m.hscenes[0];
        L.c.job0 = m.hjobs[0].dev;
        L.c.n_lds = std::min(m.lds_nodes, L.c.scene0.n_nodes);
    }
    L.lds_bytes = (size_t)(single ? L.c.n_lds : m.lds_nodes) * sizeof(Node);
    const void* fx = single ? (const void*)pw_extend_classify_kernel<true> : (const void*)pw_extend_classify_kernel<false>;
    const void* fs = single ? (const void*)pw_shade_kernel<true> : (const void*)pw_shade_kernel<false>;
    const void* fw = single ? (const void*)pw_shadow_kernel<true> : (const void*)pw_shadow_kernel<false>;
    L.gx = grid_for(fx, m.block, L.lds_bytes, m.cus);
    L.gs = grid_for(fs, m.block, 0, m.cus);
    L.gw = grid_for(fw, m.block, L.lds_bytes, m.cus);
    L.gr = m.cus * 4;
    return L;
}

// One iteration; with `ev` (6 events) the stage boundaries are recorded for profiling:
// ev[0] plan ev[1] extend_classify ev[2] shade ev[3] shadow ev[4] retire ev[5].
void PersistentEngine::launch_iteration(const LaunchPlan& L, int par, hipEvent_t* ev) {
    Impl& m = *m_;
    hipStream_t st = m.stream;
    const PCtx& c = L.c;
    if (ev) check(hipEventRecord(ev[0], st), "hipEventRecord");
    pw_plan_kernel<<<1, 1, 0, st>>>(m.st, m.ctr);
    if (ev) check(hipEventRecord(ev[1], st), "hipEventRecord");
    if (L.single) pw_extend_classify_kernel<true><<<L.gx, m.block, L.lds_bytes, st>>>(c, par);
    else pw_extend_classify_kernel<false><<<L.gx, m.block, L.lds_bytes, st>>>(c, par);
    if (ev) check(hipEventRecord(ev[2], st), "hipEventRecord");
    if (L.single) pw_shade_kernel<true><<<L.gs, m.block, 0, st>>>(c);
    else pw_shade_kernel<false><<<L.gs, m.block, 0, st>>>(c);
    if (ev) check(hipEventRecord(ev[3], st), "hipEventRecord");
    if (L.single) pw_shadow_kernel<true><<<L.gw, m.block, L.lds_bytes, st>>>(c);
    else pw_shadow_kernel<false><<<L.gw, m.block, L.lds_bytes, st>>>(c);
    if (ev) check(hipEventRecord(ev[4], st), "hipEventRecord");
    switch (m.items) {
        case 4:
            pw_retire_reduce_kernel<4><<<L.gr, kRetireThreads, 0, st>>>(c, par);
            pw_retire_scan_kernel<4><<<1, 1024, 0, st>>>(c);
            if (L.single) pw_retire_scatter_kernel<true, 4><<<L.gr, kRetireThreads, 0, st>>>(c, par);
            else pw_retire_scatter_kernel<false, 4><<<L.gr, kRetireThreads, 0, st>>>(c, par);
            break;
        case 16:
            pw_retire_reduce_kernel<16><<<L.gr, kRetireThreads, 0, st>>>(c, par);
            pw_retire_scan_kernel<16><<<1, 1024, 0, st>>>(c);
            if (L.single) pw_retire_scatter_kernel<true, 16><<<L.gr, kRetireThreads, 0, st>>>(c, par);
            else pw_retire_scatter_kernel<false, 16><<<L.gr, kRetireThreads, 0, st>>>(c, par);
            break;
        default:
            pw_retire_reduce_kernel<8><<<L.gr, kRetireThreads, 0, st>>>(c, par);
            pw_retire_scan_kernel<8><<<1, 1024, 0, st>>>(c);
            if (L.single) pw_retire_scatter_kernel<true, 8><<<L.gr, kRetireThreads, 0, st>>>(c, par);
            else pw_retire_scatter_kernel<false, 8><<<L.gr, kRetireThreads, 0, st>>>(c, par);
            break;
    }
    if (ev) check(hipEventRecord(ev[5], st), "hipEventRecord");
}

void PersistentEngine::capture(bool single) {
    Impl& m = *m_;
    m.free_graph();
    const LaunchPlan L = make_plan(single);
    check(hipStreamBeginCapture(m.stream, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
    for (int it = 0; it < m.graph_iters; ++it) launch_iteration(L, it & 1, nullptr);
    hipError_t e = hipStreamEndCapture(m.stream, &m.graph);
    check(e, "hipStreamEndCapture");
    check(hipGetLastError(), "persistent capture");
    check(hipGraphInstantiate(&m.exec, m.graph, nullptr, nullptr, 0), "hipGraphInstantiate");
    m.graph_single = single;
    m.graph_scenes = (int)m.hscenes.size();
}

// Eager iterations with HIP events around each stage (synchronising once per iteration), until
// the admitted work is done. Items per stage come from the control block of each iteration.
void PersistentEngine::run_profiled(StageTimes* out) {
    Impl& m = *m_;
    if (m.hjobs.empty()) return;
    const bool single = m.hjobs.size() == 1 && m.hscenes.size() == 1;
    const LaunchPlan L = make_plan(single);
    hipEvent_t ev[6];
    for (auto& e : ev) check(hipEventCreate(&e), "hipEventCreate");
    for (int it = 0;; ++it) {
        launch_iteration(L, it & 1, ev);
        check(hipMemcpyAsync(m.st_host, m.st, sizeof(PState), hipMemcpyDeviceToHost, m.stream), "copy state");
        check(hipStreamSynchronize(m.stream), "profile sync");
        const PState& s = *m.st_host;
        float ms[5];
        for (int k = 0; k < 5; ++k) check(hipEventElapsedTime(&ms[k], ev[k], ev[k + 1]), "hipEventElapsedTime");
        if (s.done) break;
        out->other += ms[0];
        out->extend += ms[1];
        out->shade_lambert += ms[2];
        out->shadow += ms[3];
        out->compact += ms[4];
        out->items[0] += s.it_act;
        out->items[2] += (long long)s.q[0] + s.q[1] + s.q[2];
        out->items[5] += s.q[3];
        out->items[6] += s.it_act;
        out->items[7] += 1;
        out->launches[0] += 1;
        out->launches[2] += 1;
        out->launches[5] += 1;
        out->launches[6] += 3;
        out->launches[7] += 1;
    }
    for (auto& e : ev) (void)hipEventDestroy(e);
}

void PersistentEngine::run(int launches) {
    Impl& m = *m_;
    if (m.hjobs.empty()) return;
    const bool single = m.hjobs.size() == 1 && m.hscenes.size() == 1;
    if (!m.exec || m.graph_single != single || m.graph_scenes != (int)m.hscenes.size()) capture(single);
    for (int k = 0; k < launches; ++k) check(hipGraphLaunch(m.exec, m.stream), "hipGraphLaunch");
}

PersistentProgress PersistentEngine::progress() {
    Impl& m = *m_;
    check(hipMemcpyAsync(m.st_host, m.st, sizeof(PState), hipMemcpyDeviceToHost, m.stream), "copy progress");
    check(hipStreamSynchronize(m.stream), "persistent sync");
    PersistentProgress p;
    p.next_work = m.st_host->next_work;
    p.total_work = m.st_host->total_work;
    p.min_live_work = m.st_host->min_live_work;
    p.n_live = m.st_host->n_live;
    p.done = m.st_host->done;
    p.iterations = m.st_host->iterations;
    return p;
}

bool PersistentEngine::job_complete(int job, const PersistentProgress& pr) const {
    const JobDev& d = m_->hjobs[job].dev;
    if (pr.next_work < d.work_end) return false;
    if (pr.done) return true;
    return pr.min_live_work >= d.work_end;
}

void PersistentEngine::finish_job(int job) {
    Impl& m = *m_;
    Impl::HostJob& j = m.hjobs[job];
    if (j.finished) return;
    const int WH = j.dev.width * j.dev.height;
    pw_resolve_kernel<<<(WH + 255) / 256, 256, 0, m.stream>>>(j.dev);
    check(hipGetLastError(), "persistent resolve");
    check(hipStreamSynchronize(m.stream), "persistent resolve sync");
    check(hipFree(j.dev.sbuf), "hipFree sample buffer");
    m.sbuf_bytes -= (uint64_t)WH * j.dev.spp * sizeof(float4);
    j.dev.sbuf = nullptr;
    j.finished = true;
}

void PersistentEngine::render(const PersistentJob* jobs, int n_jobs, const PersistentTuning& t, DevCounters* ctr,
                              hipStream_t stream, PersistentResult* out, StageTimes* stages) {
    PersistentResult res;
    // Split jobs into consecutive sample chunks so that one chunk's sample buffers fit the budget.
    std::vector<PersistentJob> rest(jobs, jobs + n_jobs);
    uint64_t total = 0;
    for (PersistentJob& j : rest) {
        if (j.spp < 0 || j.scene == nullptr) throw std::runtime_error("persistent: invalid job");
        total += (uint64_t)j.scene->width * j.scene->height * (uint64_t)j.spp;
    }
    // No more slots than paths in the batch.
    PersistentTuning tt = t;
    const uint64_t want = tt.pool_paths > 0 ? (uint64_t)tt.pool_paths : (uint64_t)kDefaultPool;
    tt.pool_paths = (int)std::min<uint64_t>(want, std::max<uint64_t>(1024, (total + 1023) & ~1023ull));
    open(tt, ctr, stream);
    for (;;) {
        std::vector<PersistentJob> chunk;
        uint64_t budget = sbuf_budget_;
        for (PersistentJob& j : rest) {
            if (j.spp <= 0) continue;
            const uint64_t per_sample = (uint64_t)j.scene->width * j.scene->height * sizeof(float4);
            uint64_t n = budget / per_sample;
            if (n == 0 && chunk.empty()) n = 1;
            if (n == 0) break;
            PersistentJob c = j;
            c.spp = (int)std::min<uint64_t>(n, (uint64_t)j.spp);
            chunk.push_back(c);
            budget -= std::min<uint64_t>(budget, per_sample * (uint64_t)c.spp);
            j.first_sample += c.spp;
            j.spp -= c.spp;
        }
        if (chunk.empty()) break;
        if (res.chunks > 0) reset_jobs();
        for (const PersistentJob& c : chunk) admit(c);
        PersistentProgress pr;
        if (stages) {
            run_profiled(stages);
            pr = progress();
        } else {
            for (;;) {
                run(1);
                res.graph_launches += 1;
                pr = progress();
                if (pr.done) break;
            }
        }
        res.iterations += (int)pr.iterations;
        for (int k = 0; k < (int)m_->hjobs.size(); ++k) finish_job(k);
        res.chunks += 1;
    }
    reset_jobs();
    if (out) *out = res;
}

PersistentTuning autotune_persistent(const PersistentJob* jobs, int n_jobs, hipStream_t stream, int probe_spp,
                                     std::vector<std::pair<PersistentTuning, double>>* trials) {
    DevCounters* ctr = nullptr;
    check(hipMalloc(&ctr, sizeof(DevCounters)), "hipMalloc counters");
    std::vector<PersistentJob> probe(jobs, jobs + n_jobs);
    uint64_t work = 0;
    for (PersistentJob& j : probe) {
        j.spp = std::min(j.spp, probe_spp);
        work += (uint64_t)j.scene->width * j.scene->height * j.spp;
    }
    PersistentTuning best;
    double best_ms = 1e30;
    const int blocks[] = {256, 512};
    const int pools[] = {1 << 20, 1 << 21, 1 << 22};
    const int tiles[] = {2048, 4096};
    PersistentEngine eng;
    hipEvent_t a, b;
    check(hipEventCreate(&a), "hipEventCreate");
    check(hipEventCreate(&b), "hipEventCreate");
    // Probe runs add into the jobs' films; the caller resets them afterwards.
    for (int bs : blocks)
        for (int pool : pools) {
            if ((uint64_t)pool > 2 * work && pool != pools[0]) continue;
            for (int tile : tiles) {
                PersistentTuning t;
                t.block_size = bs;
                t.pool_paths = pool;
                t.scan_tile = tile;
                PersistentResult r;
                eng.render(probe.data(), n_jobs, t, ctr, stream, &r);   // warm-up (graph capture)
                check(hipEventRecord(a, stream), "hipEventRecord");
                eng.render(probe.data(), n_jobs, t, ctr, stream, &r);
                check(hipEventRecord(b, stream), "hipEventRecord");
                check(hipEventSynchronize(b), "hipEventSynchronize");
                float ms = 0.0f;
                check(hipEventElapsedTime(&ms, a, b), "hipEventElapsedTime");
                if (trials) trials->push_back({t, (double)ms});
                if (ms < best_ms) {
                    best_ms = ms;
                    best = t;
                }
            }
        }
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    (void)hipFree(ctr);
    return best;
}

}  // namespace wf::gpu
