// Persistent batched wavefront engine; the design is described in wf_persistent.h.
// Per iteration (all counts live in the device control block PState):
//   plan    1 thread: fold the previous iteration's results, size this one
//   extend  regenerate freed slots, traverse every live path, push it by hit material
//   shade   classify + shade the four queues; re-queue survivors, finish the others
//   shadow  any-hit shadow rays; finish paths whose last bounce left a shadow ray pending
// Stage semantics and the order of L accumulation are those of the plain wavefront
// (src/gpu/wavefront.cu:88-270): emission in classify, then the unoccluded NEE contribution
// before the next classify, so every path computes the megakernel's bits.
#include <algorithm>
#include <cstring>

#include "wf_persistent.h"

namespace wf::gpu {

namespace {

constexpr int kDefaultPool = 1 << 22;
constexpr int kDefaultBlock = 512;
constexpr int kDefaultGraphIters = 16;
constexpr int kDefaultLdsKiB = 64;

struct PCtx {
    PState* st;
    JobDev* jobs;
    uint32_t* job_done;   // finished paths per job
    SceneView* scenes;
    LdsPlan* plan;
    PathPool P;
    ShadowQueue SQ;
    uint2* wk;            // (job, local work index) per slot
    uint32_t* A[2];       // active lists (ping-pong by iteration parity)
    uint32_t* F;          // free-slot stack
    uint32_t* Q[4];       // terminal, Lambert, metal, dielectric queues
    DevCounters* ctr;
    uint32_t cap;
    LdsPlan plan0;        // single mode: the LDS layout
    SceneView scene0;     // single mode: the scene
    JobDev job0;          // single mode: the job
};

template <bool SINGLE>
__device__ __forceinline__ const SceneView& scene_of(const PCtx& c, int sc) {
    if constexpr (SINGLE) return c.scene0;
    else return c.scenes[sc];
}

template <bool SINGLE>
__device__ __forceinline__ const JobDev& job_of(const PCtx& c, uint32_t jid) {
    if constexpr (SINGLE) return c.job0;
    else return c.jobs[jid];
}

__device__ __forceinline__ int find_job(const JobDev* jobs, uint32_t n_jobs, unsigned long long w) {
    int lo = 0, hi = (int)n_jobs - 1;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (jobs[mid].work_end > w) hi = mid;
        else lo = mid + 1;
    }
    return lo;
}

// Pixel of position q in a pass, visiting the image in 8-row bands of 8x8 tiles (edge tiles
// smaller), so the 64 consecutive new paths of a wave start from one 8x8 pixel tile.
__device__ __forceinline__ void tile_pixel(uint32_t q, uint32_t W, uint32_t H, uint32_t* px, uint32_t* py) {
    const uint32_t b = q / (8u * W);
    const uint32_t r = q - b * 8u * W;
    const uint32_t hb = min(8u, H - 8u * b);
    const uint32_t j = r / (8u * hb);
    const uint32_t in = r - j * 8u * hb;
    const uint32_t wb = min(8u, W - 8u * j);
    *px = 8u * j + in % wb;
    *py = 8u * b + in / wb;
}

// Stages the LDS layout (BVH tops and primitive records); the whole block calls it.
template <bool SINGLE>
__device__ __forceinline__ void stage_lds(const PCtx& c, Node* lds, LdsPlan* sp) {
    if constexpr (SINGLE) {
        if (threadIdx.x == 0) *sp = c.plan0;
    } else {
        if (threadIdx.x == 0) *sp = *c.plan;
    }
    __syncthreads();
    float4* lprim = reinterpret_cast<float4*>(lds + sp->total_nodes);
    for (int sc = 0; sc < sp->n; ++sc) {
        const SceneView& s = scene_of<SINGLE>(c, sc);
        for (int i = threadIdx.x; i < sp->node_cnt[sc]; i += blockDim.x) lds[sp->node_off[sc] + i] = s.nodes[i];
        for (int i = threadIdx.x; i < sp->prim_cnt[sc]; i += blockDim.x)
            lprim[sp->prim_off[sc] + i] = lds_prim_record(s, i);
    }
    __syncthreads();
}

__device__ __forceinline__ LdsNodes src_of(const SceneView& s, const Node* lds, const LdsPlan* sp, int sc) {
    const float4* lprim = reinterpret_cast<const float4*>(lds + sp->total_nodes);
    return lds_nodes_of(s, lds + sp->node_off[sc], sp->node_cnt[sc],
                        sp->prim_cnt[sc] > 0 ? lprim + sp->prim_off[sc] : nullptr);
}

// Adds one per lane with pred to job_done[jid], one atomic per distinct job id in the wave.
__device__ __forceinline__ void count_done(uint32_t* job_done, bool pred, uint32_t jid) {
    uint64_t m = __ballot(pred);
    while (m) {
        const int leader = __ffsll((unsigned long long)m) - 1;
        const uint32_t j0 = __shfl(jid, leader);
        const uint64_t same = __ballot(pred && jid == j0);
        if ((int)lane_id() == leader) atomicAdd(&job_done[j0], (uint32_t)__popcll(same));
        m &= ~same;
    }
}

__global__ void pw_init_kernel(PState* st, uint32_t* F, uint32_t cap) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < cap) F[i] = i;
    if (i == 0) {
        PState z{};
        z.n_free = cap;
        *st = z;
    }
}

__global__ void pw_admit_kernel(PState* st, JobDev* jobs, uint32_t* job_done, JobDev job, uint32_t index) {
    jobs[index] = job;
    job_done[index] = 0u;
    st->n_jobs = index + 1;
    st->total_work = job.work_end;
    st->done = 0;
}

__global__ void pw_plan_kernel(PState* st, DevCounters* ctr) {
    if (st->have_prev) {
        const uint32_t act = st->it_act;
        ctr->extension_rays += act;
        ctr->shadow_rays += st->q[4];
        for (int k = 0; k < 4; ++k)
            if (st->q[k + 1] > ctr->high_water[k]) ctr->high_water[k] = st->q[k + 1];
        if (act > ctr->high_water[4]) ctr->high_water[4] = act;
        ctr->overflows += st->overflow;
        st->n_live = st->n_next;
        st->n_free = st->it_free_after + st->n_freed;
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
    for (int k = 0; k < 5; ++k) st->q[k] = 0;
    st->n_next = 0;
    st->n_freed = 0;
    st->overflow = 0;
    if (st->it_act > 0) st->iterations += 1;
    st->done = st->it_act == 0 ? 1u : 0u;
}

template <bool SINGLE>
__global__ __launch_bounds__(512) void pw_extend_kernel(PCtx c, int parity) {
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
        int cat = -1;
        if (valid) {
            uint32_t jid = 0;
            Ray r;
            if (idx >= n_old) {
                const uint32_t k = idx - n_old;
                slot = c.F[free_after + k];
                A[idx] = slot;
                const unsigned long long w = wbase + k;
                if constexpr (!SINGLE) jid = (uint32_t)find_job(c.jobs, n_jobs, w);
                const JobDev& job = job_of<SINGLE>(c, jid);
                const SceneView& s = scene_of<SINGLE>(c, job.scene);
                const uint32_t local = (uint32_t)(w - job.work_begin);
                const uint32_t WH = (uint32_t)(job.width * job.height);
                const uint32_t sample = job.first_sample + local / WH;
                uint32_t px, py;
                tile_pixel(local % WH, (uint32_t)job.width, (uint32_t)job.height, &px, &py);
                const PathState ps = init_path(s, (int)px, (int)py, sample, job.seed);
                store_new_path(c.P, slot, ps);   // ray, flags, T = 1, L = 0, key
                // Sample-buffer index (sample, pixel) of this path.
                c.wk[slot] = make_uint2(jid, (local / WH) * WH + py * (uint32_t)job.width + px);
                r = ps.ray;
            } else {
                slot = A[idx];
                const float4 o = c.P.ro[slot], d = c.P.rd[slot];
                r.o = Vec3{o.x, o.y, o.z};
                r.d = Vec3{d.x, d.y, d.z};
                r.tmin = kRayEps;
                r.tmax = kInf;
                if constexpr (!SINGLE) jid = c.wk[slot].x;
            }
            const int sc = job_of<SINGLE>(c, jid).scene;
            const SceneView& s = scene_of<SINGLE>(c, sc);
            Hit h;
            intersect_with(s, src_of(s, lds, &sp, sc), r, &h, false);
            if (!hit_is_valid(s, h)) {
                inv += 1;
                h.prim = -1;
            }
            c.P.hit[slot] = make_float2(h.t, __int_as_float(h.prim));
            if (h.prim < 0) {
                cat = 0;
            } else {
                const int mt = s.materials[s.prims[h.prim].material].type;
                cat = mt == kEmissive ? 0 : mt + 1;
            }
        }
#pragma unroll
        for (int q = 0; q < 4; ++q) wave_push(c.Q[q], &st->q[q], c.cap, slot, valid && cat == q, &st->overflow);
    }
    wave_add_u32(&c.ctr->invalid_hits, inv);
}

template <bool SINGLE>
__global__ __launch_bounds__(512) void pw_shade_kernel(PCtx c, int parity) {
    PState* st = c.st;
    const uint32_t n0 = st->q[0], n1 = st->q[1], n2 = st->q[2], n3 = st->q[3];
    const uint32_t e1 = n0 + n1, e2 = e1 + n2, total = e2 + n3;
    uint32_t* An = parity ? c.A[0] : c.A[1];
    uint32_t* Ffree = c.F + st->it_free_after;
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t b0 = blockIdx.x * blockDim.x; b0 < total; b0 += stride) {
        const uint32_t idx = b0 + threadIdx.x;
        const bool valid = idx < total;
        uint32_t slot = 0, jid = 0;
        bool live = false, fin = false;
        int has = 0, sc = 0;
        Ray sh;
        Vec3 contrib;
        if (valid) {
            const uint32_t* Qp;
            uint32_t qi;
            if (idx < n0) { Qp = c.Q[0]; qi = idx; }
            else if (idx < e1) { Qp = c.Q[1]; qi = idx - n0; }
            else if (idx < e2) { Qp = c.Q[2]; qi = idx - e1; }
            else { Qp = c.Q[3]; qi = idx - e2; }
            slot = Qp[qi];
            const uint2 w = c.wk[slot];
            if constexpr (!SINGLE) jid = w.x;
            const JobDev& job = job_of<SINGLE>(c, jid);
            sc = job.scene;
            const SceneView& s = scene_of<SINGLE>(c, sc);
            PathState ps = load_path(c.P, slot, true);
            const float2 hv = c.P.hit[slot];
            const Hit h{hv.x, __float_as_int(hv.y)};
            Surface f;
            const int mat = classify(s, ps, h, job.max_depth, &f);
            if (mat == kLambert) live = shade_lambert(s, ps, f, &sh, &contrib, &has);
            else if (mat == kMetal) live = shade_metal(s, ps, f);
            else if (mat == kDielectric) live = shade_dielectric(s, ps, f);
            if (live) store_scatter(c.P, slot, ps);
            const float4 L = make_float4(ps.L.x, ps.L.y, ps.L.z, 0.0f);
            fin = !live && !has;
            if (fin) job.sbuf[w.y] = L;
            else c.P.L[slot] = L;
        }
        wave_push(An, &st->n_next, c.cap, slot, live, &st->overflow);
        wave_push(Ffree, &st->n_freed, c.cap, slot, fin, &st->overflow);
        count_done(c.job_done, fin, jid);
        const uint32_t e = wave_reserve(&st->q[4], has != 0);
        if (has) {
            c.SQ.o[e] = make_float4(sh.o.x, sh.o.y, sh.o.z, sh.tmax);
            c.SQ.d[e] = make_float4(sh.d.x, sh.d.y, sh.d.z, __int_as_float((int)(slot | (live ? 0u : 0x80000000u))));
            c.SQ.c[e] = make_float4(contrib.x, contrib.y, contrib.z, __int_as_float(sc));
        }
    }
}

template <bool SINGLE>
__global__ __launch_bounds__(512) void pw_shadow_kernel(PCtx c) {
    extern __shared__ Node lds[];
    __shared__ LdsPlan sp;
    stage_lds<SINGLE>(c, lds, &sp);
    PState* st = c.st;
    const uint32_t n = st->q[4];
    uint32_t* Ffree = c.F + st->it_free_after;
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t b0 = blockIdx.x * blockDim.x; b0 < n; b0 += stride) {
        const uint32_t idx = b0 + threadIdx.x;
        uint32_t slot = 0, jid = 0;
        bool fin = false;
        if (idx < n) {
            const float4 o = c.SQ.o[idx], d = c.SQ.d[idx], cc = c.SQ.c[idx];
            const uint32_t tag = (uint32_t)__float_as_int(d.w);
            slot = tag & 0x7fffffffu;
            fin = (tag >> 31) != 0;
            const int sc = SINGLE ? 0 : __float_as_int(cc.w);
            const SceneView& s = scene_of<SINGLE>(c, sc);
            Ray r;
            r.o = Vec3{o.x, o.y, o.z};
            r.d = Vec3{d.x, d.y, d.z};
            r.tmin = kRayEps;
            r.tmax = o.w;
            Hit h;
            const bool blocked = intersect_with(s, src_of(s, lds, &sp, sc), r, &h, true);
            if (!blocked || fin) {
                float4 l = c.P.L[slot];
                if (!blocked) {
                    l.x += cc.x;
                    l.y += cc.y;
                    l.z += cc.z;
                }
                if (fin) {
                    const uint2 w = c.wk[slot];
                    if constexpr (!SINGLE) jid = w.x;
                    job_of<SINGLE>(c, jid).sbuf[w.y] = l;
                } else {
                    c.P.L[slot] = l;
                }
            }
        }
        wave_push(Ffree, &st->n_freed, c.cap, slot, fin, &st->overflow);
        count_done(c.job_done, fin, jid);
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
    int block = kDefaultBlock, graph_iters = kDefaultGraphIters, lds_bytes = kDefaultLdsKiB * 1024, cus = 0;
    std::vector<void*> pool_allocs;
    PState* st = nullptr;
    PState* st_host = nullptr;       // pinned mirror
    uint32_t* done_host = nullptr;   // pinned mirror of job_done
    PCtx ctx{};
    LdsPlan lds_plan{};
    struct HostJob {
        JobDev dev;
        uint64_t work = 0;
        uint32_t done = 0;
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
        if (done_host) (void)hipHostFree(done_host);
    }
};

struct PersistentEngine::LaunchPlan {
    PCtx c;
    bool single;
    size_t lds_bytes;
    int gx, gs, gw;
};

PersistentEngine::PersistentEngine() : m_(std::make_unique<Impl>()) {}
PersistentEngine::~PersistentEngine() = default;

uint64_t PersistentEngine::bytes() const { return pool_bytes_ + m_->sbuf_bytes; }

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
    m.graph_iters = t.graph_iters > 0 ? ((t.graph_iters + 1) & ~1) : kDefaultGraphIters;
    const int kib = t.lds_kib > 0 ? t.lds_kib : kDefaultLdsKiB;
    if (kib > 64) throw std::runtime_error("persistent: at most 64 KiB of LDS per workgroup");
    m.lds_bytes = kib * 1024;
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
    alloc(&c.P.hit, P * 8);
    alloc(&c.SQ.o, P * 16);
    alloc(&c.SQ.d, P * 16);
    alloc(&c.SQ.c, P * 16);
    alloc(&c.wk, P * 8);
    alloc(&c.A[0], P * 4);
    alloc(&c.A[1], P * 4);
    alloc(&c.F, P * 4);
    for (int q = 0; q < 4; ++q) alloc(&c.Q[q], P * 4);
    alloc(&c.st, sizeof(PState));
    alloc(&c.jobs, sizeof(JobDev) * kMaxJobs);
    alloc(&c.job_done, sizeof(uint32_t) * kMaxJobs);
    alloc(&c.scenes, sizeof(SceneView) * kMaxScenes);
    alloc(&c.plan, sizeof(LdsPlan));
    m.st = c.st;
    if (!m.st_host) check(hipHostMalloc(&m.st_host, sizeof(PState), hipHostMallocDefault), "hipHostMalloc");
    if (!m.done_host) check(hipHostMalloc(&m.done_host, sizeof(uint32_t) * kMaxJobs, hipHostMallocDefault), "hipHostMalloc");
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
    update_lds_plan();
}

// Splits the LDS budget evenly over the scenes: BVH nodes first (the breadth-first top), then the
// scene's primitive records if all of them fit in what is left of its share.
void PersistentEngine::update_lds_plan() {
    Impl& m = *m_;
    LdsPlan p{};
    p.n = (int)m.hscenes.size();
    if (p.n > 0) {
        const int share = m.lds_bytes / p.n;
        int nodes = 0, prims = 0;
        for (int k = 0; k < p.n; ++k) {
            const SceneView& s = m.hscenes[k];
            p.node_cnt[k] = std::min(s.n_nodes, share / (int)sizeof(Node));
            const int left = share - p.node_cnt[k] * (int)sizeof(Node);
            p.prim_cnt[k] = (p.node_cnt[k] == s.n_nodes && s.n_prims * 16 <= left) ? s.n_prims : 0;
            p.node_off[k] = nodes;
            p.prim_off[k] = prims;
            nodes += p.node_cnt[k];
            prims += p.prim_cnt[k];
        }
        p.total_nodes = nodes;
        p.total_prims = prims;
    }
    m.lds_plan = p;
    check(hipMemcpyAsync(m.ctx.plan, &m.lds_plan, sizeof(LdsPlan), hipMemcpyHostToDevice, m.stream), "copy lds plan");
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
        check(hipMemcpyAsync(m.ctx.scenes + sc, job.scene, sizeof(SceneView), hipMemcpyHostToDevice, m.stream),
              "copy scene");
        update_lds_plan();
        m.free_graph();
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
    pw_admit_kernel<<<1, 1, 0, m.stream>>>(m.st, m.ctx.jobs, m.ctx.job_done, d, index);
    check(hipGetLastError(), "persistent admit");
    Impl::HostJob hj;
    hj.dev = d;
    hj.work = work;
    m.hjobs.push_back(hj);
    if (m.graph && m.graph_single) m.free_graph();   // the single-job graph bakes the job in
    return (int)index;
}

PersistentEngine::LaunchPlan PersistentEngine::make_plan(bool single) {
    Impl& m = *m_;
    LaunchPlan L{};
    L.single = single;
    L.c = m.ctx;
    L.c.plan0 = m.lds_plan;
    if (single) {
        L.c.scene0 = m.hscenes[0];
        L.c.job0 = m.hjobs[0].dev;
    }
    L.lds_bytes = (size_t)m.lds_plan.total_nodes * sizeof(Node) + (size_t)m.lds_plan.total_prims * sizeof(float4);
    const void* fx = single ? (const void*)pw_extend_kernel<true> : (const void*)pw_extend_kernel<false>;
    const void* fs = single ? (const void*)pw_shade_kernel<true> : (const void*)pw_shade_kernel<false>;
    const void* fw = single ? (const void*)pw_shadow_kernel<true> : (const void*)pw_shadow_kernel<false>;
    L.gx = grid_for(fx, m.block, L.lds_bytes, m.cus);
    L.gs = grid_for(fs, m.block, 0, m.cus);
    L.gw = grid_for(fw, m.block, L.lds_bytes, m.cus);
    return L;
}

// One iteration; with `ev` (5 events) the stage boundaries are recorded for profiling:
// ev[0] plan ev[1] extend ev[2] shade ev[3] shadow ev[4].
void PersistentEngine::launch_iteration(const LaunchPlan& L, int par, hipEvent_t* ev) {
    Impl& m = *m_;
    hipStream_t st = m.stream;
    const PCtx& c = L.c;
    if (ev) check(hipEventRecord(ev[0], st), "hipEventRecord");
    pw_plan_kernel<<<1, 1, 0, st>>>(m.st, m.ctr);
    if (ev) check(hipEventRecord(ev[1], st), "hipEventRecord");
    if (L.single) pw_extend_kernel<true><<<L.gx, m.block, L.lds_bytes, st>>>(c, par);
    else pw_extend_kernel<false><<<L.gx, m.block, L.lds_bytes, st>>>(c, par);
    if (ev) check(hipEventRecord(ev[2], st), "hipEventRecord");
    if (L.single) pw_shade_kernel<true><<<L.gs, m.block, 0, st>>>(c, par);
    else pw_shade_kernel<false><<<L.gs, m.block, 0, st>>>(c, par);
    if (ev) check(hipEventRecord(ev[3], st), "hipEventRecord");
    if (L.single) pw_shadow_kernel<true><<<L.gw, m.block, L.lds_bytes, st>>>(c);
    else pw_shadow_kernel<false><<<L.gw, m.block, L.lds_bytes, st>>>(c);
    if (ev) check(hipEventRecord(ev[4], st), "hipEventRecord");
}

void PersistentEngine::capture(bool single) {
    Impl& m = *m_;
    m.free_graph();
    const LaunchPlan L = make_plan(single);
    check(hipStreamBeginCapture(m.stream, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
    for (int it = 0; it < m.graph_iters; ++it) launch_iteration(L, it & 1, nullptr);
    check(hipStreamEndCapture(m.stream, &m.graph), "hipStreamEndCapture");
    check(hipGetLastError(), "persistent capture");
    check(hipGraphInstantiate(&m.exec, m.graph, nullptr, nullptr, 0), "hipGraphInstantiate");
    m.graph_single = single;
    m.graph_scenes = (int)m.hscenes.size();
}

void PersistentEngine::run(int launches) {
    Impl& m = *m_;
    if (m.hjobs.empty()) return;
    const bool single = m.hjobs.size() == 1 && m.hscenes.size() == 1;
    if (!m.exec || m.graph_single != single || m.graph_scenes != (int)m.hscenes.size()) capture(single);
    for (int k = 0; k < launches; ++k) check(hipGraphLaunch(m.exec, m.stream), "hipGraphLaunch");
}

// Eager iterations with HIP events around each stage (synchronising once per iteration), until
// the admitted work is done. Items per stage come from the control block of each iteration.
void PersistentEngine::run_profiled(StageTimes* out) {
    Impl& m = *m_;
    if (m.hjobs.empty()) return;
    const bool single = m.hjobs.size() == 1 && m.hscenes.size() == 1;
    const LaunchPlan L = make_plan(single);
    hipEvent_t ev[5];
    for (auto& e : ev) check(hipEventCreate(&e), "hipEventCreate");
    for (int it = 0;; ++it) {
        launch_iteration(L, it & 1, ev);
        check(hipMemcpyAsync(m.st_host, m.st, sizeof(PState), hipMemcpyDeviceToHost, m.stream), "copy state");
        check(hipStreamSynchronize(m.stream), "profile sync");
        const PState& s = *m.st_host;
        float ms[4];
        for (int k = 0; k < 4; ++k) check(hipEventElapsedTime(&ms[k], ev[k], ev[k + 1]), "hipEventElapsedTime");
        if (s.done) break;
        out->other += ms[0];
        out->extend += ms[1];
        out->shade_lambert += ms[2];
        out->shadow += ms[3];
        out->items[0] += s.it_act;
        out->items[2] += (long long)s.q[0] + s.q[1] + s.q[2] + s.q[3];
        out->items[5] += s.q[4];
        out->items[7] += 1;
        out->launches[0] += 1;
        out->launches[2] += 1;
        out->launches[5] += 1;
        out->launches[7] += 1;
    }
    for (auto& e : ev) (void)hipEventDestroy(e);
    // Fold the last iteration's counts (the plan of the final, empty iteration did so on device).
}

PersistentProgress PersistentEngine::progress() {
    Impl& m = *m_;
    const size_t nj = m.hjobs.size();
    check(hipMemcpyAsync(m.st_host, m.st, sizeof(PState), hipMemcpyDeviceToHost, m.stream), "copy progress");
    if (nj) check(hipMemcpyAsync(m.done_host, m.ctx.job_done, nj * sizeof(uint32_t), hipMemcpyDeviceToHost, m.stream),
                  "copy job progress");
    check(hipStreamSynchronize(m.stream), "persistent sync");
    for (size_t k = 0; k < nj; ++k) m.hjobs[k].done = m.done_host[k];
    PersistentProgress p;
    p.next_work = m.st_host->next_work;
    p.total_work = m.st_host->total_work;
    p.n_live = m.st_host->n_live;
    p.done = m.st_host->done;
    p.iterations = m.st_host->iterations;
    return p;
}

bool PersistentEngine::job_complete(int job) const {
    const Impl::HostJob& j = m_->hjobs[job];
    return (uint64_t)j.done == j.work;
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
    // Split jobs into consecutive sample chunks so that one chunk's sample buffers fit the budget.
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
        for (int k = 0; k < (int)m_->hjobs.size(); ++k) {
            if (!job_complete(k)) throw std::runtime_error("persistent: job incomplete after the work drained");
            finish_job(k);
        }
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
    const int ldss[] = {32, 64};
    PersistentEngine eng;
    hipEvent_t a, b;
    check(hipEventCreate(&a), "hipEventCreate");
    check(hipEventCreate(&b), "hipEventCreate");
    // Probe renders add into the jobs' films; the caller resets them afterwards.
    for (int bs : blocks)
        for (int pool : pools) {
            if ((uint64_t)pool > 2 * work && pool != pools[0]) continue;
            for (int kib : ldss) {
                PersistentTuning t;
                t.block_size = bs;
                t.pool_paths = pool;
                t.lds_kib = kib;
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
