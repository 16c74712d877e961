// Persistent batched wavefront engine; the design is described in wf_persistent.h.
// Per iteration (all counts live in the device control block PState):
//   plan    1 thread: fold the previous iteration's counts, size this one
//   extend  closest-hit traversal of items [0, n_act) (survivors, then new paths)
//   shade   classify + shade every item; compact survivors into the other buffer, finish the rest
//   shadow  any-hit shadow rays; add unoccluded contributions
// Stage semantics and the order of L accumulation are those of the plain wavefront
// (src/gpu/wavefront.cu:88-270): emission in classify, then the unoccluded NEE contribution
// before the next classify, so every path computes the megakernel's bits.
#include <algorithm>
#include <cstring>
#include <mutex>

#include "wf_persistent.h"

namespace wf::gpu {

namespace {

// Path slots when the caller gives none: as many as the admitted work, up to 2^26 (12 GB of path
// state, 184 B a slot). More slots mean fewer, fuller iterations: each iteration's launches pay
// their prologues and the straggling waves at their end, so a render at a pool of its whole work
// runs in depth + 1 iterations (sphere_grid 1280x720x64: 41 iterations at 2^22 slots, 19.1 ms;
// 9 at 2^26, 14.4 ms).
constexpr int kDefaultPool = 1 << 26;
// Default workgroup size and LDS budget of the traversal kernels, by mode. A single job's kernels
// (60-64 VGPRs) run 8 waves per SIMD: two 1024-thread workgroups with a 64 KiB tree top each fill
// a CU, where 512-thread workgroups with the same top fit only two per CU (4 waves per SIMD;
// heightfield 31.4 -> 22.7 ms per 1280x720x64 render with 1024). The multi-job extend (64 VGPRs)
// and shadow (69) kernels run four and three 512-thread workgroups per CU with 32 KiB tops, where
// one 1024-thread workgroup would use 16 of the 32 wave slots (the 64-job batch 38.4 -> 32.9 ms,
// serving at 60 requests/s p99 119 -> 89 ms). A tuning's block size or LDS budget, when given,
// applies to both modes.
constexpr int kDefaultBlock = 1024, kDefaultBlockMulti = 512;
constexpr int kDefaultGraphIters = 16;
constexpr int kDefaultLdsKiB = 64, kDefaultLdsKiBMulti = 32;
constexpr int kDefaultRefill = 32;
constexpr int kDefaultLeafBatch = 16;
constexpr int kDefaultDirBins = 1;
constexpr int kDefaultShadeBlock = 512;
// Allowance for the extend and shadow kernels' static LDS (the plan, segment tables and, in the
// multi-job kernels, the job and scene tables and finished counts: under 7.5 KiB at gfx950).
constexpr int kLdsStatic = 8 * 1024;
constexpr uint32_t kIdle = 0xFFFFFFFFu;
constexpr uint32_t kFinalFlag = 0x80000000u;

struct PCtx {
    PState* st;
    JobDev* jobs;
    uint32_t* job_done;   // finished paths per job
    SceneView* scenes;
    LdsPlan* plan;
    float2* hit;          // per item: t, leaf slot (-1 = miss)
    ShadowQueue SQ;       // o: origin, tmax; d: direction, job | final flag; c: contribution, destination
    DevCounters* ctr;
    uint32_t cap;
    int refill;
    int leaf_batch;
    int dir_bins;         // 1, or 8: survivors of a block round are grouped by direction octant
    LdsPlan plan0;        // single mode: the LDS layout
    SceneView scene0;     // single mode: the scene
};

template <bool SINGLE>
__device__ __forceinline__ const SceneView& scene_of(const PCtx& c, int sc) {
    if constexpr (SINGLE) return c.scene0;
    else return c.scenes[sc];
}

// Single mode has one unfinished job, the last one admitted: its slot comes from the control block
// (a uniform address), so a captured graph does not depend on the job.
__device__ __forceinline__ uint32_t single_slot(const PState* st) { return (st->job_hi - 1u) & kJobMask; }

template <bool SINGLE>
__device__ __forceinline__ const JobDev& job_of(const PCtx& c, uint32_t jid, uint32_t slot0) {
    if constexpr (SINGLE) return c.jobs[slot0];
    else return c.jobs[jid];
}

// First job whose work range ends after w; end(k) is job k's work_end (increasing).
template <class End>
__device__ __forceinline__ int find_job(End end, uint32_t n_jobs, unsigned long long w) {
    int lo = 0, hi = (int)n_jobs - 1;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (end(mid) > w) hi = mid;
        else lo = mid + 1;
    }
    return lo;
}

// Multi-job kernels stage the scene table and, when it has at most kLdsJobs jobs, the work ends of
// the window of jobs whose paths are not all issued ([job_lo, job_hi)) in LDS: a new path finds its
// job by a binary search over that window (from memory when it is larger), and a lane reaches its
// scene record from LDS (survivors carry their scene slot).
constexpr int kLdsJobs = 256;
struct JobTab {
    SceneView scenes[kMaxScenes];
    unsigned long long work_end[kLdsJobs];   // jobs lo, lo + 1, ... of the window
    uint32_t lo, n;                          // window: sequence number of its first job, its size
    int staged;                              // work_end holds the whole window
};

// The whole block calls it. done, tag (optional): the workgroup's finished-path counters, cleared.
__device__ __forceinline__ void stage_jobs(const PCtx& c, const PState* st, JobTab* t, uint32_t* done = nullptr,
                                           int* tag = nullptr) {
    if (done)
        for (int i = threadIdx.x; i < kLdsJobs; i += blockDim.x) {
            done[i] = 0u;
            tag[i] = -1;
        }
    const int n_scenes = c.plan->n;
    const int* src = reinterpret_cast<const int*>(c.scenes);
    int* dst = reinterpret_cast<int*>(t->scenes);
    const int words = n_scenes * (int)(sizeof(SceneView) / sizeof(int));
    for (int i = threadIdx.x; i < words; i += blockDim.x) dst[i] = src[i];
    const uint32_t lo = st->job_lo, n = st->job_hi - lo;
    const bool staged = n <= (uint32_t)kLdsJobs;
    if (staged)
        for (int i = threadIdx.x; i < (int)n; i += blockDim.x) t->work_end[i] = c.jobs[(lo + i) & kJobMask].work_end;
    if (threadIdx.x == 0) {
        t->staged = staged ? 1 : 0;
        t->lo = lo;
        t->n = n;
    }
    __syncthreads();
}

// Pixel of position q in a pass, visiting the image in 8-row bands of 8x8 tiles (edge tiles
// smaller), so the 64 consecutive new paths of a wave start from one 8x8 pixel tile. Full tiles
// (all but the last band and column) divide by shifts; band: division by 8 * W.
__device__ __forceinline__ void tile_pixel(uint32_t q, uint32_t W, uint32_t H, const DivU32& band, uint32_t* px,
                                           uint32_t* py) {
    const uint32_t b = divu32(q, band);
    const uint32_t r = q - b * 8u * W;
    const uint32_t hb = min(8u, H - 8u * b);
    const uint32_t j = hb == 8u ? r >> 6 : r / (8u * hb);
    const uint32_t in = r - j * 8u * hb;
    const uint32_t wb = min(8u, W - 8u * j);
    *px = 8u * j + (wb == 8u ? in & 7u : in % wb);
    *py = 8u * b + (wb == 8u ? in >> 3 : in / wb);
}

// Single-job kernels stage the job record in LDS once per workgroup: new paths read its camera,
// size and divisors at every refill, which would otherwise be a round trip to memory each time.
// The whole block calls it; the caller synchronises before use.
__device__ __forceinline__ void stage_job0(const PCtx& c, uint32_t slot0, JobDev* dst) {
    const int* src = reinterpret_cast<const int*>(c.jobs + slot0);
    int* d = reinterpret_cast<int*>(dst);
    for (int i = threadIdx.x; i < (int)(sizeof(JobDev) / sizeof(int)); i += blockDim.x) d[i] = src[i];
}

// The path of work index w of job `job` (index jid): its sample-buffer index (sample-major, then
// pixel) and its initial state (include/wf/path_core.hpp:59-82).
__device__ __forceinline__ PathState path_of_job(const JobDev& job, uint32_t jid, unsigned long long w,
                                                 uint32_t* jid_out, uint32_t* sidx_out) {
    const uint32_t local = (uint32_t)(w - job.work_begin);
    const uint32_t W = (uint32_t)job.width, WH = W * (uint32_t)job.height;
    const uint32_t q = divu32(local, job.div_pixels);
    uint32_t px, py;
    tile_pixel(local - q * WH, W, (uint32_t)job.height, job.div_band, &px, &py);
    *jid_out = jid;
    *sidx_out = q * WH + py * W + px;
    return init_path_keyed(job.cam, job.width, job.height, (int)px, (int)py,
                           path_key_ab(job.key_a, job.key_b, py * W + px, job.first_sample + q));
}

typedef __attribute__((address_space(4))) const JobDev const_job;

// The fields of job j that a new path needs, read through the constant address space: with j
// wave-uniform these are scalar loads.
__device__ __forceinline__ JobDev job_fields_scalar(const JobDev* jobs, uint32_t j) {
    const_job* p = (const_job*)jobs + j;
    JobDev d{};
    d.work_begin = p->work_begin;
    d.width = p->width;
    d.height = p->height;
    d.div_pixels = DivU32{p->div_pixels.m, p->div_pixels.shift, p->div_pixels.d, 0u};
    d.div_band = DivU32{p->div_band.m, p->div_band.shift, p->div_band.d, 0u};
    d.cam.origin = Vec3{p->cam.origin.x, p->cam.origin.y, p->cam.origin.z};
    d.cam.lower_left = Vec3{p->cam.lower_left.x, p->cam.lower_left.y, p->cam.lower_left.z};
    d.cam.horizontal = Vec3{p->cam.horizontal.x, p->cam.horizontal.y, p->cam.horizontal.z};
    d.cam.vertical = Vec3{p->cam.vertical.x, p->cam.vertical.y, p->cam.vertical.z};
    d.first_sample = p->first_sample;
    d.key_a = p->key_a;
    d.key_b = p->key_b;
    d.scene = p->scene;
    return d;
}

// The path of work index w: its job slot, its sample-buffer index, its scene slot and its initial
// state. t: the multi-job kernels' LDS tables; job0: the single-job kernels' LDS copy of the job in
// slot0. In the multi-job kernels the lanes that start a path together nearly always belong to
// one job: its record then comes through scalar loads (the constant address space: job records are
// read-only in these kernels) instead of a round trip to memory in every lane.
template <bool SINGLE>
__device__ __forceinline__ PathState new_path(const PCtx& c, const JobTab* t, const JobDev* job0, uint32_t slot0,
                                              unsigned long long w, uint32_t* jid_out, uint32_t* sidx_out,
                                              uint32_t* scene_out) {
    if constexpr (SINGLE) {
        *scene_out = (uint32_t)job0->scene;
        return path_of_job(*job0, slot0, w, jid_out, sidx_out);
    } else {
        uint32_t k;
        if (t->staged) k = (uint32_t)find_job([&](int i) { return t->work_end[i]; }, t->n, w);
        else k = (uint32_t)find_job([&](int i) { return c.jobs[(t->lo + i) & kJobMask].work_end; }, t->n, w);
        const uint32_t jid = (t->lo + k) & kJobMask;
        const uint32_t j0 = __builtin_amdgcn_readfirstlane(jid);
        if (__ballot(jid != j0) == 0) {
            const JobDev d = job_fields_scalar(c.jobs, j0);
            *scene_out = (uint32_t)d.scene;
            return path_of_job(d, j0, w, jid_out, sidx_out);
        }
        *scene_out = (uint32_t)c.jobs[jid].scene;
        return path_of_job(c.jobs[jid], jid, w, jid_out, sidx_out);
    }
}

// Stages the LDS layout (BVH tops and primitive records); the whole block calls it.
template <bool SINGLE>
__device__ __forceinline__ void stage_lds(const PCtx& c, Node* lds, LdsPlan* sp) {
    if constexpr (SINGLE) {
        // Word by word with constant offsets: an aggregate copy of the by-value argument can
        // be lowered through private memory.
        constexpr int kWords = (int)(sizeof(LdsPlan) / sizeof(int));
        const int* pw = reinterpret_cast<const int*>(&c.plan0);
        int* dw = reinterpret_cast<int*>(sp);
        if (threadIdx.x == 0) {
#pragma unroll
            for (int i = 0; i < kWords; ++i) dw[i] = pw[i];
        }
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

// Scene source of the multi-job traversal kernels, the same nodes and primitive tests as LdsNodes
// (src_of) in fewer registers: a lane keeps its scene slot and the place and size of its tree top
// in LDS; the scene's global node and primitive pointers and its LDS primitive records are read
// from the workgroup's LDS tables where they are needed (a node below the top, a leaf), which is
// where a visit waits on memory anyway. Three VGPRs per lane instead of nine.
struct MultiSrc {
    const JobTab* jt;           // LDS (wave-uniform)
    const LdsPlan* sp;          // LDS (wave-uniform)
    lds_float4* prim_base;      // LDS primitive records of every scene (wave-uniform)
    lds_float4* top;            // this lane's tree top
    int n_lds, sc;
    __device__ __forceinline__ void set(const Node* lds, int slot) {
        sc = slot;
        top = (lds_float4*)(lds + sp->node_off[slot]);
        n_lds = sp->node_cnt[slot];
    }
    __device__ __forceinline__ Node node_lds(int n) const {
        const f32x4 a = top[2 * n], b = top[2 * n + 1];
        Node out;
        out.lo = Vec3{a.x, a.y, a.z};
        out.skip = __float_as_int(a.w);
        out.hi = Vec3{b.x, b.y, b.z};
        out.child = __float_as_int(b.w);
        return out;
    }
    __device__ __forceinline__ Node node(int n) const {
        Node out;
        if (n < n_lds) {
            out = node_lds(n);
            __builtin_amdgcn_sched_barrier(0);
        } else {
            out = load_node_global(jt->scenes[sc].nodes, n);
        }
        return out;
    }
    __device__ __forceinline__ bool all_lds(int n_nodes) const { return n_lds >= n_nodes; }
    __device__ __forceinline__ bool hit(int slot, const Ray& r, float* t) const {
        const SceneView& s = jt->scenes[sc];
        const glb_float4* g = (const glb_float4*)s.prims + 3 * slot;
        if (sp->prim_cnt[sc] > 0) {
            const f32x4 q = prim_base[sp->prim_off[sc] + slot];
            if (q.w >= 0.0f) return hit_sphere(Vec3{q.x, q.y, q.z}, q.w, r, t);
            const f32x4 b = g[1], c = g[2];
            return hit_triangle(Vec3{q.x, q.y, q.z}, Vec3{b.x, b.y, b.z}, Vec3{c.x, c.y, c.z}, r, t);
        }
        const f32x4 a = g[0], b = g[1];
        if (__float_as_int(a.w) < s.n_spheres) return hit_sphere(Vec3{a.x, a.y, a.z}, b.w, r, t);
        const f32x4 c = g[2];
        return hit_triangle(Vec3{a.x, a.y, a.z}, Vec3{b.x, b.y, b.z}, Vec3{c.x, c.y, c.z}, r, t);
    }
};

__device__ __forceinline__ MultiSrc multi_src(const JobTab* jt, const LdsPlan* sp, const Node* lds) {
    MultiSrc m{};
    m.jt = jt;
    m.sp = sp;
    m.prim_base = (lds_float4*)(lds + sp->total_nodes);
    return m;
}

// Adds one per lane with pred to the finished-path count of job slot jid, one update per distinct
// slot in the wave, into the workgroup's LDS counters: a direct-mapped cache of kLdsJobs entries
// whose tag[e] names the slot entry e counts for (the first slot with e = slot mod kLdsJobs to
// arrive claims it); a slot whose entry another slot holds adds to its device counter directly.
// The workgroup adds its counters to the device once at its end (flush_done): waves of the whole
// grid would otherwise serialise on the device counters of the few jobs in flight.
__device__ __forceinline__ void count_done(uint32_t* done, int* tag, uint32_t* job_done, bool pred, uint32_t jid) {
    uint64_t m = __ballot(pred);
    while (m) {
        const int leader = __ffsll((unsigned long long)m) - 1;
        const uint32_t j0 = __shfl(jid, leader);
        const uint64_t same = __ballot(pred && jid == j0);
        if ((int)lane_id() == leader) {
            const int e = (int)(j0 & (kLdsJobs - 1));
            const int owner = atomicCAS(&tag[e], -1, (int)j0);
            if (owner == -1 || owner == (int)j0) atomicAdd(&done[e], (uint32_t)__popcll(same));
            else atomicAdd(&job_done[j0], (uint32_t)__popcll(same));
        }
        m &= ~same;
    }
}

// Adds the workgroup's LDS counters to the device counters; the whole block calls it at its end.
__device__ __forceinline__ void flush_done(const uint32_t* done, const int* tag, uint32_t* job_done) {
    __syncthreads();
    for (int i = threadIdx.x; i < kLdsJobs; i += blockDim.x)
        if (tag[i] >= 0 && done[i]) atomicAdd(&job_done[tag[i]], done[i]);
}

// Output window of each segment for a shade pass over n items in rounds of B: segment s holds the
// items of the rounds r with r mod kSegs == s; windows follow each other in segment order.
__device__ __forceinline__ void seg_windows(uint32_t n, uint32_t B, uint32_t* base) {
    const uint32_t full = n / B, part = n - full * B;
    uint32_t acc = 0;
    for (int s = 0; s < kSegs; ++s) {
        base[s] = acc;
        const uint32_t rounds = full > (uint32_t)s ? (full - 1 - (uint32_t)s) / kSegs + 1 : 0u;
        acc += rounds * B + ((part > 0 && full % kSegs == (uint32_t)s) ? part : 0u);
    }
}

// Position of item i of a segmented list (largest s with prefix[s] <= i).
__device__ __forceinline__ uint32_t seg_pos(const uint32_t* prefix, const uint32_t* base, uint32_t i) {
    int s = 0;
#pragma unroll
    for (int step = kSegs / 2; step > 0; step >>= 1)
        if (prefix[s + step] <= i) s += step;
    return base[s] + (i - prefix[s]);
}

// Slices of `size` items are dealt to the waves of a traversal kernel per workgroup: workgroup
// b owns the slices b, b + B, b + 2B, ... (B workgroups), which sample the whole item range, and its
// waves take them in turn from a counter in LDS, the first one by wave index.
// A wave that drew cheap rays so takes more of its workgroup's slices, and the workgroup ends when
// its work ends rather than when its unluckiest wave finishes a fixed share. Every item is traced
// once, by whichever wave. Wave-uniform state; the whole wave calls these.
struct BlockFeed {
    uint32_t cursor = 0, end = 0;   // rest of the current slice [cursor, end)
    uint32_t size = kFeedSlice;     // slice size of the launch
    bool dry = false;               // no slice left
};

// Slice sizes. A launch ends with its last wave, and a wave's last slice ends with its longest
// ray, so shorter slices end a launch sooner while longer ones keep a wave on neighbouring rays
// for longer. Extension rays keep 128 (64 made heightfield's extend slower, 256 both scenes' at
// 2^24 slots); shadow launches are smaller and their rays cheaper, and take 64 (shadow time at
// 2^22 slots: sphere_grid 3.74 -> 3.36 ms, heightfield 11.29 -> 10.94 ms).
constexpr uint32_t kExtendSlice = kFeedSlice, kShadowSlice = 64;

__device__ __forceinline__ void feed_take(BlockFeed& f, uint32_t n, uint32_t j) {
    const unsigned long long c = ((unsigned long long)blockIdx.x + (unsigned long long)gridDim.x * j) * f.size;
    if (c >= n) {
        f.dry = true;
        f.cursor = f.end = n;
    } else {
        f.cursor = (uint32_t)c;
        f.end = min(f.cursor + f.size, n);
    }
}

// `ctr` (LDS) must hold the workgroup's wave count before any wave calls feed_next.
__device__ __forceinline__ void feed_first(BlockFeed& f, uint32_t n, uint32_t size) {
    f.size = size;
    feed_take(f, n, (uint32_t)__builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6)));
}

__device__ __forceinline__ void feed_next(BlockFeed& f, uint32_t n, uint32_t* ctr) {
    uint32_t j = 0;
    if (lane_id() == 0) j = atomicAdd(ctr, 1u);
    feed_take(f, n, (uint32_t)__builtin_amdgcn_readfirstlane((int)j));
}

// True when the workgroup owns no slice (its first is blockIdx.x): it returns before staging LDS.
__device__ __forceinline__ bool feed_idle(uint32_t n, uint32_t size) {
    return (unsigned long long)blockIdx.x * size >= n;
}

__global__ void pw_init_kernel(PState* st) {
    PState z{};
    *st = z;
}

// Host-side records reach the device as kernel arguments (copied into the launch's argument
// buffer when the launch is enqueued), never as pageable host-to-device copies: the runtime stages
// a pageable copy through its own buffers, and nothing in the render path depends on that path.
template <class T>
__global__ void pw_put_kernel(T* dst, T v) {
    static_assert(sizeof(T) % sizeof(int) == 0, "records are whole words");
    const int* s = reinterpret_cast<const int*>(&v);
    int* d = reinterpret_cast<int*>(dst);
#pragma unroll
    for (int i = 0; i < (int)(sizeof(T) / sizeof(int)); ++i) d[i] = s[i];
}

// Job `seq` (its sequence number) into its ring slot.
__global__ void pw_admit_kernel(PState* st, JobDev* jobs, uint32_t* job_done, JobDev job, uint32_t seq) {
    jobs[seq & kJobMask] = job;
    job_done[seq & kJobMask] = 0u;
    st->job_hi = seq + 1u;
    st->total_work = job.work_end;
    st->done = 0;
}

__global__ void pw_plan_kernel(PState* st, const JobDev* jobs, DevCounters* ctr, uint32_t cap, uint32_t shade_block) {
    if (st->have_prev) {
        // Fold the previous iteration: its survivors become this iteration's first items, found
        // through the segment windows of its shade pass.
        uint32_t next = 0, shadow = 0;
        uint32_t win[kSegs];
        seg_windows(st->it_act, st->shade_block, win);
        for (int s = 0; s < kSegs; ++s) {
            st->in_prefix[s] = next;
            st->in_base[s] = win[s];
            next += st->seg_cnt[s];
            shadow += st->sh_cnt[s];
        }
        st->in_prefix[kSegs] = next;
        ctr->extension_rays += st->it_act;
        ctr->shadow_rays += shadow;
        if (shadow > ctr->high_water[3]) ctr->high_water[3] = shadow;
        if (st->it_act > ctr->high_water[4]) ctr->high_water[4] = st->it_act;
        st->n_live = next;
    }
    st->have_prev = 1;
    st->shade_block = shade_block;
    // Jobs whose paths have all been issued leave the window that new paths are looked up in.
    while (st->job_lo != st->job_hi && jobs[st->job_lo & kJobMask].work_end <= st->next_work) st->job_lo += 1u;
    const unsigned long long rem = st->total_work - st->next_work;
    const uint32_t room = cap - st->n_live;
    const uint32_t n_new = (uint32_t)((unsigned long long)room < rem ? (unsigned long long)room : rem);
    st->it_live = st->n_live;
    st->it_new = n_new;
    st->it_act = st->n_live + n_new;
    st->it_work_base = st->next_work;
    st->next_work += n_new;
    st->n_next = 0;
    st->n_shadow = 0;
    st->unocc = 0;
    for (int s = 0; s < kSegs; ++s) {
        st->seg_cnt[s] = 0;
        st->sh_cnt[s] = 0;
    }
    if (st->it_act > 0) st->iterations += 1;
    st->done = st->it_act == 0 ? 1u : 0u;
}

// Closest-hit traversal of items [0, it_act), "while-while" with dynamic fetch per wave.
template <bool SINGLE>
__global__ __launch_bounds__(1024) void pw_extend_kernel(PCtx c, PathBuf in) {
    extern __shared__ Node lds[];
    __shared__ LdsPlan sp;
    __shared__ JobTab jt;   // multi-job only
    __shared__ __align__(16) unsigned char job0_raw[sizeof(JobDev)];   // single-job only
    JobDev& job0 = *reinterpret_cast<JobDev*>(job0_raw);
    const PState* st = c.st;
    const uint32_t n_act = st->it_act, n_live = st->it_live, slot0 = single_slot(st);
    const unsigned long long wbase = st->it_work_base;
    const uint32_t fsize = kExtendSlice;
    if (feed_idle(n_act, fsize)) return;
    __shared__ uint32_t pre[kSegs + 1], bas[kSegs], feed_ctr;
    if (threadIdx.x == 0) feed_ctr = blockDim.x >> 6;
    if (threadIdx.x <= kSegs) pre[threadIdx.x] = st->in_prefix[threadIdx.x];
    if (threadIdx.x < kSegs) bas[threadIdx.x] = st->in_base[threadIdx.x];
    if constexpr (!SINGLE) stage_jobs(c, st, &jt);
    else stage_job0(c, slot0, &job0);   // stage_lds synchronises
    stage_lds<SINGLE>(c, lds, &sp);   // synchronises
    BlockFeed f;
    feed_first(f, n_act, fsize);
    LdsNodes src{};
    MultiSrc msrc = multi_src(&jt, &sp, lds);
    int nn = 0, sc = 0;
    if constexpr (SINGLE) {
        src = src_of(c.scene0, lds, &sp, 0);
        nn = c.scene0.n_nodes;
    }
    uint32_t item = kIdle, inv = 0;
    Traversal tr;
    tr.node = -1;
    tr.prim = -1;
    int leaf = kNoLeaf;
    for (;;) {
        const uint64_t idle = __ballot(item == kIdle);
        if (idle != 0 && f.cursor >= f.end && !f.dry) feed_next(f, n_act, &feed_ctr);
        if (idle != 0 && f.cursor < f.end) {
            const uint32_t cand = f.cursor + mbcnt64(idle);
            if (item == kIdle && cand < f.end) {
                item = cand;
                Ray r;
                if (item < n_live) {
                    const uint32_t p = seg_pos(pre, bas, item);
                    const float4 o = in.ro[p], d = in.rd[p];
                    r.o = Vec3{o.x, o.y, o.z};
                    r.d = Vec3{d.x, d.y, d.z};
                    r.tmin = kRayEps;
                    r.tmax = kInf;
                    if constexpr (!SINGLE) sc = (int)(((uint32_t)__float_as_int(d.w) >> kSceneShift) & 7u);
                } else {
                    uint32_t jid, sidx, scn;
                    r = new_path<SINGLE>(c, &jt, &job0, slot0, wbase + (item - n_live), &jid, &sidx, &scn).ray;
                    if constexpr (!SINGLE) sc = (int)scn;
                }
                if constexpr (!SINGLE) {
                    msrc.set(lds, sc);
                    nn = jt.scenes[sc].n_nodes;
                }
                tr = traversal_begin(r);
                if (nn <= 0) tr.node = -1;
            }
            f.cursor = min(f.cursor + (uint32_t)__popcll(idle), f.end);
        }
        if (__ballot(item != kIdle) == 0) {
            if (f.dry) break;
            continue;   // slice used up: take the next one
        }
        const int limit = (f.dry && f.cursor >= f.end) ? 1 : c.refill;
        if constexpr (SINGLE) wave_traverse(nn, src, tr, leaf, false, limit, c.leaf_batch);
        else wave_traverse(nn, msrc, tr, leaf, false, limit, c.leaf_batch);
        if (item != kIdle && tr.node < 0 && !leaf_parked(leaf)) {
            Hit h{tr.r.tmax, tr.prim};
            const SceneView* hs;
            if constexpr (SINGLE) hs = &c.scene0;
            else hs = &jt.scenes[sc];
            if (!hit_is_valid(*hs, h)) {
                inv += 1;
                h.prim = -1;
            }
            c.hit[item] = make_float2(h.t, __int_as_float(h.prim));
            item = kIdle;
        }
    }
    wave_add_u32(&c.ctr->invalid_hits, inv);
}

// Inputs of one shade item: its hit and, for a survivor of the previous iteration, its state.
struct ShadeIn {
    float2 hv;
    float4 o, d, t, l;
};

__device__ __forceinline__ void shade_load(const PCtx& c, const PathBuf& in, const uint32_t* pre, const uint32_t* bas,
                                           uint32_t item, uint32_t n, uint32_t n_live, ShadeIn* x) {
    if (item >= n) return;
    x->hv = c.hit[item];
    if (item < n_live) {
        const uint32_t p = seg_pos(pre, bas, item);
        x->o = in.ro[p];
        x->d = in.rd[p];
        x->t = in.T[p];
        x->l = in.L[p];
    }
}

// Classify + shade of items [0, it_act). Survivors are compacted into `out`; slots for survivors
// and shadow rays are reserved once per block round. Inside a round the survivors are grouped by
// the octant of their new direction (dir_bins = 8), each group in item order, so the waves of the
// next extend see rays of nearby pixels heading the same way.
template <bool SINGLE>
__global__ __launch_bounds__(1024) __attribute__((amdgpu_waves_per_eu(6))) void pw_shade_kernel(PCtx c, PathBuf in, PathBuf out) {
    __shared__ uint32_t cnt[2][9][16];    // per round parity, per wave: 8 direction bins, shadow
    __shared__ uint32_t bases[2][2];
    __shared__ uint32_t pre[kSegs + 1], bas[kSegs], win[kSegs];
    __shared__ JobTab jt;                   // multi-job only
    __shared__ uint32_t done_cnt[kLdsJobs];   // multi-job only
    __shared__ int done_tag[kLdsJobs];        // multi-job only
    __shared__ __align__(16) unsigned char job0_raw[sizeof(JobDev)];   // single-job only
    JobDev& job0 = *reinterpret_cast<JobDev*>(job0_raw);
    PState* st = c.st;
    const uint32_t n = st->it_act, n_live = st->it_live, slot0 = single_slot(st);
    const unsigned long long wbase = st->it_work_base;
    const uint32_t wave = threadIdx.x >> 6, nwaves = blockDim.x >> 6;
    const uint32_t stride = gridDim.x * blockDim.x;
    if (blockIdx.x * blockDim.x >= n) return;   // no round for this workgroup
    if constexpr (!SINGLE) stage_jobs(c, st, &jt, done_cnt, done_tag);
    else stage_job0(c, slot0, &job0);   // synchronised below
    if (threadIdx.x <= kSegs) pre[threadIdx.x] = st->in_prefix[threadIdx.x];
    if (threadIdx.x < kSegs) bas[threadIdx.x] = st->in_base[threadIdx.x];
    if (threadIdx.x == 0) seg_windows(n, blockDim.x, win);
    __syncthreads();
    int round = 0;
    for (uint32_t b0 = blockIdx.x * blockDim.x; b0 < n; b0 += stride, round ^= 1) {
        const uint32_t seg = (b0 / blockDim.x) % kSegs;
        const uint32_t item = b0 + threadIdx.x;
        // Loading the next round's inputs ahead (a software pipeline) measured slower at the
        // default 512-thread blocks: the extra live registers cost more occupancy than it hid.
        ShadeIn cur{};
        shade_load(c, in, pre, bas, item, n, n_live, &cur);
        bool live = false, fin = false;
        int has = 0;
        uint32_t jid = 0, sidx = 0, sc = 0;
        PathState ps{};
        PathKey key{0u, 0u};
        Ray sh{};
        Vec3 contrib{0.0f, 0.0f, 0.0f};
        if (item < n) {
            const float2 hv = cur.hv;
            if (item < n_live) {
                const float4 o = cur.o, d = cur.d, t = cur.t, l = cur.l;
                ps.ray.o = Vec3{o.x, o.y, o.z};
                ps.ray.d = Vec3{d.x, d.y, d.z};
                ps.ray.tmin = kRayEps;
                ps.ray.tmax = kInf;
                ps.prev_pdf = o.w;
                const uint32_t bits = (uint32_t)__float_as_int(d.w);
                ps.depth = (int)(bits & ((1u << kSpecShift) - 1u));
                ps.specular = (int)((bits >> kSpecShift) & 1u);
                ps.throughput = Vec3{t.x, t.y, t.z};
                ps.L = Vec3{l.x, l.y, l.z};
                jid = (bits >> kJobShift) & kJobMask;
                sc = (bits >> kSceneShift) & 7u;
                sidx = (uint32_t)__float_as_int(l.w);
                // k1 from the job's seed hash, the pixel and the sample (the sample-buffer index
                // is sample offset * width * height + pixel).
                const JobDev& jb = SINGLE ? job0 : c.jobs[jid];
                const uint32_t q = divu32(sidx, jb.div_pixels);
                const uint32_t pix = sidx - q * (uint32_t)jb.width * (uint32_t)jb.height;
                ps.key = PathKey{(uint32_t)__float_as_int(t.w), path_key_k1(jb.key_b, pix, jb.first_sample + q)};
            } else {
                ps = new_path<SINGLE>(c, &jt, &job0, slot0, wbase + (item - n_live), &jid, &sidx, &sc);
            }
            key = ps.key;
            const JobDev& job = SINGLE ? job0 : c.jobs[jid];
            const SceneView* sv;
            if constexpr (SINGLE) sv = &c.scene0;
            else sv = &jt.scenes[sc];
            const int depth = job.max_depth;
            const SceneView& s = *sv;
            const Hit h{hv.x, __float_as_int(hv.y)};
            Surface f;
            const int mat = classify(s, ps, h, depth, &f);
            if (mat >= kLambert && mat <= kDielectric) live = shade_surface(s, ps, f, mat, &sh, &contrib, &has);
            fin = !live && !has;
            if (!live) job.sbuf[sidx] = make_float4(ps.L.x, ps.L.y, ps.L.z, 0.0f);
        }
        const int nb = c.dir_bins;
        int bin = 0;
        if (nb > 1)
            bin = (ps.ray.d.x < 0.0f ? 1 : 0) | (ps.ray.d.y < 0.0f ? 2 : 0) | (ps.ray.d.z < 0.0f ? 4 : 0);
        uint32_t(*cn)[16] = cnt[round];
        uint64_t ml = 0;
        for (int b = 0; b < nb; ++b) {
            const uint64_t m = __ballot(live && bin == b);
            if (bin == b) ml = m;
            if (lane_id() == 0) cn[b][wave] = (uint32_t)__popcll(m);
        }
        const uint64_t ms = __ballot(has != 0);
        if (lane_id() == 0) cn[8][wave] = (uint32_t)__popcll(ms);
        __syncthreads();
        // Wave 0 turns the counts into offsets: survivors bin-major then wave, shadow rays by wave
        // (exclusive prefix sums over the lanes), and reserves the round's slots in its segment.
        // A single job's finished paths are counted once per iteration (shadow kernel).
        if (wave == 0) {
            const uint32_t l = lane_id();
            uint32_t* flat = &cn[0][0];   // [b][w] at b * 16 + w; lane l holds slots 2l, 2l + 1
            const uint32_t s0 = 2u * l, s1 = s0 + 1u;
            const bool v0 = (s0 >> 4) < (uint32_t)nb && (s0 & 15u) < nwaves;
            const bool v1 = (s1 >> 4) < (uint32_t)nb && (s1 & 15u) < nwaves;
            const uint32_t a0 = v0 ? flat[s0] : 0u, a1 = v1 ? flat[s1] : 0u;
            const uint32_t c8 = l < nwaves ? cn[8][l] : 0u;
            uint32_t x = a0 + a1, y = c8;
#pragma unroll
            for (int d = 1; d < 64; d <<= 1) {
                const uint32_t xu = (uint32_t)__shfl_up((int)x, d), yu = (uint32_t)__shfl_up((int)y, d);
                if ((int)l >= d) {
                    x += xu;
                    y += yu;
                }
            }
            const uint32_t t0 = (uint32_t)__shfl((int)x, 63), t1 = (uint32_t)__shfl((int)y, 63);
            const uint32_t e = x - a0 - a1;
            if (v0) flat[s0] = e;
            if (v1) flat[s1] = e + a0;
            if (l < nwaves) cn[8][l] = y - c8;
            uint32_t r = 0u;
            if (l == 0 && t0) r = atomicAdd(&st->seg_cnt[seg], t0);
            if (l == 1 && t1) r = atomicAdd(&st->sh_cnt[seg], t1);
            if (l < 2) bases[round][l] = win[seg] + r;
        }
        __syncthreads();
        uint32_t o = 0;
        if (live) {
            o = bases[round][0] + cn[bin][wave] + mbcnt64(ml);
            out.ro[o] = make_float4(ps.ray.o.x, ps.ray.o.y, ps.ray.o.z, ps.prev_pdf);
            out.rd[o] = make_float4(ps.ray.d.x, ps.ray.d.y, ps.ray.d.z,
                                    __int_as_float((int)pack_rdw((uint32_t)ps.depth, (uint32_t)ps.specular, jid, sc)));
            out.T[o] = make_float4(ps.throughput.x, ps.throughput.y, ps.throughput.z, __int_as_float((int)key.k0));
            out.L[o] = make_float4(ps.L.x, ps.L.y, ps.L.z, __int_as_float((int)sidx));
        }
        if (has) {
            const uint32_t e = bases[round][1] + cn[8][wave] + mbcnt64(ms);
            c.SQ.o[e] = make_float4(sh.o.x, sh.o.y, sh.o.z, sh.tmax);
            c.SQ.d[e] = make_float4(sh.d.x, sh.d.y, sh.d.z,
                                    __int_as_float((int)(jid | (sc << kTagSceneShift) | (live ? 0u : kFinalFlag))));
            c.SQ.c[e] = make_float4(contrib.x, contrib.y, contrib.z, __int_as_float((int)(live ? o : sidx)));
        }
        if constexpr (!SINGLE) count_done(done_cnt, done_tag, c.job_done, fin, jid);
    }
    if constexpr (!SINGLE) flush_done(done_cnt, done_tag, c.job_done);
}

// Any-hit traversal of the shadow rays, "while-while" like extend. An unoccluded ray adds its
// contribution to the survivor's radiance in `out`, or to the sample of a path that ended with it.
// At most 64 VGPRs (8 waves per SIMD): the multi-job form needs 69 otherwise, which fits three
// 512-thread workgroups per CU instead of four; capped, it keeps 20 bytes per lane in scratch and
// the 64-job batch takes 29.5 instead of 31.1 ms. The single-job form is at 63 either way.
template <bool SINGLE>
__global__ __launch_bounds__(1024) __attribute__((amdgpu_waves_per_eu(8))) void pw_shadow_kernel(PCtx c, PathBuf out) {
    extern __shared__ Node lds[];
    __shared__ LdsPlan sp;
    __shared__ uint32_t pre[kSegs + 1], win[kSegs], shc[kSegs], svc[kSegs];
    __shared__ JobTab jt;                   // multi-job only
    __shared__ uint32_t done_cnt[kLdsJobs];   // multi-job only
    __shared__ int done_tag[kLdsJobs];        // multi-job only
    __shared__ uint32_t blk_unocc, feed_ctr;
    PState* st = c.st;
    const uint32_t slot0 = single_slot(st);
    if (threadIdx.x < kSegs) {
        shc[threadIdx.x] = st->sh_cnt[threadIdx.x];
        svc[threadIdx.x] = st->seg_cnt[threadIdx.x];
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        feed_ctr = blockDim.x >> 6;
        // The shadow rays of this iteration, segment by segment (windows of the shade pass).
        uint32_t acc = 0, next = 0;
        for (int s = 0; s < kSegs; ++s) {
            pre[s] = acc;
            acc += shc[s];
            next += svc[s];
        }
        pre[kSegs] = acc;
        seg_windows(st->it_act, st->shade_block, win);
        if (blockIdx.x == 0) {
            st->n_next = next;
            st->n_shadow = acc;
            // A single job: every path that did not survive this iteration is finished once this
            // kernel is done (a pending shadow ray is the last thing it waits for).
            if constexpr (SINGLE) c.job_done[slot0] += st->it_act - next;
        }
    }
    __syncthreads();
    const uint32_t n = pre[kSegs];
    const uint32_t fsize = kShadowSlice;
    if (feed_idle(n, fsize)) return;
    if constexpr (!SINGLE) stage_jobs(c, st, &jt, done_cnt, done_tag);
    stage_lds<SINGLE>(c, lds, &sp);
    BlockFeed f;
    feed_first(f, n, fsize);
    LdsNodes src{};
    MultiSrc msrc = multi_src(&jt, &sp, lds);
    int nn = 0;
    if constexpr (SINGLE) {
        src = src_of(c.scene0, lds, &sp, 0);
        nn = c.scene0.n_nodes;
    }
    uint32_t item = kIdle, pos = 0, tag = 0;
    uint32_t unocc = 0;   // unoccluded rays of this wave (scalar), for the byte count of a profiled iteration
    if (threadIdx.x == 0) blk_unocc = 0u;
    Traversal tr;
    tr.node = -1;
    tr.prim = -1;
    int leaf = kNoLeaf;
    for (;;) {
        const uint64_t idle = __ballot(item == kIdle);
        if (idle != 0 && f.cursor >= f.end && !f.dry) feed_next(f, n, &feed_ctr);
        if (idle != 0 && f.cursor < f.end) {
            const uint32_t cand = f.cursor + mbcnt64(idle);
            if (item == kIdle && cand < f.end) {
                item = cand;
                pos = seg_pos(pre, win, item);
                const float4 o = c.SQ.o[pos], d = c.SQ.d[pos];
                tag = (uint32_t)__float_as_int(d.w);
                Ray r;
                r.o = Vec3{o.x, o.y, o.z};
                r.d = Vec3{d.x, d.y, d.z};
                r.tmin = kRayEps;
                r.tmax = o.w;
                if constexpr (!SINGLE) {
                    const int sc = (int)((tag >> kTagSceneShift) & 7u);
                    msrc.set(lds, sc);
                    nn = jt.scenes[sc].n_nodes;
                }
                tr = traversal_begin(r);
                if (nn <= 0) tr.node = -1;
            }
            f.cursor = min(f.cursor + (uint32_t)__popcll(idle), f.end);
        }
        if (__ballot(item != kIdle) == 0) {
            if (f.dry) break;
            continue;   // slice used up: take the next one
        }
        const int limit = (f.dry && f.cursor >= f.end) ? 1 : c.refill;
        if constexpr (SINGLE) wave_traverse(nn, src, tr, leaf, true, limit, c.leaf_batch);
        else wave_traverse(nn, msrc, tr, leaf, true, limit, c.leaf_batch);
        bool fin = false, clear = false;
        uint32_t jid = 0;
        if (item != kIdle && tr.node < 0 && !leaf_parked(leaf)) {
            fin = (tag & kFinalFlag) != 0;
            if constexpr (!SINGLE) jid = tag & kJobMask;
            if (tr.prim < 0) {
                clear = true;
                const float4 cc = c.SQ.c[pos];
                const uint32_t dest = (uint32_t)__float_as_int(cc.w);
                float4* dst = fin ? &job_of<SINGLE>(c, jid, slot0).sbuf[dest] : &out.L[dest];
                float4 l = *dst;
                l.x += cc.x;
                l.y += cc.y;
                l.z += cc.z;
                *dst = l;
            }
            item = kIdle;
        }
        unocc += (uint32_t)__popcll(__ballot(clear));
        if constexpr (!SINGLE) count_done(done_cnt, done_tag, c.job_done, fin, jid);
    }
    if (lane_id() == 0 && unocc) atomicAdd(&blk_unocc, unocc);
    __syncthreads();
    if (threadIdx.x == 0 && blk_unocc) atomicAdd(&st->unocc, blk_unocc);
    if constexpr (!SINGLE) flush_done(done_cnt, done_tag, c.job_done);
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

uint64_t job_pixels(const PersistentJob& j) {
    return j.has_view ? (uint64_t)j.width * (uint64_t)j.height : (uint64_t)j.scene->width * j.scene->height;
}

// Large device blocks (path state, sample arenas) of engines that have ended, kept for the next
// engine that asks for the same size on the same device. A work-sized pool is up to 12 GB in a
// dozen blocks; freeing and allocating that much for every batch render made the device stall
// now and then for 0.1-2 s (a batch render's teardown went 15 -> 335 ms, its render 34 ms -> 2.2 s,
// after about 16 renders in one process). Blocks below kMinBytes go straight back to the runtime;
// at most kMaxHeld bytes are kept, per process; wf_release_cached_memory returns them.
class BlockCache {
public:
    static constexpr size_t kMinBytes = 64ull << 20, kMaxHeld = 32ull << 30;

    static void* get(size_t bytes, const char* what) {
        if (bytes >= kMinBytes) {
            int dev = 0;
            check(hipGetDevice(&dev), "hipGetDevice");
            std::lock_guard<std::mutex> lock(mu());
            std::vector<Entry>& v = entries();
            for (size_t i = 0; i < v.size(); ++i)
                if (v[i].dev == dev && v[i].bytes == bytes) {
                    void* p = v[i].p;
                    held() -= bytes;
                    v.erase(v.begin() + (std::ptrdiff_t)i);
                    return p;
                }
        }
        void* q = nullptr;
        check(hipMalloc(&q, bytes), what);
        return q;
    }

    // No kernel may still use p (the caller has synchronised the device).
    static void put(void* p, size_t bytes) {
        if (!p) return;
        int dev = 0;
        if (bytes >= kMinBytes && hipGetDevice(&dev) == hipSuccess) {
            std::lock_guard<std::mutex> lock(mu());
            if (held() + bytes <= kMaxHeld) {
                entries().push_back(Entry{dev, bytes, p});
                held() += bytes;
                return;
            }
        }
        (void)hipFree(p);
    }

    // Frees every kept block (of every device); returns the bytes freed.
    static uint64_t release() {
        std::lock_guard<std::mutex> lock(mu());
        uint64_t n = 0;
        int cur = 0;
        (void)hipGetDevice(&cur);
        for (const Entry& e : entries()) {
            (void)hipSetDevice(e.dev);
            (void)hipFree(e.p);
            n += e.bytes;
        }
        (void)hipSetDevice(cur);
        entries().clear();
        held() = 0;
        return n;
    }

    static uint64_t held_bytes() {
        std::lock_guard<std::mutex> lock(mu());
        return held();
    }

private:
    struct Entry {
        int dev;
        size_t bytes;
        void* p;
    };
    // Never destroyed: an engine may end during process exit, after static destructors have run.
    static std::mutex& mu() {
        static std::mutex* m = new std::mutex;
        return *m;
    }
    static std::vector<Entry>& entries() {
        static std::vector<Entry>* v = new std::vector<Entry>;
        return *v;
    }
    static size_t& held() {
        static size_t h = 0;
        return h;
    }
};

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
    int block_multi = kDefaultBlockMulti, lds_multi = kDefaultLdsKiBMulti * 1024;   // multi-job mode
    int lds_max = 0;      // LDS a workgroup may use on this device (bytes)
    int lds_static = 0;   // allowance for the traversal kernels' own LDS (bytes)
    int refill = kDefaultRefill;
    int leaf_batch = kDefaultLeafBatch;
    int dir_bins = kDefaultDirBins;
    int shade_block = kDefaultShadeBlock;
    std::vector<std::pair<void*, size_t>> pool_allocs;   // blocks of the path pool and their sizes
    PState* st = nullptr;
    PState* st_host = nullptr;       // pinned mirror
    uint32_t* done_host = nullptr;   // pinned mirror of job_done
    PCtx ctx{};
    PathBuf S[2];
    int par = 0;   // buffer holding the survivors of the last iteration
    LdsPlan lds_plan{};
    struct HostJob {
        JobDev dev;
        uint64_t work = 0;
        uint32_t done = 0;
        int scene = 0;             // scene slot
        bool finished = false;
        bool owns_buffer = true;   // false: a slice of the render arena
    };
    // Ring of job slots: job q (sequence number) is ring[q & kJobMask]; the unfinished-or-younger
    // jobs are [seq_lo, seq_hi).
    std::vector<HostJob> ring = std::vector<HostJob>(kMaxJobs);
    int64_t seq_lo = 0, seq_hi = 0;
    // Scene slots: a slot's key is null once forgotten; refs counts the unfinished jobs on it.
    std::vector<SceneView> hscenes;
    std::vector<const SceneView*> scene_keys;
    std::vector<int> scene_refs;
    hipMemPool_t mem_pool = nullptr;   // sample buffers of admitted jobs (null: the default pool)
    unsigned long long total_work = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    // What a captured graph depends on: single/multi mode, the scene table (baked in single mode,
    // and the LDS size of every launch), the launch shape and scheduling parameters.
    struct GraphKey {
        bool single = false;
        int single_scene = -1;
        uint64_t scene_gen = 0;
        int block = 0, refill = 0, leaf_batch = 0, dir_bins = 0, shade_block = 0, cap = 0, iters = 0;
        size_t lds_bytes = 0;
        const void* ctr = nullptr;
        bool operator==(const GraphKey& o) const {
            return single == o.single && single_scene == o.single_scene && scene_gen == o.scene_gen &&
                   block == o.block && refill == o.refill &&
                   leaf_batch == o.leaf_batch && dir_bins == o.dir_bins && shade_block == o.shade_block && cap == o.cap && iters == o.iters && lds_bytes == o.lds_bytes &&
                   ctr == o.ctr;
        }
    } graph_key;
    uint64_t scene_gen = 1;   // bumped whenever the scene table changes
    uint64_t sbuf_bytes = 0;
    // render(): one allocation holds the sample buffers of a chunk of jobs; kept across renders.
    float4* arena = nullptr;
    uint64_t arena_bytes = 0;
    // Progress snapshots (snapshot/collect): pinned copies and the event that marks them done.
    struct Snap {
        PState* st = nullptr;
        uint32_t* done = nullptr;
        DevCounters* ctr = nullptr;
        hipEvent_t ev = nullptr;
        int64_t seq_hi = 0;     // jobs admitted when the snapshot was taken
        uint64_t jobs_gen = 0;
        bool with_ctr = false, outstanding = false;
    } snap[2];
    int snap_next = 0;
    uint64_t jobs_gen = 0;   // bumped by reset_jobs: a snapshot of an older job table applies to none

    void free_graph() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        exec = nullptr;
        graph = nullptr;
    }
    // Blocks go to BlockCache once no kernel can still use them.
    void free_arena() {
        if (arena) {
            (void)hipDeviceSynchronize();
            BlockCache::put(arena, (size_t)arena_bytes);
        }
        arena = nullptr;
        arena_bytes = 0;
    }
    void free_pool() {
        if (!pool_allocs.empty()) (void)hipDeviceSynchronize();
        for (const auto& [p, bytes] : pool_allocs) BlockCache::put(p, bytes);
        pool_allocs.clear();
    }
    ~Impl() {
        free_graph();
        for (int64_t q = seq_lo; q < seq_hi; ++q) {
            HostJob& j = ring[q & kJobMask];
            if (j.dev.sbuf && j.owns_buffer) (void)hipFree(j.dev.sbuf);   // the caller's stream may be gone
        }
        free_arena();
        free_pool();
        if (st_host) (void)hipHostFree(st_host);
        if (done_host) (void)hipHostFree(done_host);
        for (Snap& s : snap) {
            if (s.ev) (void)hipEventSynchronize(s.ev);
            if (s.ev) (void)hipEventDestroy(s.ev);
            if (s.st) (void)hipHostFree(s.st);
            if (s.done) (void)hipHostFree(s.done);
            if (s.ctr) (void)hipHostFree(s.ctr);
        }
    }
};

struct PersistentEngine::LaunchPlan {
    PCtx c;
    bool single;
    size_t lds_bytes;
    int block;   // traversal workgroup size of the mode
    int gx, gs, gw;
};

PersistentEngine::PersistentEngine() : m_(std::make_unique<Impl>()) {}
PersistentEngine::~PersistentEngine() = default;

uint64_t PersistentEngine::bytes() const { return pool_bytes_ + m_->sbuf_bytes + m_->arena_bytes; }

void PersistentEngine::ensure_pool(const PersistentTuning& t) {
    Impl& m = *m_;
    if (m.cus == 0) {
        int dev = 0;
        check(hipGetDevice(&dev), "hipGetDevice");
        check(hipDeviceGetAttribute(&m.cus, hipDeviceAttributeMultiprocessorCount, dev), "hipDeviceGetAttribute");
        check(hipDeviceGetAttribute(&m.lds_max, hipDeviceAttributeMaxSharedMemoryPerBlock, dev),
              "hipDeviceGetAttribute");
        m.lds_static = kLdsStatic;
    }
    int cap = t.pool_paths > 0 ? t.pool_paths : kDefaultPool;
    cap = (cap + 1023) & ~1023;
    m.block = t.block_size > 0 ? t.block_size : kDefaultBlock;
    m.block_multi = t.block_size > 0 ? t.block_size : kDefaultBlockMulti;
    if (m.block != 256 && m.block != 512 && m.block != 1024)
        throw std::runtime_error("persistent: block size must be 256, 512 or 1024");
    m.refill = t.refill > 0 ? t.refill : kDefaultRefill;
    if (m.refill > 64) throw std::runtime_error("persistent: refill threshold must be 1..64 lanes");
    m.leaf_batch = t.leaf_batch > 0 ? t.leaf_batch : kDefaultLeafBatch;
    if (m.leaf_batch > 64) throw std::runtime_error("persistent: leaf batch must be 1..64 lanes");
    m.dir_bins = t.dir_bins > 0 ? t.dir_bins : kDefaultDirBins;
    m.shade_block = t.shade_block > 0 ? t.shade_block : kDefaultShadeBlock;
    if (m.shade_block != 128 && m.shade_block != 256 && m.shade_block != 512 && m.shade_block != 1024)
        throw std::runtime_error("persistent: shade block size must be 128, 256, 512 or 1024");
    if (m.dir_bins != 1 && m.dir_bins != 8) throw std::runtime_error("persistent: direction bins must be 1 or 8");
    m.ctx.refill = m.refill;
    m.ctx.leaf_batch = m.leaf_batch;
    m.ctx.dir_bins = m.dir_bins;
    if (t.graph_iters > 1024) throw std::runtime_error("persistent: graph iterations must be 1..1024");
    m.graph_iters = t.graph_iters > 0 ? ((t.graph_iters + 1) & ~1) : kDefaultGraphIters;
    const int kib = t.lds_kib > 0 ? t.lds_kib : kDefaultLdsKiB;
    const int kib_multi = t.lds_kib > 0 ? t.lds_kib : kDefaultLdsKiBMulti;
    // The budget is the dynamic LDS; the kernels' own (the plan and the segment tables)
    // comes on top.
    if (kib * 1024 > m.lds_max - m.lds_static)
        throw std::runtime_error("persistent: LDS budget above what a workgroup may use on this device");
    m.lds_bytes = kib * 1024;   // single mode plans its scene's LDS per launch plan
    if (m.lds_multi != kib_multi * 1024) {
        m.lds_multi = kib_multi * 1024;
        if (!m.hscenes.empty()) {
            update_lds_plan();
            m.scene_gen += 1;
        }
    }
    if (cap == cap_ && !m.pool_allocs.empty()) return;
    m.free_graph();
    m.free_pool();
    // The device scene table lives in the pool allocation (open() then clears the job table).
    m.scene_keys.clear();
    m.hscenes.clear();
    m.scene_refs.clear();
    m.scene_gen += 1;
    pool_bytes_ = 0;
    cap_ = cap;
    auto alloc = [&](auto** p, size_t bytes) {
        void* q = BlockCache::get(bytes, "hipMalloc persistent pool");
        m.pool_allocs.push_back({q, bytes});
        *p = (std::remove_reference_t<decltype(**p)>*)q;
        pool_bytes_ += bytes;
    };
    const size_t P = (size_t)cap;
    PCtx& c = m.ctx;
    c = PCtx{};
    c.refill = m.refill;
    c.leaf_batch = m.leaf_batch;
    c.dir_bins = m.dir_bins;
    for (PathBuf& b : m.S) {
        alloc(&b.ro, P * 16);
        alloc(&b.rd, P * 16);
        alloc(&b.T, P * 16);
        alloc(&b.L, P * 16);
    }
    alloc(&c.hit, P * 8);
    alloc(&c.SQ.o, P * 16);
    alloc(&c.SQ.d, P * 16);
    alloc(&c.SQ.c, P * 16);
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
    for (int64_t q = m.seq_lo; q < m.seq_hi; ++q) {
        Impl::HostJob& j = m.ring[q & kJobMask];
        if (j.dev.sbuf && j.owns_buffer && !j.finished) (void)hipFreeAsync(j.dev.sbuf, m.stream);
        j = Impl::HostJob{};
    }
    m.seq_lo = m.seq_hi = 0;
    for (int& r : m.scene_refs) r = 0;
    m.jobs_gen += 1;
    m.sbuf_bytes = 0;
    m.total_work = 0;
    pw_init_kernel<<<1, 1, 0, m.stream>>>(m.st);
    m.par = 0;
    check(hipGetLastError(), "persistent init");
}

// Splits the LDS budget over the scenes. Smallest need first, a scene gets its whole tree (then
// its primitive records, if they fit too) when that leaves kMinTopBytes for each other scene not
// yet whole (those after it and those already left partial):
// a whole tree runs the all-LDS visit, worth more than a few more top nodes of a large tree. The
// scenes left over share the rest evenly and store the breadth-first top of their trees.
constexpr long long kMinTopBytes = 4 * 1024;

LdsPlan plan_lds(const int* n_nodes, const int* n_prims, int n, long long budget) {
    if (n < 0 || n > kMaxScenes) throw std::invalid_argument("lds plan: scene count outside 0..8");
    LdsPlan p{};
    p.n = n;
    if (n == 0) return p;
    auto need = [&](int k) { return (long long)n_nodes[k] * (long long)sizeof(Node) + (long long)n_prims[k] * 16; };
    std::vector<int> order(n);
    for (int k = 0; k < n; ++k) order[k] = k;
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return need(a) < need(b); });
    long long left = budget;
    std::vector<int> rest;
    for (int i = 0; i < n; ++i) {
        const int k = order[i];
        // What stays for the scenes after this one and for those already left partial.
        const long long avail = left - kMinTopBytes * (long long)(n - 1 - i + (int)rest.size());
        const long long nodes = (long long)n_nodes[k] * (long long)sizeof(Node);
        if (nodes > avail) {
            rest.push_back(k);
            continue;
        }
        p.node_cnt[k] = n_nodes[k];
        p.prim_cnt[k] = (long long)n_prims[k] * 16 <= avail - nodes ? n_prims[k] : 0;
        left -= nodes + (long long)p.prim_cnt[k] * 16;
    }
    for (size_t i = 0; i < rest.size(); ++i) {
        const int k = rest[i];
        const long long share = left / (long long)(rest.size() - i);
        p.node_cnt[k] = (int)std::min<long long>(n_nodes[k], std::max<long long>(share, 0) / (long long)sizeof(Node));
        p.prim_cnt[k] = 0;
        left -= (long long)p.node_cnt[k] * (long long)sizeof(Node);
    }
    int nodes = 0, prims = 0;
    for (int k = 0; k < n; ++k) {
        p.node_off[k] = nodes;
        p.prim_off[k] = prims;
        nodes += p.node_cnt[k];
        prims += p.prim_cnt[k];
    }
    p.total_nodes = nodes;
    p.total_prims = prims;
    return p;
}

void PersistentEngine::update_lds_plan() {
    Impl& m = *m_;
    int nn[kMaxScenes] = {}, np[kMaxScenes] = {};
    const int n = (int)m.hscenes.size();
    for (int k = 0; k < n; ++k) {   // a forgotten slot's record is zero: it gets no LDS
        nn[k] = m.hscenes[k].n_nodes;
        np[k] = m.hscenes[k].n_prims;
    }
    m.lds_plan = plan_lds(nn, np, n, m.lds_multi);
    pw_put_kernel<LdsPlan><<<1, 1, 0, m.stream>>>(m.ctx.plan, m.lds_plan);
    check(hipGetLastError(), "put lds plan");
}

int PersistentEngine::scene_slot(const SceneView* scene) const {
    const Impl& m = *m_;
    if (!scene) return -1;
    for (size_t k = 0; k < m.scene_keys.size(); ++k)
        if (m.scene_keys[k] == scene && std::memcmp(&m.hscenes[k], scene, sizeof(SceneView)) == 0) return (int)k;
    return -1;
}

// scene null: whether a scene not yet in the table would find a slot.
bool PersistentEngine::can_admit(const SceneView* scene) const {
    const Impl& m = *m_;
    if (m.seq_hi - m.seq_lo >= kMaxJobs) return false;
    if (scene_slot(scene) >= 0 || (int)m.scene_keys.size() < kMaxScenes) return true;
    for (int r : m.scene_refs)
        if (r == 0) return true;
    return false;
}

int64_t PersistentEngine::live_jobs() const { return m_->seq_hi - m_->seq_lo; }

void PersistentEngine::set_mem_pool(hipMemPool_t pool) { m_->mem_pool = pool; }

// The record of a forgotten slot is zero (no nodes, no primitives), so no kernel stages or reads
// the scene's device memory once the table update has run in stream order.
void PersistentEngine::forget_scene(const SceneView* scene) {
    Impl& m = *m_;
    const int k = scene_slot(scene);
    if (k < 0) return;
    if (m.scene_refs[k] > 0) throw std::logic_error("persistent: forgetting a scene that unfinished jobs use");
    m.scene_keys[k] = nullptr;
    m.hscenes[k] = SceneView{};
    pw_put_kernel<SceneView><<<1, 1, 0, m.stream>>>(m.ctx.scenes + k, m.hscenes[k]);
    check(hipGetLastError(), "put scene record");
    update_lds_plan();
    m.scene_gen += 1;
}

// Validates a job, enters its scene in the scene table and returns its device record, its work
// range following the jobs admitted so far, and its sequence number in *seq; sbuf: its sample
// buffer, or null to allocate one.
JobDev PersistentEngine::prepare_job(const PersistentJob& job, float4* sbuf, int64_t* seq) {
    Impl& m = *m_;
    if (m.seq_hi - m.seq_lo >= kMaxJobs) throw std::runtime_error("persistent: job table full");
    if (job.spp < 1 || job.scene == nullptr) throw std::runtime_error("persistent: invalid job");
    if (job.max_depth < 0 || job.max_depth > kMaxDepth) throw std::runtime_error("persistent: max_depth out of range");
    const int jw = job.has_view ? job.width : job.scene->width, jh = job.has_view ? job.height : job.scene->height;
    if (jw < 1 || jh < 1) throw std::runtime_error("persistent: job image size must be positive");
    const uint64_t WH = (uint64_t)jw * (uint64_t)jh;
    const uint64_t work = WH * (uint64_t)job.spp;
    if (work >= (1ull << 32)) throw std::runtime_error("persistent: job has 2^32 or more paths");
    // Scene table: kept across job resets; an entry matches by address and by contents (a view
    // rebuilt at the same address is a different scene). A new scene takes a free slot: one past
    // the end, or one that no unfinished job uses.
    int sc = scene_slot(job.scene);
    if (sc < 0) {
        if ((int)m.scene_keys.size() < kMaxScenes) {
            sc = (int)m.scene_keys.size();
            m.scene_keys.push_back(nullptr);
            m.hscenes.push_back(SceneView{});
            m.scene_refs.push_back(0);
        } else {
            for (int k = 0; k < kMaxScenes && sc < 0; ++k)
                if (m.scene_refs[k] == 0) sc = k;
            if (sc < 0) throw std::runtime_error("persistent: too many distinct scenes in flight");
        }
        m.scene_keys[sc] = job.scene;
        m.hscenes[sc] = *job.scene;
        pw_put_kernel<SceneView><<<1, 1, 0, m.stream>>>(m.ctx.scenes + sc, *job.scene);
        check(hipGetLastError(), "put scene record");
        update_lds_plan();
        m.scene_gen += 1;
    }
    JobDev d{};
    d.scene = sc;
    d.width = jw;
    d.height = jh;
    d.cam = job.has_view ? job.cam : job.scene->camera;
    d.max_depth = job.max_depth;
    d.first_sample = (uint32_t)job.first_sample;
    d.spp = (uint32_t)job.spp;
    d.key_a = path_key_a(job.seed);
    d.key_b = path_key_b(job.seed);
    d.work_begin = m.total_work;
    d.work_end = m.total_work + work;
    d.div_pixels = make_divu32((uint32_t)WH);
    d.div_band = make_divu32(8u * (uint32_t)jw);
    d.film = job.film;
    Impl::HostJob hj;
    if (sbuf) {
        d.sbuf = sbuf;
        hj.owns_buffer = false;
    } else {
        // Stream-ordered: no device-wide synchronisation, and finish_job's free returns the memory
        // to the pool for the next job in stream order.
        const size_t bytes = (size_t)work * sizeof(float4);
        if (m.mem_pool) check(hipMallocFromPoolAsync((void**)&d.sbuf, bytes, m.mem_pool, m.stream), "hipMallocFromPoolAsync sample buffer");
        else check(hipMallocAsync((void**)&d.sbuf, bytes, m.stream), "hipMallocAsync sample buffer");
        m.sbuf_bytes += (uint64_t)work * sizeof(float4);
    }
    m.total_work = d.work_end;
    m.scene_refs[sc] += 1;
    hj.dev = d;
    hj.work = work;
    hj.scene = sc;
    *seq = m.seq_hi++;
    m.ring[*seq & kJobMask] = hj;
    return d;
}

int64_t PersistentEngine::admit(const PersistentJob& job) {
    Impl& m = *m_;
    int64_t seq = 0;
    const JobDev d = prepare_job(job, nullptr, &seq);
    pw_admit_kernel<<<1, 1, 0, m.stream>>>(m.st, m.ctx.jobs, m.ctx.job_done, d, (uint32_t)seq);
    check(hipGetLastError(), "persistent admit");
    return seq;
}

// Admits a chunk of jobs with their sample buffers in the arena (one allocation, kept across
// renders, instead of an allocation per job).
void PersistentEngine::admit_chunk(const std::vector<PersistentJob>& chunk) {
    Impl& m = *m_;
    uint64_t need = 0;
    for (const PersistentJob& j : chunk) need += job_pixels(j) * (uint64_t)j.spp * sizeof(float4);
    if (need > m.arena_bytes) {
        m.free_arena();
        // Whole 64 MiB units, so that batches of similar sizes reuse each other's cached arena.
        const uint64_t unit = 64ull << 20;
        const uint64_t bytes = (need + unit - 1) / unit * unit;
        m.arena = (float4*)BlockCache::get((size_t)bytes, "hipMalloc sample arena");
        m.arena_bytes = bytes;
    }
    uint64_t off = 0;
    for (const PersistentJob& j : chunk) {
        int64_t seq = 0;
        const JobDev d = prepare_job(j, m.arena + off, &seq);
        off += job_pixels(j) * (uint64_t)j.spp;
        pw_admit_kernel<<<1, 1, 0, m.stream>>>(m.st, m.ctx.jobs, m.ctx.job_done, d, (uint32_t)seq);
        check(hipGetLastError(), "persistent admit");
    }
}

// Single mode (one unfinished job) plans the LDS for that job's scene alone and bakes its record
// into the launches; multi mode reads the scene table and the plan of every slot from memory.
PersistentEngine::LaunchPlan PersistentEngine::make_plan(bool single) {
    Impl& m = *m_;
    LaunchPlan L{};
    L.single = single;
    L.c = m.ctx;
    L.c.plan0 = m.lds_plan;
    const LdsPlan* plan = &m.lds_plan;
    if (single) {
        const SceneView& sv = m.hscenes[m.ring[(m.seq_hi - 1) & kJobMask].scene];
        L.c.scene0 = sv;
        L.c.plan0 = plan_lds(&sv.n_nodes, &sv.n_prims, 1, m.lds_bytes);
        plan = &L.c.plan0;
    }
    L.lds_bytes = (size_t)plan->total_nodes * sizeof(Node) + (size_t)plan->total_prims * sizeof(float4);
    const void* fx = single ? (const void*)pw_extend_kernel<true> : (const void*)pw_extend_kernel<false>;
    const void* fs = single ? (const void*)pw_shade_kernel<true> : (const void*)pw_shade_kernel<false>;
    const void* fw = single ? (const void*)pw_shadow_kernel<true> : (const void*)pw_shadow_kernel<false>;
    if (L.lds_bytes > 64 * 1024)
        for (const void* fn : {fx, fw})
            check(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)L.lds_bytes),
                  "hipFuncSetAttribute");
    L.block = single ? m.block : m.block_multi;
    L.gx = grid_for(fx, L.block, L.lds_bytes, m.cus);
    L.gs = grid_for(fs, m.shade_block, 0, m.cus);
    L.gw = grid_for(fw, L.block, L.lds_bytes, m.cus);
    return L;
}

// One iteration reading survivors from buffer `par` and writing them to the other; with `ev`
// (5 events) the stage boundaries are recorded for profiling:
// ev[0] plan ev[1] extend ev[2] shade ev[3] shadow ev[4].
void PersistentEngine::launch_iteration(const LaunchPlan& L, int par, hipEvent_t* ev) {
    Impl& m = *m_;
    hipStream_t st = m.stream;
    const PCtx& c = L.c;
    const PathBuf& in = m.S[par & 1];
    const PathBuf& out = m.S[(par & 1) ^ 1];
    if (ev) check(hipEventRecord(ev[0], st), "hipEventRecord");
    pw_plan_kernel<<<1, 1, 0, st>>>(m.st, c.jobs, m.ctr, c.cap, (uint32_t)m.shade_block);
    if (ev) check(hipEventRecord(ev[1], st), "hipEventRecord");
    if (L.single) pw_extend_kernel<true><<<L.gx, L.block, L.lds_bytes, st>>>(c, in);
    else pw_extend_kernel<false><<<L.gx, L.block, L.lds_bytes, st>>>(c, in);
    if (ev) check(hipEventRecord(ev[2], st), "hipEventRecord");
    if (L.single) pw_shade_kernel<true><<<L.gs, m.shade_block, 0, st>>>(c, in, out);
    else pw_shade_kernel<false><<<L.gs, m.shade_block, 0, st>>>(c, in, out);
    if (ev) check(hipEventRecord(ev[3], st), "hipEventRecord");
    if (L.single) pw_shadow_kernel<true><<<L.gw, L.block, L.lds_bytes, st>>>(c, out);
    else pw_shadow_kernel<false><<<L.gw, L.block, L.lds_bytes, st>>>(c, out);
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
}

void PersistentEngine::run(int launches) {
    Impl& m = *m_;
    if (m.seq_hi == m.seq_lo) return;
    Impl::GraphKey k;
    k.single = m.seq_hi - m.seq_lo == 1;
    k.single_scene = k.single ? m.ring[(m.seq_hi - 1) & kJobMask].scene : -1;
    k.scene_gen = m.scene_gen;
    k.block = k.single ? m.block : m.block_multi;
    k.refill = m.refill;
    k.leaf_batch = m.leaf_batch;
    k.dir_bins = m.dir_bins;
    k.shade_block = m.shade_block;
    k.cap = cap_;
    k.iters = m.graph_iters;
    k.ctr = m.ctx.ctr;
    k.lds_bytes = (size_t)m.lds_plan.total_nodes * sizeof(Node) + (size_t)m.lds_plan.total_prims * sizeof(float4);
    k.lds_bytes = k.single ? (size_t)m.lds_bytes : k.lds_bytes;   // single mode plans within the budget
    if (!m.exec || !(k == m.graph_key)) {
        capture(k.single);
        m.graph_key = k;
    }
    for (int i = 0; i < launches; ++i) check(hipGraphLaunch(m.exec, m.stream), "hipGraphLaunch");
}

// Eager iterations with HIP events around each stage (synchronising once per iteration), until
// the admitted work is done. Items per stage come from the control block of each iteration.
void PersistentEngine::run_profiled(StageTimes* out) {
    Impl& m = *m_;
    if (m.seq_hi == m.seq_lo) return;
    const bool single = m.seq_hi - m.seq_lo == 1;
    const LaunchPlan L = make_plan(single);
    hipEvent_t ev[5];
    for (auto& e : ev) check(hipEventCreate(&e), "hipEventCreate");
    for (;;) {
        launch_iteration(L, m.par, ev);
        m.par ^= 1;
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
        out->items[2] += s.it_act;
        out->items[5] += s.n_shadow;
        // Algorithmic bytes (PathBuf and ShadowQueue layouts; BVH and primitives come from LDS):
        //   extend  survivor ray 32 (new paths are generated), hit 8
        //   shade   hit 8; survivor state 64 in, 64 out; sample 16 per finished path; shadow ray 48
        //   shadow  ray 32 (origin, direction) and, for an unoccluded ray, its contribution 16 and
        //           the radiance read-modify-write 32
        const long long live = s.it_live, act = s.it_act, next = s.n_next, sh = s.n_shadow, un = s.unocc;
        out->bytes[0] += 32 * live + 8 * act;
        out->bytes[2] += 8 * act + 64 * live + 64 * next + 16 * (act - next) + 48 * sh;
        out->bytes[5] += 32 * sh + 48 * un;
        out->items[7] += 1;
        out->launches[0] += 1;
        out->launches[2] += 1;
        out->launches[5] += 1;
        out->threads[0] = (long long)L.gx * L.block;
        out->threads[2] = (long long)L.gs * m.shade_block;
        out->threads[5] = (long long)L.gw * L.block;
        out->launches[7] += 1;
    }
    // The captured graph runs an even number of iterations from buffer 0; the work has drained,
    // so one more (empty) iteration restores that parity.
    if (m.par) {
        launch_iteration(L, m.par, nullptr);
        m.par = 0;
    }
    for (auto& e : ev) (void)hipEventDestroy(e);
}

PersistentProgress PersistentEngine::progress() {
    Impl& m = *m_;
    check(hipMemcpyAsync(m.st_host, m.st, sizeof(PState), hipMemcpyDeviceToHost, m.stream), "copy progress");
    if (m.seq_hi > m.seq_lo)
        check(hipMemcpyAsync(m.done_host, m.ctx.job_done, kMaxJobs * sizeof(uint32_t), hipMemcpyDeviceToHost, m.stream),
              "copy job progress");
    check(hipStreamSynchronize(m.stream), "persistent sync");
    for (int64_t q = m.seq_lo; q < m.seq_hi; ++q)
        if (!m.ring[q & kJobMask].finished) m.ring[q & kJobMask].done = m.done_host[q & kJobMask];
    PersistentProgress p;
    p.next_work = m.st_host->next_work;
    p.total_work = m.st_host->total_work;
    p.n_live = m.st_host->n_live;
    p.done = m.st_host->done;
    p.iterations = m.st_host->iterations;
    return p;
}

int PersistentEngine::snapshot(const DevCounters* ctr) {
    Impl& m = *m_;
    if (!m.st) throw std::logic_error("persistent snapshot: engine not open");
    const int slot = m.snap_next;
    Impl::Snap& s = m.snap[slot];
    if (s.outstanding) throw std::logic_error("persistent snapshot: both slots are waiting to be collected");
    if (!s.st) {
        check(hipHostMalloc(&s.st, sizeof(PState), hipHostMallocDefault), "hipHostMalloc");
        check(hipHostMalloc(&s.done, sizeof(uint32_t) * kMaxJobs, hipHostMallocDefault), "hipHostMalloc");
        check(hipHostMalloc(&s.ctr, sizeof(DevCounters), hipHostMallocDefault), "hipHostMalloc");
        check(hipEventCreateWithFlags(&s.ev, hipEventDisableTiming), "hipEventCreate");
    }
    s.seq_hi = m.seq_hi;
    s.jobs_gen = m.jobs_gen;
    s.with_ctr = ctr != nullptr;
    check(hipMemcpyAsync(s.st, m.st, sizeof(PState), hipMemcpyDeviceToHost, m.stream), "copy snapshot");
    if (m.seq_hi > m.seq_lo)
        check(hipMemcpyAsync(s.done, m.ctx.job_done, kMaxJobs * sizeof(uint32_t), hipMemcpyDeviceToHost, m.stream),
              "copy snapshot");
    if (ctr) check(hipMemcpyAsync(s.ctr, ctr, sizeof(DevCounters), hipMemcpyDeviceToHost, m.stream), "copy snapshot");
    check(hipEventRecord(s.ev, m.stream), "hipEventRecord snapshot");
    s.outstanding = true;
    m.snap_next ^= 1;
    return slot;
}

PersistentProgress PersistentEngine::collect(int slot, DevCounters* ctr_out) {
    Impl& m = *m_;
    if (slot < 0 || slot > 1 || !m.snap[slot].outstanding) throw std::logic_error("persistent collect: no such snapshot");
    Impl::Snap& s = m.snap[slot];
    check(hipEventSynchronize(s.ev), "snapshot wait");
    s.outstanding = false;
    // A job admitted after the snapshot is not in it; a slot reused since holds a younger job.
    if (s.jobs_gen == m.jobs_gen)
        for (int64_t q = m.seq_lo; q < s.seq_hi && q < m.seq_hi; ++q)
            if (!m.ring[q & kJobMask].finished) m.ring[q & kJobMask].done = s.done[q & kJobMask];
    if (ctr_out && s.with_ctr) *ctr_out = *s.ctr;
    PersistentProgress p;
    p.next_work = s.st->next_work;
    p.total_work = s.st->total_work;
    p.n_live = s.st->n_live;
    p.done = s.st->done;
    p.iterations = s.st->iterations;
    return p;
}

hipEvent_t PersistentEngine::snapshot_event(int slot) const { return m_->snap[slot & 1].ev; }

bool PersistentEngine::job_complete(int64_t job) const {
    const Impl& m = *m_;
    if (job < m.seq_lo) return true;   // finished and retired
    if (job >= m.seq_hi) throw std::logic_error("persistent: no such job");
    const Impl::HostJob& j = m.ring[job & kJobMask];
    return j.finished || (uint64_t)j.done == j.work;
}

// Jobs finish in any order; their slots return to the ring oldest first.
void PersistentEngine::retire_jobs() {
    Impl& m = *m_;
    while (m.seq_lo < m.seq_hi && m.ring[m.seq_lo & kJobMask].finished) {
        Impl::HostJob& j = m.ring[m.seq_lo & kJobMask];
        m.scene_refs[j.scene] -= 1;
        j = Impl::HostJob{};
        m.seq_lo += 1;
    }
}

void PersistentEngine::finish_job(int64_t job, hipStream_t stream) {
    Impl& m = *m_;
    if (job < m.seq_lo) return;
    if (job >= m.seq_hi) throw std::logic_error("persistent: no such job");
    Impl::HostJob& j = m.ring[job & kJobMask];
    if (j.finished) return;
    hipStream_t on = stream ? stream : m.stream;
    const int WH = j.dev.width * j.dev.height;
    pw_resolve_kernel<<<(WH + 255) / 256, 256, 0, on>>>(j.dev);
    check(hipGetLastError(), "persistent resolve");
    if (j.owns_buffer) {
        check(hipFreeAsync(j.dev.sbuf, on), "hipFreeAsync sample buffer");   // after the resolve
        m.sbuf_bytes -= (uint64_t)WH * j.dev.spp * sizeof(float4);
    }
    j.dev.sbuf = nullptr;
    j.finished = true;
    retire_jobs();
}

void PersistentEngine::render(const PersistentJob* jobs, int n_jobs, const PersistentTuning& t, DevCounters* ctr,
                              hipStream_t stream, PersistentResult* out, StageTimes* stages) {
    PersistentResult res;
    std::vector<PersistentJob> rest(jobs, jobs + n_jobs);
    uint64_t total = 0;
    for (PersistentJob& j : rest) {
        if (j.spp < 0 || j.scene == nullptr) throw std::runtime_error("persistent: invalid job");
        total += job_pixels(j) * (uint64_t)j.spp;
    }
    // No more slots than paths in the batch.
    PersistentTuning tt = t;
    const uint64_t want = tt.pool_paths > 0 ? (uint64_t)tt.pool_paths : (uint64_t)kDefaultPool;
    const uint64_t need = std::min<uint64_t>(want, std::max<uint64_t>(1024, (total + 1023) & ~1023ull));
    // A pool sized to the work is kept while later renders need no more: results do not depend on
    // the pool size, and a new pool would mean freeing and allocating up to 12 GB, uploading the
    // scene table again and capturing a new graph.
    const bool keep = t.pool_paths == 0 && !m_->pool_allocs.empty() && (uint64_t)cap_ >= need && cap_ <= kDefaultPool;
    tt.pool_paths = keep ? cap_ : (int)need;
    // When every sample fits the pool, all paths start in the first iteration and have ended after
    // max_depth + 1; the plan of the next iteration finds no work and marks the render done. A
    // graph of max_depth + 2 iterations then finishes in one launch without idle iterations (each
    // of which still launches four persistent grids).
    if (tt.graph_iters == 0 && total <= (uint64_t)tt.pool_paths) {
        int depth = 0;
        for (const PersistentJob& j : rest) depth = std::max(depth, j.max_depth);
        tt.graph_iters = std::min(kDefaultGraphIters, std::max(depth, 0) + 2);
    }
    open(tt, ctr, stream);
    // Split jobs into consecutive sample chunks so that one chunk's sample buffers fit the budget.
    for (;;) {
        std::vector<PersistentJob> chunk;
        uint64_t budget = sbuf_budget_;
        for (PersistentJob& j : rest) {
            if (j.spp <= 0) continue;
            const uint64_t per_sample = job_pixels(j) * sizeof(float4);
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
        admit_chunk(chunk);
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
        const int64_t lo = m_->seq_lo, hi = m_->seq_hi;
        for (int64_t q = lo; q < hi; ++q)
            if (!job_complete(q)) throw std::runtime_error("persistent: job incomplete after the work drained");
        for (int64_t q = lo; q < hi; ++q) finish_job(q);
        res.chunks += 1;
    }
    reset_jobs();
    if (out) *out = res;
}

PersistentTuning autotune_persistent(const PersistentJob* jobs, int n_jobs, hipStream_t stream, int probe_spp,
                                     std::vector<std::pair<PersistentTuning, double>>* trials, uint64_t state_budget) {
    DevCounters* ctr = nullptr;
    check(hipMalloc(&ctr, sizeof(DevCounters)), "hipMalloc counters");
    std::vector<PersistentJob> probe(jobs, jobs + n_jobs);
    uint64_t work = 0;
    for (PersistentJob& j : probe) {
        j.spp = std::min(j.spp, probe_spp);
        work += job_pixels(j) * j.spp;
    }
    // Queue capacity: the work-sized pool (every path of the batch in flight at once, up to
    // kDefaultPool) unless its path state exceeds the budget; then the largest power of two that
    // fits. Stage 5 times smaller pools too.
    uint64_t full = 0;
    for (int k = 0; k < n_jobs; ++k) full += job_pixels(jobs[k]) * (uint64_t)std::max(jobs[k].spp, 0);
    const uint64_t work_sized = std::min<uint64_t>(std::max<uint64_t>(1024, (full + 1023) & ~1023ull), kDefaultPool);
    auto fits = [&](uint64_t slots) { return state_budget == 0 || slots * PersistentEngine::kBytesPerSlot <= state_budget; };
    int base_pool = 0;
    if (!fits(work_sized)) {
        uint64_t p = 1024;
        if (!fits(p)) throw std::invalid_argument("autotune: the path-state budget holds fewer than 1024 slots");
        while (p * 2 < work_sized && fits(p * 2)) p *= 2;
        base_pool = (int)p;
    }
    PersistentEngine eng;
    hipEvent_t a, b;
    check(hipEventCreate(&a), "hipEventCreate");
    check(hipEventCreate(&b), "hipEventCreate");
    PersistentTuning best;
    double best_ms = 1e30;
    std::vector<std::pair<PersistentTuning, double>> seen;
    // The fastest of `reps` renders of `js` after one untimed render (graph capture).
    auto render_ms = [&](const PersistentJob* js, const PersistentTuning& t, int reps) {
        PersistentResult r;
        eng.render(js, n_jobs, t, ctr, stream, &r);
        float fastest = 1e30f;
        for (int k = 0; k < reps; ++k) {
            check(hipEventRecord(a, stream), "hipEventRecord");
            eng.render(js, n_jobs, t, ctr, stream, &r);
            check(hipEventRecord(b, stream), "hipEventRecord");
            check(hipEventSynchronize(b), "hipEventSynchronize");
            float ms = 0.0f;
            check(hipEventElapsedTime(&ms, a, b), "hipEventElapsedTime");
            fastest = std::min(fastest, ms);
        }
        return (double)fastest;
    };
    // Probe renders add into the jobs' films; the caller resets them afterwards.
    auto time = [&](const PersistentTuning& t) {
        const double ms = render_ms(probe.data(), t, 1);
        seen.push_back({t, ms});
        if (trials) trials->push_back({t, ms});
        if (ms < best_ms) {
            best_ms = ms;
            best = t;
        }
    };
    // Coordinate search. The pool is not searched: it is sized to the work of the render (up to
    // kDefaultPool), which the short probe renders cannot show, and more slots were faster at
    // every size measured. Stage 1: workgroup size.
    (void)work;
    for (int bs : {256, 512, 1024}) {
        PersistentTuning t;
        t.block_size = bs;
        t.refill = kDefaultRefill;
        t.leaf_batch = kDefaultLeafBatch;
        t.dir_bins = kDefaultDirBins;
        t.pool_paths = base_pool;
        t.lds_kib = 64;
        time(t);
    }
    // Stage 2: traversal scheduling (refill threshold x leaf batch) for that workgroup size, then
    // survivor grouping.
    PersistentTuning base = best;
    for (int rf : {16, 32, 48})
        for (int lb : {8, 16, 24, 32}) {
            if (rf == base.refill && lb == base.leaf_batch) continue;
            PersistentTuning t = base;
            t.refill = rf;
            t.leaf_batch = lb;
            time(t);
        }
    base = best;
    for (int sb : {256, 1024}) {
        PersistentTuning t = base;
        t.shade_block = sb;
        time(t);
    }
    base = best;
    for (int nb : {1, 8}) {
        if (nb == (base.dir_bins > 0 ? base.dir_bins : kDefaultDirBins)) continue;
        PersistentTuning t = base;
        t.dir_bins = nb;
        time(t);
    }
    // Stage 3: LDS budget, jointly with the workgroup size: the two decide together how many
    // workgroups fit a CU (a 128 KiB tree top admits one, 48 KiB three 512-thread ones), and the
    // multi-job kernels' 72 VGPRs leave room for 28 waves per CU, so a 512-thread workgroup with a
    // smaller tree top can beat the workgroup size stage 1 found at 64 KiB (the 64-job batch:
    // 512 threads with 48 KiB, 35.6 ms, against 1024 with 128 KiB, 38.4 ms).
    base = best;
    for (int bs : {512, 1024})
        for (int kib : {32, 40, 48, 64, 128}) {
            if (bs == base.block_size && kib == base.lds_kib) continue;
            PersistentTuning t = base;
            t.block_size = bs;
            t.lds_kib = kib;
            time(t);
        }
    // Stage 4: a probe's iterations hold less work than the whole batch's, and one probe render
    // is one sample of a noisy time, so the three fastest distinct candidates and the default
    // tuning are timed again on the whole batch (fastest of two renders each), and the fastest
    // of those is kept: the result is never measurably slower than the defaults.
    auto same = [](const PersistentTuning& x, const PersistentTuning& y) {
        return x.block_size == y.block_size && x.pool_paths == y.pool_paths && x.lds_kib == y.lds_kib &&
               x.graph_iters == y.graph_iters && x.refill == y.refill && x.leaf_batch == y.leaf_batch &&
               x.dir_bins == y.dir_bins && x.shade_block == y.shade_block;
    };
    std::stable_sort(seen.begin(), seen.end(), [](const auto& x, const auto& y) { return x.second < y.second; });
    std::vector<PersistentTuning> finalists;
    for (const auto& c : seen) {
        bool dup = false;
        for (const PersistentTuning& f : finalists) dup = dup || same(f, c.first);
        if (!dup) finalists.push_back(c.first);
        if (finalists.size() == 3) break;
    }
    PersistentTuning dflt;   // block size and LDS budget 0: the engine's defaults for the mode
    dflt.pool_paths = base_pool;
    dflt.refill = kDefaultRefill;
    dflt.leaf_batch = kDefaultLeafBatch;
    dflt.dir_bins = kDefaultDirBins;
    bool have_default = false;
    for (const PersistentTuning& f : finalists) have_default = have_default || same(f, dflt);
    if (!have_default) finalists.push_back(dflt);
    best_ms = 1e30;
    for (const PersistentTuning& f : finalists) {
        const double ms = render_ms(jobs, f, 2);
        if (ms < best_ms) {
            best_ms = ms;
            best = f;
        }
    }
    // Stage 5: queue capacity. The whole batch at 2^24 and 2^22 slots (where those are smaller
    // than the pool above and within the budget) against that pool: more slots mean fewer, fuller
    // iterations but more path state; the fastest is kept.
    const uint64_t top = base_pool ? (uint64_t)base_pool : work_sized;
    for (int p : {1 << 24, 1 << 22}) {
        if ((uint64_t)p >= top || !fits((uint64_t)p)) continue;
        PersistentTuning t = best;
        t.pool_paths = p;
        const double ms = render_ms(jobs, t, 2);
        if (trials) trials->push_back({t, ms});
        if (ms < best_ms) {
            best_ms = ms;
            best = t;
        }
    }
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    (void)hipFree(ctr);
    return best;
}

uint64_t cached_block_bytes() { return BlockCache::held_bytes(); }
uint64_t release_cached_blocks() { return BlockCache::release(); }

}  // namespace wf::gpu
