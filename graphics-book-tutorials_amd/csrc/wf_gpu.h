// Device-side building blocks shared by the megakernel, the wavefront engines and the queue
// primitives: device scene upload, film, counters, wave64 helpers and the LDS node source.
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "wf_core.h"
#include "wf_scene.h"

namespace wf::gpu {

inline void check(hipError_t e, const char* what) {
    if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

// Owns the device copies of a finalized scene's device-layout arrays.
struct DeviceScene {
    SceneView view{};
    uint64_t bytes = 0;
    std::vector<void*> allocs;
    void upload(const Scene& s);
    void release();
    ~DeviceScene() { release(); }
};

// Per-pixel float sums (rgb interleaved, then luminance and its square).
struct DeviceFilm {
    float* rgb = nullptr;
    float* lum = nullptr;
    float* lum2 = nullptr;
    int pixels = 0;
};

// Counters shared by every engine (src/gpu/engine.cuh:35-41). high_water order:
// lambert, metal, dielectric, shadow, active.
struct DevCounters {
    unsigned long long extension_rays;
    unsigned long long shadow_rays;
    unsigned int invalid_hits;
    unsigned int overflows;
    unsigned int high_water[5];
    unsigned int pad[3];
};

__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }

// Rank of this lane among the set lanes of a wave64 ballot.
__device__ __forceinline__ uint32_t mbcnt64(uint64_t m) {
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

__device__ __forceinline__ uint32_t wave_sum_u32(uint32_t v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

// Wave-aggregated queue push: the lanes with pred reserve consecutive slots with one atomic per
// wave (ballot + popcount, rank by mbcnt), then store `value`. Called by every active lane of the
// wave with the same control flow. A slot at or past `capacity` is dropped and counted in
// *overflow (one atomic per wave). Returns the slot or 0xFFFFFFFF. Contract as in
// src/gpu/queues.cuh:12-27, for 64-wide waves.
__device__ __forceinline__ uint32_t wave_push(uint32_t* items, uint32_t* counter, uint32_t capacity,
                                              uint32_t value, bool pred, uint32_t* overflow) {
    const uint64_t m = __ballot(pred);
    if (m == 0) return 0xFFFFFFFFu;
    const uint32_t n = (uint32_t)__popcll(m);
    const int leader = __ffsll((unsigned long long)m) - 1;
    uint32_t base = 0;
    if ((int)lane_id() == leader) {
        base = atomicAdd(counter, n);
        const uint32_t keep = base >= capacity ? 0u : (capacity - base < n ? capacity - base : n);
        if (keep < n) atomicAdd(overflow, n - keep);
    }
    base = __shfl(base, leader);
    if (!pred) return 0xFFFFFFFFu;
    const uint32_t slot = base + mbcnt64(m);
    if (slot >= capacity) return 0xFFFFFFFFu;
    items[slot] = value;
    return slot;
}

// Reserve `pred ? 1 : 0` slots per lane with one atomic per wave; returns the slot or ~0u.
__device__ __forceinline__ uint32_t wave_reserve(uint32_t* counter, bool pred) {
    const uint64_t m = __ballot(pred);
    if (m == 0) return 0xFFFFFFFFu;
    const int leader = __ffsll((unsigned long long)m) - 1;
    uint32_t base = 0;
    if ((int)lane_id() == leader) base = atomicAdd(counter, (uint32_t)__popcll(m));
    base = __shfl(base, leader);
    return pred ? base + mbcnt64(m) : 0xFFFFFFFFu;
}

// Adds per-lane counts to 64-bit device counters with one atomic per wave.
__device__ __forceinline__ void wave_add_u64(unsigned long long* dst, uint32_t v) {
    uint32_t s = wave_sum_u32(v);
    if (lane_id() == 0 && s) atomicAdd(dst, (unsigned long long)s);
}
__device__ __forceinline__ void wave_add_u32(unsigned int* dst, uint32_t v) {
    uint32_t s = wave_sum_u32(v);
    if (lane_id() == 0 && s) atomicAdd(dst, s);
}

// Address-space-qualified pointers to a native 4-float vector: loads through lds_float4 are
// ds_read instructions, through glb_float4 global loads (a generic pointer read from memory, such
// as a scene table entry, would otherwise give flat loads).
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef __attribute__((address_space(3))) const f32x4 lds_float4;
typedef __attribute__((address_space(1))) const f32x4 glb_float4;

static_assert(sizeof(Node) == 32 && offsetof(Node, skip) == 12 && offsetof(Node, child) == 28, "Node layout");
static_assert(sizeof(Prim) == 48 && offsetof(Prim, id) == 12 && offsetof(Prim, radius) == 28 &&
                  offsetof(Prim, material) == 44, "Prim layout");

// Node n of a device node array (32-byte records {lo, skip, hi, child}).
__device__ __forceinline__ Node load_node_global(const Node* nodes, int n) {
    const glb_float4* q = (const glb_float4*)nodes + 2 * n;
    const f32x4 a = q[0], b = q[1];
    Node out;
    out.lo = Vec3{a.x, a.y, a.z};
    out.skip = __float_as_int(a.w);
    out.hi = Vec3{b.x, b.y, b.z};
    out.child = __float_as_int(b.w);
    return out;
}

// Scene source that serves the first n_lds nodes (the breadth-first top of the tree) from LDS and,
// when has_lds_prims is set, every leaf primitive from a compact LDS record: centre and radius of a
// sphere, or the first vertex of a triangle with w = -1 (the other two vertices then come from
// memory). Spheres always have a positive radius (validated), so w < 0 marks a triangle.
// A node is one 32-byte record {lo, skip, hi, child}, read from LDS as two 16-byte loads. The
// scheduling barrier closing the LDS branch keeps the compiler from merging the LDS and memory
// loads into one flat load through a selected pointer (a flat load waits on both the LDS and the
// memory counters).
struct LdsNodes {
    lds_float4* lds;
    const Node* mem;
    int n_lds;
    const Prim* prims;
    int n_spheres;
    lds_float4* lds_prims;
    int has_lds_prims;
    __device__ __forceinline__ Node node_lds(int n) const {
        const f32x4 a = lds[2 * n], b = lds[2 * n + 1];
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
            out = load_node_global(mem, n);
        }
        return out;
    }
    // Every node of a tree of n_nodes is served from LDS.
    __device__ __forceinline__ bool all_lds(int n_nodes) const { return n_lds >= n_nodes; }
    // Prim records are 48 bytes: {p0, id}, {p1, radius}, {p2, material}.
    __device__ __forceinline__ bool hit(int slot, const Ray& r, float* t) const {
        const glb_float4* g = (const glb_float4*)prims + 3 * slot;
        if (has_lds_prims) {
            const f32x4 q = lds_prims[slot];
            if (q.w >= 0.0f) return hit_sphere(Vec3{q.x, q.y, q.z}, q.w, r, t);
            const f32x4 b = g[1], c = g[2];
            return hit_triangle(Vec3{q.x, q.y, q.z}, Vec3{b.x, b.y, b.z}, Vec3{c.x, c.y, c.z}, r, t);
        }
        const f32x4 a = g[0], b = g[1];
        if (__float_as_int(a.w) < n_spheres) return hit_sphere(Vec3{a.x, a.y, a.z}, b.w, r, t);
        const f32x4 c = g[2];
        return hit_triangle(Vec3{a.x, a.y, a.z}, Vec3{b.x, b.y, b.z}, Vec3{c.x, c.y, c.z}, r, t);
    }
};

// Wave-scheduled traversal. The per-ray sequence is exactly traversal_step's (same nodes, same
// primitive tests, same order), split into two phases so that a wave runs its expensive leaf
// tests together: a visit (trav_visit, trav_visit_signed) moves a lane through one node and, when
// the node is a hit leaf, parks the leaf in `leaf` as the node stores it, child = -(code + 1)
// with code = first << 2 | count - 1 (so a parked leaf is negative, and kNoLeaf is none), with
// the continuation already in tr.node; trav_leaf tests a parked leaf's primitives. A lane is live
// while tr.node >= 0 or a leaf is parked.
constexpr int kNoLeaf = 0;
__device__ __forceinline__ bool leaf_parked(int leaf) { return leaf < 0; }

template <class Src>
__device__ __forceinline__ void trav_visit(int n_nodes, const Src& src, Traversal& tr, int& leaf) {
    const Node node = src.node(tr.node);
    int n;
    if (hit_box(node.lo, node.hi, tr.r.o, tr.inv_d, tr.r.tmin, tr.r.tmax)) {
        if (node.child < 0) {
            leaf = node.child;
            n = node.skip;
        } else {
            n = node.child;
        }
    } else {
        n = node.skip;
    }
    tr.node = (n >= 0 && n < n_nodes) ? n : -1;
}

template <class Src>
__device__ __forceinline__ void trav_leaf(const Src& src, Traversal& tr, int& leaf, bool anyhit) {
    const int code = ~leaf;   // -(child + 1)
    const int first = code >> 2, cnt = (code & 3) + 1;
    leaf = kNoLeaf;
    for (int i = 0; i < cnt; ++i) {
        float t;
        if (src.hit(first + i, tr.r, &t)) {
            tr.r.tmax = t;
            tr.prim = first + i;
            if (anyhit) {
                tr.node = -1;
                return;
            }
        }
    }
}

// trav_visit for the rays of slab_signed_ok (hit_box_signed). Device trees are checked when they are
// laid out (every skip is -1 or a node, every child a node or a leaf code), so the next node needs
// no range check: -1 ends the traversal. ALL_LDS: every node is in LDS.
template <bool ALL_LDS, class Src>
__device__ __forceinline__ void trav_visit_signed(const Src& src, Traversal& tr, int& leaf, bool neg_x, bool neg_y,
                                                  bool neg_z) {
    Node node;
    if constexpr (ALL_LDS) node = src.node_lds(tr.node);
    else node = src.node(tr.node);
    const bool hit = hit_box_signed(node.lo, node.hi, tr.r.o, tr.inv_d, tr.r.tmin, tr.r.tmax, neg_x, neg_y, neg_z);
    const bool inner = node.child >= 0;   // one compare; its negation is a mask operation
    if (hit && !inner) leaf = node.child;
    tr.node = (hit && inner) ? node.child : node.skip;
}

// Node visits a moving lane makes per round of the while-while loop. Each lane still visits its
// nodes in the same order and parks the same leaves, so the results are those of one visit per
// round; what changes is how often the wave stops to count its parked and moving lanes (two
// ballots and a branch per round). Three visits took sphere_grid 1280x720x64 from 13.19 to
// 12.33 ms per render (extend 7.59 -> 7.02 ms, shadow 1.79 -> 1.48 ms); two gave 12.42, four 12.34.
// The first visit stays outside the loop: folded into it (`k == 0 ||`), the single-job shadow
// kernel took 65 VGPRs instead of 63, which leaves room for only one 1024-thread workgroup per CU
// (4 waves per SIMD instead of 8) and cost 0.35 ms per render.
constexpr int kVisitsPerRound = 3;

// MODE 0: trav_visit (any ray); 1: trav_visit_signed; 2: trav_visit_signed with every node in LDS.
template <int MODE, class Src>
__device__ __forceinline__ void wave_traverse_loop(int n_nodes, const Src& src, Traversal& tr, int& leaf, bool anyhit,
                                                   int limit, int leaf_batch, bool neg_x, bool neg_y, bool neg_z) {
    for (;;) {
        const bool parked = leaf_parked(leaf);
        const bool moving = tr.node >= 0 && !parked;
        const int np = __popcll(__ballot(parked));
        const int nm = __popcll(__ballot(moving));
        if (np + nm < limit || np + nm == 0) break;
        if (np > 0 && (nm == 0 || np >= leaf_batch)) {
            if (parked) trav_leaf(src, tr, leaf, anyhit);
        } else if (moving) {
            // Up to kVisitsPerRound visits before the wave counts its lanes again: a lane that
            // parks a leaf or finishes stops; the others go on.
            if constexpr (MODE == 0) trav_visit(n_nodes, src, tr, leaf);
            else trav_visit_signed<MODE == 2>(src, tr, leaf, neg_x, neg_y, neg_z);
#pragma unroll
            for (int k = 1; k < kVisitsPerRound; ++k) {
                if (tr.node >= 0 && !leaf_parked(leaf)) {
                    if constexpr (MODE == 0) trav_visit(n_nodes, src, tr, leaf);
                    else trav_visit_signed<MODE == 2>(src, tr, leaf, neg_x, neg_y, neg_z);
                }
            }
        }
    }
}

// Runs the wave's live lanes until fewer than `limit` of them are live (limit 1: until all are
// done). Each round either visits nodes in every lane that has no parked leaf, or, once at
// least `leaf_batch` lanes (or every live lane) have parked a leaf, tests those leaves together.
// The rays do not change inside, so the visit form is chosen once per call for the whole wave:
// the signed slab test unless a live ray fails slab_signed_ok.
template <class Src>
__device__ __forceinline__ void wave_traverse(int n_nodes, const Src& src, Traversal& tr, int& leaf, bool anyhit,
                                              int limit, int leaf_batch) {
    limit = __builtin_amdgcn_readfirstlane(limit);
    leaf_batch = __builtin_amdgcn_readfirstlane(leaf_batch);
    const bool live = tr.node >= 0 || leaf_parked(leaf);
    const bool finite = slab_signed_ok(tr.r.o, tr.inv_d, tr.r.tmin, tr.r.tmax);
    const bool neg_x = tr.inv_d.x < 0.0f, neg_y = tr.inv_d.y < 0.0f, neg_z = tr.inv_d.z < 0.0f;
    if (__ballot(live && !finite) != 0)
        wave_traverse_loop<0>(n_nodes, src, tr, leaf, anyhit, limit, leaf_batch, neg_x, neg_y, neg_z);
    else if (__ballot(live && !src.all_lds(n_nodes)) == 0)
        wave_traverse_loop<2>(n_nodes, src, tr, leaf, anyhit, limit, leaf_batch, neg_x, neg_y, neg_z);
    else
        wave_traverse_loop<1>(n_nodes, src, tr, leaf, anyhit, limit, leaf_batch, neg_x, neg_y, neg_z);
}

// Rays of a traversal kernel are dealt to the waves in slices of `size` consecutive items,
// round-robin: wave w takes slices w, w + W, w + 2W, ... of the W waves of the grid. Each slice
// is spatially coherent (consecutive items are neighbouring pixels or their survivors) while every
// wave samples the whole image, so the grid stays balanced however the cost of a ray varies across
// the image, without atomics. Large launches (the persistent engine) use 128-ray slices; launches
// of a few hundred thousand rays (the plain wavefront's bounces) use 64, one ray per lane, so that
// they still occupy every SIMD. Wave-uniform state.
constexpr uint32_t kFeedSlice = 128;

struct WaveFeed {
    uint32_t cursor = 0, end = 0;   // rest of the current slice [cursor, end)
    uint32_t slice = 0, stride = 0, size = kFeedSlice;
    bool dry = false;               // no slice left
};

__device__ __forceinline__ void feed_first(WaveFeed& f, uint32_t n, uint32_t size = kFeedSlice) {
    const uint32_t wpb = blockDim.x >> 6;
    f.size = size;
    f.stride = gridDim.x * wpb;
    f.slice = blockIdx.x * wpb + (threadIdx.x >> 6);
    const unsigned long long c = (unsigned long long)f.slice * size;
    f.cursor = c < n ? (uint32_t)c : n;
    f.end = min(f.cursor + size, n);
    f.dry = f.cursor >= n;
}

__device__ __forceinline__ void feed_next(WaveFeed& f, uint32_t n) {
    f.slice += f.stride;
    const unsigned long long c = (unsigned long long)f.slice * f.size;
    if (c >= n) {
        f.dry = true;
        f.cursor = f.end = n;
    } else {
        f.cursor = (uint32_t)c;
        f.end = min(f.cursor + f.size, n);
    }
}

// True when no wave of this block has a slice (the block returns before staging LDS).
__device__ __forceinline__ bool block_idle(uint32_t n, uint32_t size = kFeedSlice) {
    return (unsigned long long)blockIdx.x * (blockDim.x >> 6) * size >= n;
}

// Wave64 queue push to K queues at once: the lane's value goes to queue q (q < 0: none). The first
// active lane issues the K atomics back to back (one round trip for all queues instead of one per
// queue); slots at or past `capacity` are dropped and counted in *overflow. Called by every active
// lane with the same control flow. Returns the slot or ~0u.
template <int K>
__device__ __forceinline__ uint32_t wave_push_multi(uint32_t* const* items, uint32_t* counters, uint32_t capacity,
                                                    uint32_t value, int q, uint32_t* overflow) {
    uint64_t m[K];
    uint32_t base[K];
#pragma unroll
    for (int k = 0; k < K; ++k) {
        m[k] = __ballot(q == k);
        base[k] = 0u;
    }
    const int leader = __ffsll((unsigned long long)__ballot(true)) - 1;
    if ((int)lane_id() == leader) {
#pragma unroll
        for (int k = 0; k < K; ++k) base[k] = m[k] ? atomicAdd(&counters[k], (uint32_t)__popcll(m[k])) : 0u;
        uint32_t dropped = 0;
#pragma unroll
        for (int k = 0; k < K; ++k) {
            const uint32_t n = (uint32_t)__popcll(m[k]);
            const uint32_t keep = base[k] >= capacity ? 0u : (capacity - base[k] < n ? capacity - base[k] : n);
            dropped += n - keep;
        }
        if (dropped) atomicAdd(overflow, dropped);
    }
    uint32_t slot = 0xFFFFFFFFu;
#pragma unroll
    for (int k = 0; k < K; ++k) {
        const uint32_t b = (uint32_t)__shfl((int)base[k], leader);
        if (q == k) {
            const uint32_t sl = b + mbcnt64(m[k]);
            if (sl < capacity) {
                items[k][sl] = value;
                slot = sl;
            }
        }
    }
    return slot;
}

// Workgroup queue push to K queues: as wave_push_multi, but the waves' ballot counts are combined
// in LDS and one thread issues one atomic per queue for the whole workgroup (16x fewer atomics on
// the shared counters at 1024 threads). Every thread of the workgroup calls it (it synchronises);
// `lds_cnt` holds K x 16 words, `lds_base` K words. A lane's slot is its queue's workgroup base +
// the counts of the waves before it + its rank in its wave (mbcnt). Returns the slot or ~0u.
template <int K>
__device__ __forceinline__ uint32_t block_push_multi(uint32_t* const* items, uint32_t* counters, uint32_t capacity,
                                                     uint32_t value, int q, uint32_t* overflow,
                                                     uint32_t (*lds_cnt)[16], uint32_t* lds_base) {
    const uint32_t wave = threadIdx.x >> 6, nwaves = blockDim.x >> 6;
    uint64_t mine = 0;
#pragma unroll
    for (int k = 0; k < K; ++k) {
        const uint64_t m = __ballot(q == k);
        if (q == k) mine = m;
        if (lane_id() == 0) lds_cnt[k][wave] = (uint32_t)__popcll(m);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t dropped = 0;
#pragma unroll
        for (int k = 0; k < K; ++k) {
            uint32_t t = 0;
            for (uint32_t w = 0; w < nwaves; ++w) {
                const uint32_t c = lds_cnt[k][w];
                lds_cnt[k][w] = t;
                t += c;
            }
            const uint32_t base = t ? atomicAdd(&counters[k], t) : 0u;
            lds_base[k] = base;
            const uint32_t keep = base >= capacity ? 0u : (capacity - base < t ? capacity - base : t);
            dropped += t - keep;
        }
        if (dropped) atomicAdd(overflow, dropped);
    }
    __syncthreads();
    uint32_t slot = 0xFFFFFFFFu;
    for (int k = 0; k < K; ++k)
        if (q == k) {
            const uint32_t sl = lds_base[k] + lds_cnt[k][wave] + mbcnt64(mine);
            if (sl < capacity) {
                items[k][sl] = value;
                slot = sl;
            }
        }
    return slot;
}

// Workgroup slot reservation on one counter (block_push_multi without the item store): lanes with
// pred get consecutive slots; a slot at or past `capacity` is dropped, counted in *overflow and
// returned as ~0u. Every thread of the workgroup calls it.
__device__ __forceinline__ uint32_t block_reserve(uint32_t* counter, uint32_t capacity, bool pred, uint32_t* overflow,
                                                  uint32_t (*lds_cnt)[16], uint32_t* lds_base) {
    const uint32_t wave = threadIdx.x >> 6, nwaves = blockDim.x >> 6;
    const uint64_t m = __ballot(pred);
    if (lane_id() == 0) lds_cnt[0][wave] = (uint32_t)__popcll(m);
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t t = 0;
        for (uint32_t w = 0; w < nwaves; ++w) {
            const uint32_t c = lds_cnt[0][w];
            lds_cnt[0][w] = t;
            t += c;
        }
        const uint32_t base = t ? atomicAdd(counter, t) : 0u;
        lds_base[0] = base;
        const uint32_t keep = base >= capacity ? 0u : (capacity - base < t ? capacity - base : t);
        if (t > keep) atomicAdd(overflow, t - keep);
    }
    __syncthreads();
    if (!pred) return 0xFFFFFFFFu;
    const uint32_t sl = lds_base[0] + lds_cnt[0][wave] + mbcnt64(m);
    return sl < capacity ? sl : 0xFFFFFFFFu;
}

// Traces items [0, n) with the slice feed and the wave schedule of wave_traverse (one source for
// every lane). load(item, &tag) gives an item's ray and a per-lane tag; done(item, tag, tr) runs
// once per item when its traversal has ended (divergent). Called by every lane of the wave.
template <class Src, class Load, class Done>
__device__ __forceinline__ void trace_items(uint32_t n, uint32_t slice, int n_nodes, const Src& src, bool anyhit,
                                            int refill, int leaf_batch, Load load, Done done) {
    constexpr uint32_t kIdleItem = 0xFFFFFFFFu;
    WaveFeed f;
    feed_first(f, n, slice);
    uint32_t item = kIdleItem, tag = 0;
    Traversal tr;
    tr.node = -1;
    tr.prim = -1;
    int leaf = kNoLeaf;
    for (;;) {
        const uint64_t idle = __ballot(item == kIdleItem);
        if (idle != 0 && f.cursor >= f.end && !f.dry) feed_next(f, n);
        if (idle != 0 && f.cursor < f.end) {
            const uint32_t cand = f.cursor + mbcnt64(idle);
            if (item == kIdleItem && cand < f.end) {
                item = cand;
                tr = traversal_begin(load(item, &tag));
                if (n_nodes <= 0) tr.node = -1;
            }
            f.cursor = min(f.cursor + (uint32_t)__popcll(idle), f.end);
        }
        if (__ballot(item != kIdleItem) == 0) {
            if (f.dry) break;
            continue;
        }
        wave_traverse(n_nodes, src, tr, leaf, anyhit, (f.dry && f.cursor >= f.end) ? 1 : refill, leaf_batch);
        if (item != kIdleItem && tr.node < 0 && !leaf_parked(leaf)) {
            done(item, tag, tr);
            item = kIdleItem;
        }
    }
}

// `lds` and `lds_prims` must point into LDS (__shared__ memory).
__device__ __forceinline__ LdsNodes lds_nodes_of(const SceneView& s, const Node* lds, int n_lds,
                                                 const float4* lds_prims = nullptr) {
    return LdsNodes{(lds_float4*)lds, s.nodes, n_lds, s.prims, s.n_spheres,
                    (lds_float4*)(lds_prims ? lds_prims : reinterpret_cast<const float4*>(lds)),
                    lds_prims ? 1 : 0};
}

// Compact LDS record of leaf slot `slot` (see LdsNodes).
__device__ __forceinline__ float4 lds_prim_record(const SceneView& s, int slot) {
    const Prim& p = s.prims[slot];
    return p.id < s.n_spheres ? make_float4(p.p0.x, p.p0.y, p.p0.z, p.radius)
                              : make_float4(p.p0.x, p.p0.y, p.p0.z, -1.0f);
}

// Stages the first n nodes into LDS; the whole block must call it.
__device__ __forceinline__ void stage_nodes(Node* lds, const Node* nodes, int n) {
    for (int i = threadIdx.x; i < n; i += blockDim.x) lds[i] = nodes[i];
    __syncthreads();
}

// Workgroups are dispatched round-robin over the 8 XCDs; map them so that each XCD (and so each
// L2) works on one contiguous range of logical block ids.
__device__ __forceinline__ int xcd_remap(int bid, int nblocks) {
    const int x = bid & 7, k = bid >> 3;
    const int per = nblocks >> 3, rem = nblocks & 7;
    return x * per + (x < rem ? x : rem) + k;
}

}  // namespace wf::gpu
