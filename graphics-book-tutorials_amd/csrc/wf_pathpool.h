// Structure-of-arrays path state shared by the wavefront engines. Every per-path field is a 16-,
// 8- or 4-byte element so one lane issues one load per field and a wave64 reads contiguous
// 256..1024-byte runs when consecutive lanes handle consecutive slots.
//   ro  : ray origin xyz, prev_pdf          rd : ray direction xyz, depth | specular << 16
//   T   : throughput xyz                    L  : radiance xyz
//   key : PathKey                           hit: t, leaf slot
//   sp  : surface point xyz, material | inside << 31
//   sn  : facing normal xyz
// Extension rays always have tmin = kRayEps and tmax = kInf (include/wf/path_core.hpp:312-317),
// so those two words are not stored.
#pragma once
#include "wf_gpu.h"

namespace wf::gpu {

struct PathPool {
    float4* ro = nullptr;
    float4* rd = nullptr;
    float4* T = nullptr;
    float4* L = nullptr;
    uint2* key = nullptr;
    float2* hit = nullptr;
    float4* sp = nullptr;
    float4* sn = nullptr;
};

struct ShadowQueue {
    float4* o = nullptr;   // origin xyz, tmax
    float4* d = nullptr;   // direction xyz, owning path slot
    float4* c = nullptr;   // contribution xyz (added to L if unoccluded)
};

__device__ __forceinline__ PathState load_path(const PathPool& P, uint32_t i, bool with_L) {
    PathState ps;
    const float4 o = P.ro[i], d = P.rd[i], t = P.T[i];
    const uint2 k = P.key[i];
    ps.ray.o = Vec3{o.x, o.y, o.z};
    ps.ray.d = Vec3{d.x, d.y, d.z};
    ps.ray.tmin = kRayEps;
    ps.ray.tmax = kInf;
    ps.prev_pdf = o.w;
    const int bits = __float_as_int(d.w);
    ps.depth = bits & 0xffff;
    ps.specular = bits >> 16;
    ps.throughput = Vec3{t.x, t.y, t.z};
    if (with_L) {
        const float4 l = P.L[i];
        ps.L = Vec3{l.x, l.y, l.z};
    } else {
        ps.L = Vec3{0.0f, 0.0f, 0.0f};
    }
    ps.key = PathKey{k.x, k.y};
    return ps;
}

// Stores ray, flags, prev_pdf and throughput (not L, not the key).
__device__ __forceinline__ void store_scatter(const PathPool& P, uint32_t i, const PathState& ps) {
    P.ro[i] = make_float4(ps.ray.o.x, ps.ray.o.y, ps.ray.o.z, ps.prev_pdf);
    P.rd[i] = make_float4(ps.ray.d.x, ps.ray.d.y, ps.ray.d.z, __int_as_float(ps.depth | (ps.specular << 16)));
    P.T[i] = make_float4(ps.throughput.x, ps.throughput.y, ps.throughput.z, 0.0f);
}

__device__ __forceinline__ void store_new_path(const PathPool& P, uint32_t i, const PathState& ps) {
    store_scatter(P, i, ps);
    P.L[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    P.key[i] = make_uint2(ps.key.k0, ps.key.k1);
}

__device__ __forceinline__ Surface load_surface(const PathPool& P, uint32_t i) {
    Surface f;
    const float4 a = P.sp[i], b = P.sn[i];
    const uint32_t bits = (uint32_t)__float_as_int(a.w);
    f.p = Vec3{a.x, a.y, a.z};
    f.n = Vec3{b.x, b.y, b.z};
    f.n_out = f.n;   // not used after classify
    f.material = (int)(bits & 0x7fffffffu);
    f.inside = (int)(bits >> 31);
    return f;
}

__device__ __forceinline__ void store_surface(const PathPool& P, uint32_t i, const Surface& f) {
    const uint32_t bits = (uint32_t)f.material | ((uint32_t)f.inside << 31);
    P.sp[i] = make_float4(f.p.x, f.p.y, f.p.z, __int_as_float((int)bits));
    P.sn[i] = make_float4(f.n.x, f.n.y, f.n.z, 0.0f);
}

// Bytes per path of the pool arrays above (ro, rd, T, L, key, hit, sp, sn).
constexpr size_t kPathStateBytes = 16 + 16 + 16 + 16 + 8 + 8 + 16 + 16;
constexpr size_t kShadowEntryBytes = 48;

}  // namespace wf::gpu
