// Estimator test hooks (include/wfamd.h): the same shared functions evaluated on the host and on
// gfx950, so tests can show bit-for-bit agreement of the two compilations.
#include <algorithm>
#include <cstring>
#include <stdexcept>
#include <vector>

#include "../../include/wfamd.h"
#include "wf_gpu.h"
#include "wf_handles.h"
#include "wf_persistent.h"

namespace wf::gpu {
namespace {

__global__ void sincos_kernel(const float* x, int n, float* s, float* c) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) sincos_f32(x[i], &s[i], &c[i]);
}

__device__ __host__ inline void grid_angle(uint32_t k, float* s, float* c) {
    const float u = (float)k * (1.0f / 16777216.0f);
    const float phi = 2.0f * kPi * u;
    sincos_f32(phi, s, c);
}

// Per-block FNV-1a over the (s, c) bits of a contiguous range; blocks are combined on the host.
__global__ void sincos_grid_kernel(uint32_t k0, uint32_t k1, uint32_t per_block, uint64_t* out) {
    __shared__ uint32_t bits[2 * 256];
    const uint32_t base = k0 + blockIdx.x * per_block;
    uint64_t h = 14695981039346656037ULL;
    for (uint32_t off = 0; off < per_block; off += 256) {
        const uint32_t k = base + off + threadIdx.x;
        float s = 0.0f, c = 0.0f;
        const bool ok = k < k1 && off + threadIdx.x < per_block;
        if (ok) grid_angle(k, &s, &c);
        bits[2 * threadIdx.x] = ok ? (uint32_t)__float_as_int(s) : 0u;
        bits[2 * threadIdx.x + 1] = ok ? (uint32_t)__float_as_int(c) : 0u;
        __syncthreads();
        if (threadIdx.x == 0) {
            for (uint32_t j = 0; j < 256; ++j) {
                const uint32_t kk = base + off + j;
                if (kk >= k1 || off + j >= per_block) break;
                for (int w = 0; w < 2; ++w) {
                    const uint32_t v = bits[2 * j + w];
                    for (int b = 0; b < 4; ++b) {
                        h ^= (v >> (8 * b)) & 0xffu;
                        h *= 1099511628211ULL;
                    }
                }
            }
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) out[blockIdx.x] = h;
}

// Ray of an 8-float record (o xyz, tmin, d xyz, tmax).
__device__ __host__ inline Ray ray_of(const float* q) {
    Ray r;
    r.o = Vec3{q[0], q[1], q[2]};
    r.tmin = q[3];
    r.d = Vec3{q[4], q[5], q[6]};
    r.tmax = q[7];
    return r;
}

// One slab case of 14 floats (lo xyz, hi xyz, o xyz, d xyz, tmin, tmax): out[0] = hit_box,
// out[1] = hit_box_signed where slab_signed_ok holds, else -1.
__device__ __host__ inline void slab_case(const float* q, int* out) {
    const Vec3 lo{q[0], q[1], q[2]}, hi{q[3], q[4], q[5]}, o{q[6], q[7], q[8]}, d{q[9], q[10], q[11]};
    const float tmin = q[12], tmax = q[13];
    const Vec3 inv{1.0f / d.x, 1.0f / d.y, 1.0f / d.z};
    out[0] = hit_box(lo, hi, o, inv, tmin, tmax) ? 1 : 0;
    out[1] = slab_signed_ok(o, inv, tmin, tmax)
                 ? (hit_box_signed(lo, hi, o, inv, tmin, tmax, inv.x < 0.0f, inv.y < 0.0f, inv.z < 0.0f) ? 1 : 0)
                 : -1;
}

__global__ void divu_kernel(const uint32_t* n, int count, DivU32 v, uint32_t* q) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) q[i] = divu32(n[i], v);
}

__global__ void slab_kernel(const float* q, int n, int* out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) slab_case(q + 14 * i, out + 2 * i);
}

// The engines' wave-scheduled traversal (slice feed, wave_traverse with the signed, mixed and
// exact visit forms) over the first n_lds nodes in LDS, and n_lds_prims primitive records.
constexpr int kWaveTestThreads = 512;
__global__ __launch_bounds__(kWaveTestThreads) void wave_intersect_kernel(SceneView s, const float* rays, int n,
                                                                          int anyhit, float* t, int* prim, int n_lds,
                                                                          int n_lds_prims) {
    extern __shared__ Node lds[];
    if (block_idle((uint32_t)n, 64)) return;
    for (int k = threadIdx.x; k < n_lds; k += blockDim.x) lds[k] = s.nodes[k];
    float4* lprim = reinterpret_cast<float4*>(lds + n_lds);
    for (int k = threadIdx.x; k < n_lds_prims; k += blockDim.x) lprim[k] = lds_prim_record(s, k);
    __syncthreads();
    const LdsNodes src = lds_nodes_of(s, lds, n_lds, n_lds_prims > 0 ? lprim : nullptr);
    trace_items(
        (uint32_t)n, 64, s.n_nodes, src, anyhit != 0, 32, 16,
        [&](uint32_t i, uint32_t* tag) {
            *tag = i;
            return ray_of(rays + 8 * i);
        },
        [&](uint32_t i, uint32_t, const Traversal& tr) {
            t[i] = tr.r.tmax;
            prim[i] = tr.prim >= 0 ? s.prims[tr.prim].id : -1;
        });
}

__global__ void intersect_kernel(SceneView s, const float* rays, int n, int anyhit, float* t, int* prim) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    Hit h;
    intersect(s, ray_of(rays + 8 * i), &h, anyhit != 0);
    t[i] = h.t;
    prim[i] = h.prim >= 0 ? s.prims[h.prim].id : -1;
}

__global__ void trace_kernel(SceneView s, const int* pxys, int n, uint32_t seed, int max_depth, float* L,
                             uint32_t* rays) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    RayCount rc{0, 0, 0};
    const Vec3 v = trace_path(s, pxys[3 * i], pxys[3 * i + 1], (uint32_t)pxys[3 * i + 2], seed, max_depth, &rc);
    L[3 * i] = v.x;
    L[3 * i + 1] = v.y;
    L[3 * i + 2] = v.z;
    rays[2 * i] = rc.extension;
    rays[2 * i + 1] = rc.shadow;
}

// One classify + shade step. A case is kShadeCaseWords words (ints as their bits): ray o xyz,
// d xyz, hit t, hit primitive (device order, -1 for a miss), throughput xyz, L xyz, prev_pdf, depth,
// specular, key k0, k1, max_depth. The output is kShadeOutWords words: alive, has_shadow, classify's
// type, the path's ray (o xyz, d xyz, tmin, tmax), throughput xyz, L xyz, prev_pdf, depth, specular,
// the shadow ray (o xyz, d xyz, tmin, tmax), its contribution xyz, then the hit primitive and t.
// MERGED picks shade_surface over the three per-material functions.
constexpr int kShadeCaseWords = 20, kShadeOutWords = 34;

template <bool MERGED>
__device__ __host__ inline void shade_case(const SceneView& s, const uint32_t* q, uint32_t* o) {
    auto f = [&](int i) { return __builtin_bit_cast(float, q[i]); };
    PathState ps;
    ps.ray.o = Vec3{f(0), f(1), f(2)};
    ps.ray.d = Vec3{f(3), f(4), f(5)};
    ps.ray.tmin = kRayEps;
    ps.ray.tmax = kInf;
    const Hit h{f(6), (int)q[7]};
    ps.throughput = Vec3{f(8), f(9), f(10)};
    ps.L = Vec3{f(11), f(12), f(13)};
    ps.prev_pdf = f(14);
    ps.depth = (int)q[15];
    ps.specular = (int)q[16];
    ps.key = PathKey{q[17], q[18]};
    Surface sf;
    Ray sh{};
    Vec3 c{0.0f, 0.0f, 0.0f};
    int has = 0;
    bool alive = false;
    const int mat = classify(s, ps, h, (int)q[19], &sf);
    if (MERGED) {
        if (mat >= kLambert && mat <= kDielectric) alive = shade_surface(s, ps, sf, mat, &sh, &c, &has);
    } else if (mat == kLambert) {
        alive = shade_lambert(s, ps, sf, &sh, &c, &has);
    } else if (mat == kMetal) {
        alive = shade_metal(s, ps, sf);
    } else if (mat == kDielectric) {
        alive = shade_dielectric(s, ps, sf);
    }
    const float w[31] = {ps.ray.o.x,  ps.ray.o.y,  ps.ray.o.z,   ps.ray.d.x,   ps.ray.d.y,  ps.ray.d.z,  ps.ray.tmin,
                         ps.ray.tmax, ps.throughput.x, ps.throughput.y, ps.throughput.z, ps.L.x, ps.L.y, ps.L.z,
                         ps.prev_pdf, 0.0f,        0.0f,         sh.o.x,       sh.o.y,      sh.o.z,      sh.d.x,
                         sh.d.y,      sh.d.z,      sh.tmin,      sh.tmax,      c.x,         c.y,         c.z,
                         0.0f,        h.t,         0.0f};
    o[0] = alive ? 1u : 0u;
    o[1] = (uint32_t)has;
    o[2] = (uint32_t)mat;
    for (int i = 0; i < 31; ++i) o[3 + i] = __builtin_bit_cast(uint32_t, w[i]);
    o[18] = (uint32_t)ps.depth;
    o[19] = (uint32_t)ps.specular;
    o[31] = (uint32_t)h.prim;
    o[33] = 0u;
}

template <bool MERGED>
__global__ void shade_case_kernel(SceneView s, const uint32_t* q, int n, uint32_t* out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) shade_case<MERGED>(s, q + kShadeCaseWords * i, out + kShadeOutWords * i);
}

// Plain 16-byte-per-lane copy: the calibration kernel for hardware traffic counters (a kernel
// whose device-memory bytes are known exactly: n16 * 16 read, n16 * 16 written).
__global__ void copy_probe_kernel(const float4* __restrict__ src, float4* __restrict__ dst, size_t n16) {
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n16; i += stride) dst[i] = src[i];
}

template <class T>
struct DevBuf {
    T* p = nullptr;
    explicit DevBuf(size_t n) { check(hipMalloc(&p, (n ? n : 1) * sizeof(T)), "hipMalloc test buffer"); }
    ~DevBuf() {
        if (p) (void)hipFree(p);
    }
    void put(const T* h, size_t n) { check(hipMemcpy(p, h, n * sizeof(T), hipMemcpyHostToDevice), "hipMemcpy"); }
    void get(T* h, size_t n) { check(hipMemcpy(h, p, n * sizeof(T), hipMemcpyDeviceToHost), "hipMemcpy"); }
};

uint64_t fnv_words(uint64_t h, uint32_t v) {
    for (int b = 0; b < 4; ++b) {
        h ^= (v >> (8 * b)) & 0xffu;
        h *= 1099511628211ULL;
    }
    return h;
}

}  // namespace
}  // namespace wf::gpu

namespace {
template <class F>
int hook(F&& f) {
    try {
        f();
        return WF_OK;
    } catch (const std::invalid_argument& e) {
        return wf::set_last_error(WF_ERR_INVALID, e.what());
    } catch (const std::exception& e) {
        return wf::set_last_error(WF_ERR_RUNTIME, e.what());
    }
}
}  // namespace

using namespace wf;
using namespace wf::gpu;

extern "C" {

int wf_test_sincos(const float* x, int n, float* s, float* c, int device) {
    return hook([&] {
        if (n < 0 || (n > 0 && (!x || !s || !c))) throw std::invalid_argument("bad argument");
        if (!device) {
            for (int i = 0; i < n; ++i) sincos_f32(x[i], &s[i], &c[i]);
            return;
        }
        DevBuf<float> dx(n), ds(n), dc(n);
        dx.put(x, n);
        if (n) sincos_kernel<<<(n + 255) / 256, 256>>>(dx.p, n, ds.p, dc.p);
        check(hipGetLastError(), "sincos kernel");
        check(hipDeviceSynchronize(), "sincos sync");
        ds.get(s, n);
        dc.get(c, n);
    });
}

int wf_test_sincos_grid_hash(uint32_t k0, uint32_t k1, int device, uint64_t* hash) {
    return hook([&] {
        if (!hash || k1 < k0 || k1 > (1u << 24)) throw std::invalid_argument("bad range");
        const uint32_t per_block = 4096;
        const uint32_t nb = (k1 - k0 + per_block - 1) / per_block;
        std::vector<uint64_t> parts(nb);
        if (device) {
            DevBuf<uint64_t> d(nb);
            if (nb) sincos_grid_kernel<<<nb, 256>>>(k0, k1, per_block, d.p);
            check(hipGetLastError(), "sincos grid kernel");
            check(hipDeviceSynchronize(), "sincos grid sync");
            d.get(parts.data(), nb);
        } else {
            for (uint32_t b = 0; b < nb; ++b) {
                uint64_t h = 14695981039346656037ULL;
                for (uint32_t k = k0 + b * per_block; k < k1 && k < k0 + (b + 1) * per_block; ++k) {
                    float s, c;
                    grid_angle(k, &s, &c);
                    uint32_t sb, cb;
                    std::memcpy(&sb, &s, 4);
                    std::memcpy(&cb, &c, 4);
                    h = fnv_words(fnv_words(h, sb), cb);
                }
                parts[b] = h;
            }
        }
        uint64_t h = 14695981039346656037ULL;
        for (uint64_t p : parts) {
            h = fnv_words(h, (uint32_t)p);
            h = fnv_words(h, (uint32_t)(p >> 32));
        }
        *hash = h;
    });
}

int wf_test_lds_plan(const int* n_nodes, const int* n_prims, int n, int64_t budget_bytes, int* node_cnt,
                     int* prim_cnt) {
    return hook([&] {
        if (n < 0 || (n > 0 && (!n_nodes || !n_prims || !node_cnt || !prim_cnt))) throw std::invalid_argument("bad argument");
        const LdsPlan p = plan_lds(n_nodes, n_prims, n, budget_bytes);
        for (int k = 0; k < n; ++k) {
            node_cnt[k] = p.node_cnt[k];
            prim_cnt[k] = p.prim_cnt[k];
        }
    });
}

int wf_test_divu32(const uint32_t* n, int count, uint32_t d, uint32_t* q, int device) {
    return hook([&] {
        if (d == 0 || count < 0 || (count > 0 && (!n || !q))) throw std::invalid_argument("bad argument");
        const DivU32 v = make_divu32(d);
        if (!device) {
            for (int i = 0; i < count; ++i) q[i] = divu32(n[i], v);
            return;
        }
        DevBuf<uint32_t> dn(count), dq(count);
        dn.put(n, count);
        if (count) divu_kernel<<<(count + 255) / 256, 256>>>(dn.p, count, v, dq.p);
        check(hipGetLastError(), "divu kernel");
        check(hipDeviceSynchronize(), "divu sync");
        dq.get(q, count);
    });
}

int wf_test_slab(const float* cases, int n, int* out, int device) {
    return hook([&] {
        if (n < 0 || (n > 0 && (!cases || !out))) throw std::invalid_argument("bad argument");
        if (!device) {
            for (int i = 0; i < n; ++i) slab_case(cases + 14 * i, out + 2 * i);
            return;
        }
        DevBuf<float> dq((size_t)14 * n);
        DevBuf<int> dout((size_t)2 * n);
        dq.put(cases, (size_t)14 * n);
        if (n) slab_kernel<<<(n + 255) / 256, 256>>>(dq.p, n, dout.p);
        check(hipGetLastError(), "slab kernel");
        check(hipDeviceSynchronize(), "slab sync");
        dout.get(out, (size_t)2 * n);
    });
}

int wf_test_intersect(const wf_scene* scene, const float* rays, int n, int anyhit, float* t, int* prim, int device) {
    return hook([&] {
        if (!scene || n < 0 || (n > 0 && (!rays || !t || !prim))) throw std::invalid_argument("bad argument");
        if (device < 0 || device > 3) throw std::invalid_argument("device must be 0..3");
        const Scene& sc = scene->s;
        if (!device) {
            const SceneView v = host_view(sc);
            for (int i = 0; i < n; ++i) {
                Hit h;
                intersect(v, ray_of(rays + 8 * i), &h, anyhit != 0);
                t[i] = h.t;
                prim[i] = h.prim >= 0 ? sc.dprims[h.prim].id : -1;
            }
            return;
        }
        DeviceScene ds;
        ds.upload(sc);
        DevBuf<float> dr((size_t)8 * n), dt(n);
        DevBuf<int> dp(n);
        dr.put(rays, (size_t)8 * n);
        if (n && device == 1) intersect_kernel<<<(n + 255) / 256, 256>>>(ds.view, dr.p, n, anyhit, dt.p, dp.p);
        if (n && device >= 2) {
            const int nn = ds.view.n_nodes, np = ds.view.n_prims;
            const int n_lds = device == 2 ? std::min(nn, 1024) : std::min(nn, 8);
            const int n_lds_prims = (device == 2 && n_lds == nn && (size_t)np * 16 + (size_t)n_lds * 32 <= 64 * 1024) ? np : 0;
            const size_t lds = (size_t)n_lds * sizeof(Node) + (size_t)n_lds_prims * 16;
            const int per_block = 64 * (kWaveTestThreads / 64);
            const int blocks = std::min((n + per_block - 1) / per_block, 1024);
            wave_intersect_kernel<<<blocks, kWaveTestThreads, lds>>>(ds.view, dr.p, n, anyhit, dt.p, dp.p, n_lds,
                                                                     n_lds_prims);
        }
        check(hipGetLastError(), "intersect kernel");
        check(hipDeviceSynchronize(), "intersect sync");
        dt.get(t, n);
        dp.get(prim, n);
    });
}

int wf_test_trace_paths(const wf_scene* scene, const int* pxys, int n, uint32_t seed, int max_depth, float* L,
                        uint32_t* rays, int device) {
    return hook([&] {
        if (!scene || n < 0 || (n > 0 && (!pxys || !L || !rays))) throw std::invalid_argument("bad argument");
        const Scene& sc = scene->s;
        for (int i = 0; i < n; ++i)
            if (pxys[3 * i] < 0 || pxys[3 * i] >= sc.width || pxys[3 * i + 1] < 0 || pxys[3 * i + 1] >= sc.height)
                throw std::invalid_argument("pixel outside the image");
        if (!device) {
            const SceneView v = host_view(sc);
            for (int i = 0; i < n; ++i) {
                RayCount rc{0, 0, 0};
                Vec3 r = trace_path(v, pxys[3 * i], pxys[3 * i + 1], (uint32_t)pxys[3 * i + 2], seed, max_depth, &rc);
                L[3 * i] = r.x;
                L[3 * i + 1] = r.y;
                L[3 * i + 2] = r.z;
                rays[2 * i] = rc.extension;
                rays[2 * i + 1] = rc.shadow;
            }
            return;
        }
        DeviceScene ds;
        ds.upload(sc);
        DevBuf<int> dq((size_t)3 * n);
        DevBuf<float> dl((size_t)3 * n);
        DevBuf<uint32_t> dc((size_t)2 * n);
        dq.put(pxys, (size_t)3 * n);
        if (n) trace_kernel<<<(n + 255) / 256, 256>>>(ds.view, dq.p, n, seed, max_depth, dl.p, dc.p);
        check(hipGetLastError(), "trace kernel");
        check(hipDeviceSynchronize(), "trace sync");
        dl.get(L, (size_t)3 * n);
        dc.get(rays, (size_t)2 * n);
    });
}

int wf_test_shade(const wf_scene* scene, const uint32_t* cases, int n, int form, uint32_t* out) {
    return hook([&] {
        if (!scene || n < 0 || (n > 0 && (!cases || !out))) throw std::invalid_argument("bad argument");
        if (form < 0 || form > 3) throw std::invalid_argument("form must be 0..3");
        const Scene& sc = scene->s;
        const SceneView v = host_view(sc);
        std::vector<uint32_t> q(cases, cases + (size_t)kShadeCaseWords * n);
        for (int i = 0; i < n; ++i) {
            uint32_t* c = q.data() + (size_t)kShadeCaseWords * i;
            const int p = (int)c[7];
            if (p == -2) {   // closest hit of the case's ray, found on the host
                Ray r;
                r.o = Vec3{__builtin_bit_cast(float, c[0]), __builtin_bit_cast(float, c[1]), __builtin_bit_cast(float, c[2])};
                r.d = Vec3{__builtin_bit_cast(float, c[3]), __builtin_bit_cast(float, c[4]), __builtin_bit_cast(float, c[5])};
                r.tmin = kRayEps;
                r.tmax = kInf;
                Hit h;
                intersect(v, r, &h, false);
                c[6] = __builtin_bit_cast(uint32_t, h.t);
                c[7] = (uint32_t)h.prim;
            } else if (p < -1 || p >= v.n_prims) {
                throw std::invalid_argument("primitive index out of range");
            }
        }
        if (form < 2) {
            for (int i = 0; i < n; ++i) {
                const uint32_t* c = q.data() + (size_t)kShadeCaseWords * i;
                uint32_t* o = out + (size_t)kShadeOutWords * i;
                if (form == 1) shade_case<true>(v, c, o);
                else shade_case<false>(v, c, o);
            }
            return;
        }
        DeviceScene ds;
        ds.upload(sc);
        DevBuf<uint32_t> dq((size_t)kShadeCaseWords * n), dout((size_t)kShadeOutWords * n);
        dq.put(q.data(), (size_t)kShadeCaseWords * n);
        if (n && form == 3) shade_case_kernel<true><<<(n + 255) / 256, 256>>>(ds.view, dq.p, n, dout.p);
        if (n && form == 2) shade_case_kernel<false><<<(n + 255) / 256, 256>>>(ds.view, dq.p, n, dout.p);
        check(hipGetLastError(), "shade kernel");
        check(hipDeviceSynchronize(), "shade sync");
        dout.get(out, (size_t)kShadeOutWords * n);
    });
}

int wf_copy_probe(uint64_t bytes, int reps, double* ms_per_copy) {
    return hook([&] {
        if (bytes < 16 || bytes % 16 != 0 || reps < 1 || !ms_per_copy) throw std::invalid_argument("bad argument");
        const size_t n16 = bytes / 16;
        DevBuf<float4> src(n16), dst(n16);
        check(hipMemset(src.p, 0, bytes), "hipMemset");
        int dev = 0, cus = 0;
        check(hipGetDevice(&dev), "hipGetDevice");
        check(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev), "hipDeviceGetAttribute");
        const int grid = cus * 8;
        copy_probe_kernel<<<grid, 256>>>(src.p, dst.p, n16);   // warm-up
        hipEvent_t a, b;
        check(hipEventCreate(&a), "hipEventCreate");
        check(hipEventCreate(&b), "hipEventCreate");
        check(hipEventRecord(a, nullptr), "hipEventRecord");
        for (int k = 0; k < reps; ++k) copy_probe_kernel<<<grid, 256>>>(src.p, dst.p, n16);
        check(hipEventRecord(b, nullptr), "hipEventRecord");
        check(hipEventSynchronize(b), "hipEventSynchronize");
        check(hipGetLastError(), "copy probe");
        float ms = 0.0f;
        check(hipEventElapsedTime(&ms, a, b), "hipEventElapsedTime");
        *ms_per_copy = (double)ms / reps;
        (void)hipEventDestroy(a);
        (void)hipEventDestroy(b);
    });
}

}  // extern "C"
