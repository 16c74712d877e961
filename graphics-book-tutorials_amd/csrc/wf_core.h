// Shared path-tracing estimator: vector math, counter-based RNG, scene view, intersection,
// stackless skip-link traversal and the per-bounce estimator stages. Every function here is
// compiled twice from the same source -- for the x86 host (the CPU arm) and for gfx950 (the
// megakernel and every wavefront stage) -- and the build pins IEEE binary32 semantics on both:
// no FP contraction (-ffp-contract=off), correctly rounded divide and sqrt, denormals kept.
// Transcendentals that the estimator needs (sin/cos of 2*pi*u) are evaluated by sincos_f32 below
// in binary64 with exact argument reduction and rounded once, so host and device agree bit for
// bit and both agree with a correctly rounded sinf/cosf on every argument the estimator produces.
//
// Semantics follow the reference estimator:
//   math      include/wf/math.hpp:18-67      rng        include/wf/rng.hpp:8-42
//   geometry  include/wf/bvh.hpp:25-116      estimator  include/wf/path_core.hpp:13-304
//   camera    include/wf/scene.hpp:129-141   hit check  src/gpu/engine.cuh:44-48
// The float operation order of each formula is kept exactly (for example a/s is a*(1/s),
// include/wf/math.hpp:29), because the arms are compared bitwise. Extensions beyond the
// reference (metal fuzz, sky-gradient emitter with next-event estimation, directional sun) are
// off for the reference scenes, which then execute exactly the reference's operations.
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#define WF_HD __host__ __device__ __forceinline__

namespace wf {

constexpr float kPi = 3.14159265358979323846f;
constexpr float kInvPi = 0.31830988618379067154f;
constexpr float kInv4Pi = 0.07957747154594767f;
constexpr float kInf = 3.402823466e+38f;
constexpr float kRayEps = 1e-4f;

// ---------------------------------------------------------------------------------------------
// L0: vectors (include/wf/math.hpp)
// ---------------------------------------------------------------------------------------------
struct Vec3 {
    float x, y, z;
};

WF_HD Vec3 make_vec3(float x, float y, float z) { return Vec3{x, y, z}; }
WF_HD Vec3 operator+(Vec3 a, Vec3 b) { return Vec3{a.x + b.x, a.y + b.y, a.z + b.z}; }
WF_HD Vec3 operator-(Vec3 a, Vec3 b) { return Vec3{a.x - b.x, a.y - b.y, a.z - b.z}; }
WF_HD Vec3 operator-(Vec3 a) { return Vec3{-a.x, -a.y, -a.z}; }
WF_HD Vec3 operator*(Vec3 a, Vec3 b) { return Vec3{a.x * b.x, a.y * b.y, a.z * b.z}; }
WF_HD Vec3 operator*(Vec3 a, float s) { return Vec3{a.x * s, a.y * s, a.z * s}; }
WF_HD Vec3 operator*(float s, Vec3 a) { return Vec3{a.x * s, a.y * s, a.z * s}; }
WF_HD Vec3 operator/(Vec3 a, float s) {
    float r = 1.0f / s;
    return Vec3{a.x * r, a.y * r, a.z * r};
}
WF_HD void add_to(Vec3& a, Vec3 b) { a.x += b.x; a.y += b.y; a.z += b.z; }
WF_HD float dot(Vec3 a, Vec3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
WF_HD Vec3 cross(Vec3 a, Vec3 b) {
    return Vec3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
WF_HD float length2(Vec3 a) { return dot(a, a); }
WF_HD float length(Vec3 a) { return sqrtf(dot(a, a)); }
WF_HD Vec3 normalize(Vec3 a) {
    float l = length(a);
    return l > 0.0f ? a / l : a;
}
WF_HD float min2(float a, float b) { return a < b ? a : b; }
WF_HD float max2(float a, float b) { return a > b ? a : b; }
WF_HD float clampf(float v, float lo, float hi) { return v < lo ? lo : (v > hi ? hi : v); }
WF_HD float max_component(Vec3 a) { return max2(a.x, max2(a.y, a.z)); }
WF_HD Vec3 reflect(Vec3 d, Vec3 n) { return d - n * (2.0f * dot(d, n)); }
WF_HD float luminance(Vec3 c) { return 0.2126f * c.x + 0.7152f * c.y + 0.0722f * c.z; }
// NaN and +-inf both fail v - v == 0.
WF_HD bool is_finite1(float v) { return v - v == 0.0f; }
WF_HD bool is_finite3(Vec3 a) { return is_finite1(a.x) && is_finite1(a.y) && is_finite1(a.z); }

// Orthonormal frame around unit n, branching on the sign of n.z (Frisvad's construction as
// used at include/wf/math.hpp:54-60).
WF_HD void make_basis(Vec3 n, Vec3* t, Vec3* b) {
    float sign = n.z >= 0.0f ? 1.0f : -1.0f;
    float a = -1.0f / (sign + n.z);
    float c = n.x * n.y * a;
    *t = Vec3{1.0f + sign * n.x * n.x * a, sign * c, -sign * n.x};
    *b = Vec3{c, sign + n.y * n.y * a, -n.y};
}

struct Ray {
    Vec3 o;
    float tmin;
    Vec3 d;
    float tmax;
};

// sin and cos of a binary32 angle (|x| < 2^20). Cody-Waite reduction by pi/2 split into three
// parts (the first two with 33 significant bits, so k*C1 and k*C2 are exact for small k), then
// Taylor polynomials of sin (to x^15) and cos (to x^16) on |r| <= pi/4 evaluated in binary64,
// then one rounding to binary32. The truncation error (< 1e-16 relative) is far below half a
// binary32 ulp, so the result equals the correctly rounded value unless the exact value lies
// within ~1e-16 of a binary32 rounding boundary. Every operation is an IEEE binary64 multiply or
// add, so the host and gfx950 results are identical.
WF_HD void sincos_f32(float xf, float* s_out, float* c_out) {
    const double x = (double)xf;
    const double kd = floor(x * 0.63661977236758134 + 0.5);
    const double r = ((x - kd * 1.5707963267341256) - kd * 6.077100506303966e-11) -
                     kd * 2.0222662487959506e-21;
    const double r2 = r * r;
    double ps = -7.647163731819816e-13;
    ps = ps * r2 + 1.6059043836821613e-10;
    ps = ps * r2 + -2.505210838544172e-08;
    ps = ps * r2 + 2.7557319223985893e-06;
    ps = ps * r2 + -0.0001984126984126984;
    ps = ps * r2 + 0.008333333333333333;
    ps = ps * r2 + -0.16666666666666666;
    const double sn = r + (r * r2) * ps;
    double pc = 4.779477332387385e-14;
    pc = pc * r2 + -1.1470745597729725e-11;
    pc = pc * r2 + 2.08767569878681e-09;
    pc = pc * r2 + -2.755731922398589e-07;
    pc = pc * r2 + 2.48015873015873e-05;
    pc = pc * r2 + -0.001388888888888889;
    pc = pc * r2 + 0.041666666666666664;
    pc = pc * r2 + -0.5;
    const double cs = 1.0 + r2 * pc;
    const int q = ((int)kd) & 3;
    double s, c;
    if (q == 0) { s = sn; c = cs; }
    else if (q == 1) { s = cs; c = -sn; }
    else if (q == 2) { s = -sn; c = -cs; }
    else { s = -cs; c = sn; }
    *s_out = (float)s;
    *c_out = (float)c;
}

// Unsigned 32-bit division by an invariant divisor d >= 1 as a multiply-high and shifts
// (Granlund and Montgomery, "Division by invariant integers using multiplication", 1994: the
// round-up multiplier with the add-back step): with l = ceil(log2 d) and
// m = floor(2^32 (2^l - d) / d) + 1, n / d = (t + ((n - t) >> 1)) >> (l - 1) with t = mulhi(n, m),
// for every 32-bit n; d = 1 is passed through.
struct DivU32 {
    uint32_t m, shift, d, pad;
};

inline DivU32 make_divu32(uint32_t d) {
    if (d <= 1) return DivU32{0u, 0u, 1u, 0u};
    uint32_t l = 0;
    while ((1ull << l) < d) ++l;
    const unsigned long long m = ((1ull << 32) * ((1ull << l) - d)) / d + 1;
    return DivU32{(uint32_t)m, l - 1, d, 0u};
}

WF_HD uint32_t divu32(uint32_t n, const DivU32& v) {
    const uint32_t t = (uint32_t)(((unsigned long long)n * v.m) >> 32);
    const uint32_t q = (t + ((n - t) >> 1)) >> v.shift;
    return v.d == 1u ? n : q;   // a select, so callers can load the divisor with their other data
}

// ---------------------------------------------------------------------------------------------
// L0: counter-based RNG (include/wf/rng.hpp). Every draw is a pure function of
// (seed, pixel, sample, depth, dim), so arms that schedule paths differently draw identical numbers.
// ---------------------------------------------------------------------------------------------
WF_HD uint32_t hash_u32(uint32_t x) {  // lowbias32
    x ^= x >> 16;
    x *= 0x7feb352dU;
    x ^= x >> 15;
    x *= 0x846ca68bU;
    x ^= x >> 16;
    return x;
}

struct PathKey {
    uint32_t k0, k1;
};

// k1 of a path's key from the seed's hash b = path_key_b(seed): a batched engine keeps b per job
// and rebuilds k1 from the path's pixel and sample instead of storing it.
WF_HD uint32_t path_key_b(uint32_t seed) { return hash_u32(seed + 0x85ebca6bU); }
WF_HD uint32_t path_key_k1(uint32_t b, uint32_t pixel, uint32_t sample) {
    return hash_u32(hash_u32(b ^ (pixel * 0x9e3779b1U)) + (sample * 0x85ebca77U + 0xc2b2ae3dU));
}

// The key from the seed's two hashes a = path_key_a(seed) and b = path_key_b(seed), which a batched
// engine computes once per job.
WF_HD uint32_t path_key_a(uint32_t seed) { return hash_u32(seed ^ 0x9e3779b9U); }
WF_HD PathKey path_key_ab(uint32_t a, uint32_t b, uint32_t pixel, uint32_t sample) {
    uint32_t k0 = hash_u32(hash_u32(a + pixel * 0xc2b2ae35U) ^ (sample * 0x27d4eb2fU + 0x165667b1U));
    return PathKey{k0, path_key_k1(b, pixel, sample)};
}

WF_HD PathKey make_path_key(uint32_t seed, uint32_t pixel, uint32_t sample) {
    return path_key_ab(path_key_a(seed), path_key_b(seed), pixel, sample);
}

// Dimension slots per depth (< 16). 0-8 are the reference's (include/wf/rng.hpp:30-35);
// 9-10 are the metal-fuzz direction, an extension.
enum RngDim : uint32_t {
    kDimPixelX = 0, kDimPixelY = 1,
    kDimLightPick = 2, kDimLightU = 3, kDimLightV = 4,
    kDimBsdfU = 5, kDimBsdfV = 6, kDimRoulette = 7, kDimDielectric = 8,
    kDimFuzzU = 9, kDimFuzzV = 10,
    kDimCount = 16
};

// Uniform float in [0,1) on the 2^-24 grid.
WF_HD float rng(PathKey key, uint32_t depth, uint32_t dim) {
    uint32_t ctr = depth * kDimCount + dim;
    uint32_t h = hash_u32(key.k0 ^ hash_u32(ctr * 0x9e3779b9U + key.k1));
    return (h >> 8) * (1.0f / 16777216.0f);
}

// ---------------------------------------------------------------------------------------------
// L1: scene view in the device layout (shared by the CPU arm and the kernels).
// ---------------------------------------------------------------------------------------------
enum MaterialType : int { kLambert = 0, kMetal = 1, kDielectric = 2, kEmissive = 3 };

// Reference Material (include/wf/scene.hpp:12-20) with the metal fuzz radius in its first pad word.
struct Material {
    int type;
    float ior;
    float fuzz;
    float pad0;
    Vec3 albedo;
    float pad1;
    Vec3 emission;
    float pad2;
};

// BVH node, 32 bytes. The logical tree, box values and skip links are the reference's preorder
// skip-link tree (include/wf/scene.hpp:49-56, src/bvh_build.cpp:191-226); only the storage order
// changes: the left child is explicit so the top levels can sit first (breadth-first) and be
// staged in LDS. child >= 0: internal node, index of the left child. child < 0: leaf with
// code = -(child + 1), first primitive slot = code >> 2, count = (code & 3) + 1.
struct alignas(16) Node {
    Vec3 lo;
    int skip;   // next node on a miss or after a leaf; -1 terminates
    Vec3 hi;
    int child;
};

// Primitive record in BVH leaf order (slot k holds the reference's prim_order[k]). 48 bytes.
// Sphere: p0 = centre, radius. Triangle: p0, p1, p2 are the original vertex positions.
struct alignas(16) Prim {
    Vec3 p0;
    int id;        // primitive id in the shared space (spheres first, then triangles)
    Vec3 p1;
    float radius;
    Vec3 p2;
    int material;
};

struct Light {   // emissive primitive, in the reference's light order (src/scene.cpp:68-91)
    int slot;    // leaf slot of the primitive
    float area;
    int id;
    int pad;
};

struct CameraRig {
    Vec3 origin;
    Vec3 lower_left;
    Vec3 horizontal;
    Vec3 vertical;
};

// Escaped-ray radiance and non-geometric lights. sky_mode 0: constant env (the reference's
// include/wf/path_core.hpp:153-156). sky_mode 1: gradient lerp(sky_lo, sky_hi, 0.5*(d.y+1)).
// sky_light: the sky also takes part in next-event estimation (uniform-sphere sampling, MIS).
// sun_light: directional delta light toward sun_dir with radiance sun_E (no MIS, never hit).
struct EnvParams {
    Vec3 env;
    int sky_mode;
    Vec3 sky_lo;
    int sky_light;
    Vec3 sky_hi;
    int sun_light;
    Vec3 sun_dir;
    float pad0;
    Vec3 sun_E;
    float pad1;
};

struct SceneView {
    const Node* nodes;
    const Prim* prims;
    const Material* materials;
    const Light* lights;
    int n_nodes, n_prims, n_spheres, n_materials;
    int n_lights;   // emissive primitives
    int n_sel;      // lights in the selection pool: n_lights + sky_light + sun_light
    float sky_pdf;  // kInv4Pi / n_sel: the sky's solid-angle pdf with its selection probability
    EnvParams env;
    CameraRig camera;
    int width, height;
};

struct Hit {
    float t;
    int prim;   // leaf slot, -1 for a miss
};

// ---------------------------------------------------------------------------------------------
// L1: intersection (include/wf/bvh.hpp:25-87)
// ---------------------------------------------------------------------------------------------
// Projected-distance discriminant; nearest root in (tmin, tmax).
// The reference's mid = -b / a is taken as -(b / a): round-to-nearest is symmetric in sign, so the
// two are the same float (and a NaN quotient misses either way); this saves a second division.
WF_HD bool hit_sphere(Vec3 center, float radius, const Ray& r, float* t_out) {
    Vec3 oc = r.o - center;
    float a = dot(r.d, r.d);
    float b = dot(oc, r.d);
    const float q = b / a;
    Vec3 w = oc - r.d * q;
    float disc = radius * radius - dot(w, w);
    if (disc < 0.0f) return false;
    float mid = -q;
    float sq = sqrtf(disc / a);
    float t = mid - sq;
    if (!(t > r.tmin && t < r.tmax)) {
        t = mid + sq;
        if (!(t > r.tmin && t < r.tmax)) return false;
    }
    *t_out = t;
    return true;
}

// Moller-Trumbore; det == 0 misses; barycentric bounds inclusive.
WF_HD bool hit_triangle(Vec3 p0, Vec3 p1, Vec3 p2, const Ray& r, float* t_out) {
    Vec3 e1 = p1 - p0, e2 = p2 - p0;
    Vec3 pv = cross(r.d, e2);
    float det = dot(e1, pv);
    if (det == 0.0f) return false;
    float inv = 1.0f / det;
    Vec3 tv = r.o - p0;
    float u = dot(tv, pv) * inv;
    if (u < 0.0f || u > 1.0f) return false;
    Vec3 qv = cross(tv, e1);
    float v = dot(r.d, qv) * inv;
    if (v < 0.0f || u + v > 1.0f) return false;
    float t = dot(e2, qv) * inv;
    if (!(t > r.tmin && t < r.tmax)) return false;
    *t_out = t;
    return true;
}

// Slab test against [tmin, tmax].
// The reference returns early once t0 > t1 after an axis and updates t0 = tn > t0 ? tn : t0,
// t1 = tf < t1 ? tf : t1. This form evaluates all three axes and compares once (no divergent
// branches on a GPU) and folds the updates into fmaxf/fminf; the result is the same:
//  - after t0 > t1 the remaining axes can only raise t0 or lower t1, so the final t0 <= t1 is
//    false exactly when an early exit would have been taken;
//  - t0 and t1 start finite and never become NaN, and a NaN slab bound (0 * inf) leaves them
//    unchanged both in the select form (its comparison is false) and in fmaxf/fminf (which return
//    the other operand); the two forms can differ only in the sign of a zero, which no comparison
//    sees. The swap of tn and tf stays a select: a NaN there must not move the other bound.
WF_HD bool hit_box(Vec3 lo, Vec3 hi, Vec3 o, Vec3 inv_d, float tmin, float tmax) {
    float tn = (lo.x - o.x) * inv_d.x, tf = (hi.x - o.x) * inv_d.x;
    const float ax = tn > tf ? tf : tn, bx = tn > tf ? tn : tf;
    tn = (lo.y - o.y) * inv_d.y;
    tf = (hi.y - o.y) * inv_d.y;
    const float ay = tn > tf ? tf : tn, by = tn > tf ? tn : tf;
    tn = (lo.z - o.z) * inv_d.z;
    tf = (hi.z - o.z) * inv_d.z;
    const float az = tn > tf ? tf : tn, bz = tn > tf ? tn : tf;
    const float t0 = fmaxf(fmaxf(fmaxf(tmin, ax), ay), az);
    const float t1 = fminf(fminf(fminf(tmax, bx), by), bz);
    return t0 <= t1;
}

// Slab test used by the GPU traversal. For a box with lo <= hi and a ray with a finite origin, 1/d
// components that are finite and nonzero and a tmin and tmax that are not NaN (slab_signed_ok), it
// equals hit_box bit for bit:
//  - no slab bound is NaN: with a finite origin and 1/d finite and nonzero, b - o is finite or
//    infinite (a box plane may be infinite) and its product with 1/d is never 0 * inf;
//  - hit_box's swap tn > tf ? tf : tn takes the far plane's bound exactly when 1/d < 0: for
//    1/d > 0, lo <= hi gives tn <= tf (subtracting o and multiplying by a positive number round
//    monotonically); for 1/d < 0, tn >= tf, and where they are equal both choices have the same
//    value up to the sign of a zero, which no comparison sees. neg_x/y/z are those signs, fixed
//    per ray;
//  - with no NaN anywhere, t0 <= fminf(fminf(fminf(tmax, bx), by), bz) is
//    t0 <= tmax && t0 <= min(bx, by, bz).
WF_HD bool hit_box_signed(const Vec3& lo, const Vec3& hi, const Vec3& o, const Vec3& inv_d, float tmin, float tmax,
                          bool neg_x, bool neg_y, bool neg_z) {
    const float lx = (lo.x - o.x) * inv_d.x, hx = (hi.x - o.x) * inv_d.x;
    const float ly = (lo.y - o.y) * inv_d.y, hy = (hi.y - o.y) * inv_d.y;
    const float lz = (lo.z - o.z) * inv_d.z, hz = (hi.z - o.z) * inv_d.z;
    const float ax = neg_x ? hx : lx, bx = neg_x ? lx : hx;
    const float ay = neg_y ? hy : ly, by = neg_y ? ly : hy;
    const float az = neg_z ? hz : lz, bz = neg_z ? lz : hz;
    const float t0 = fmaxf(fmaxf(fmaxf(tmin, ax), ay), az);
    const float t1 = fminf(fminf(bx, by), bz);
    return t0 <= tmax && t0 <= t1;
}

// The rays hit_box_signed serves (a traversal only ever lowers tmax to a hit distance, never NaN).
WF_HD bool slab_signed_ok(const Vec3& o, const Vec3& inv_d, float tmin, float tmax) {
    const float inf = __builtin_inff();
    return fabsf(inv_d.x) < inf && fabsf(inv_d.y) < inf && fabsf(inv_d.z) < inf && inv_d.x != 0.0f &&
           inv_d.y != 0.0f && inv_d.z != 0.0f && fabsf(o.x) < inf && fabsf(o.y) < inf && fabsf(o.z) < inf &&
           tmin == tmin && tmax == tmax;
}

WF_HD bool prim_is_sphere(const SceneView& s, const Prim& p) { return p.id < s.n_spheres; }

WF_HD bool hit_prim(const SceneView& s, const Prim& p, const Ray& r, float* t) {
    if (prim_is_sphere(s, p)) return hit_sphere(p.p0, p.radius, r, t);
    return hit_triangle(p.p0, p.p1, p.p2, r, t);
}

// Scene sources for traversal: where nodes come from and how a leaf primitive is tested. MemNodes
// reads both from the scene arrays; the device adds sources that serve the top of the tree and
// compact primitive records from LDS. Every source tests with the same hit_sphere/hit_triangle.
struct MemNodes {
    const Node* nodes;
    const Prim* prims;
    int n_spheres;
    WF_HD Node node(int n) const { return nodes[n]; }
    WF_HD bool hit(int slot, const Ray& r, float* t) const {
        const Prim& p = prims[slot];
        if (p.id < n_spheres) return hit_sphere(p.p0, p.radius, r, t);
        return hit_triangle(p.p0, p.p1, p.p2, r, t);
    }
};

WF_HD MemNodes mem_nodes(const SceneView& s) { return MemNodes{s.nodes, s.prims, s.n_spheres}; }

// Stackless skip-link traversal (include/wf/bvh.hpp:91-116): on a box hit descend to the left
// child (leaf: test its primitives, then follow skip); on a miss follow skip. tmax shrinks with
// each hit; a strictly nearer hit wins, so ties keep the earlier primitive in traversal order.
// anyhit returns on the first hit (shadow rays).
//
// The traversal is written as a resumable step so that a device kernel can interleave rays in one
// lane (a lane whose ray finished takes a new ray while the rest of its wave is still traversing)
// and still visit exactly the nodes and primitives, in exactly the order, of the loop below.
// Traversal state: the ray (its tmax is the closest hit distance so far), 1/d, the closest slot
// and the next node; node -1 means finished.
struct Traversal {
    Ray r;
    Vec3 inv_d;
    int prim;
    int node;
};

WF_HD Traversal traversal_begin(const Ray& r) {
    Traversal tr;
    tr.r = r;
    tr.inv_d = Vec3{1.0f / r.d.x, 1.0f / r.d.y, 1.0f / r.d.z};
    tr.prim = -1;
    tr.node = 0;
    return tr;
}

// Visits one node (a leaf with its primitives); sets node = -1 once the traversal is over.
template <class NodeSrc>
WF_HD void traversal_step(int n_nodes, const NodeSrc& src, Traversal& tr, bool anyhit) {
    const Node node = src.node(tr.node);
    int n;
    if (hit_box(node.lo, node.hi, tr.r.o, tr.inv_d, tr.r.tmin, tr.r.tmax)) {
        if (node.child < 0) {
            const int code = -(node.child + 1);
            const int first = code >> 2, cnt = (code & 3) + 1;
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
            n = node.skip;
        } else {
            n = node.child;
        }
    } else {
        n = node.skip;
    }
    tr.node = (n >= 0 && n < n_nodes) ? n : -1;
}

template <class NodeSrc>
WF_HD bool intersect_with(const SceneView& s, const NodeSrc& src, Ray r, Hit* h, bool anyhit) {
    Traversal tr = traversal_begin(r);
    if (s.n_nodes <= 0) tr.node = -1;
    while (tr.node >= 0) traversal_step(s.n_nodes, src, tr, anyhit);
    h->t = tr.r.tmax;
    h->prim = tr.prim;
    return h->prim >= 0;
}

WF_HD bool intersect(const SceneView& s, const Ray& r, Hit* h, bool anyhit = false) {
    return intersect_with(s, mem_nodes(s), r, h, anyhit);
}

// Oracle over every primitive in slot order. Slot order differs from id order, so ties between
// exactly equal t can resolve differently from traversal order; tests use it on tie-free rays.
WF_HD bool intersect_bruteforce(const SceneView& s, Ray r, Hit* h) {
    h->t = r.tmax;
    h->prim = -1;
    for (int p = 0; p < s.n_prims; ++p) {
        float t;
        if (hit_prim(s, s.prims[p], r, &t)) {
            r.tmax = t;
            h->t = t;
            h->prim = p;
        }
    }
    return h->prim >= 0;
}

// Valid hit: in range, finite, positive t (src/gpu/engine.cuh:44-48).
WF_HD bool hit_is_valid(const SceneView& s, const Hit& h) {
    if (h.prim < 0) return true;
    return h.prim < s.n_prims && is_finite1(h.t) && h.t > 0.0f;
}

// ---------------------------------------------------------------------------------------------
// L2: estimator (include/wf/path_core.hpp)
// ---------------------------------------------------------------------------------------------
struct PathState {
    Ray ray;
    Vec3 throughput;
    Vec3 L;
    int depth;
    int specular;    // 1 if the last event was a delta event (or the camera)
    float prev_pdf;  // solid-angle pdf of the last BSDF sample, 0 for delta
    PathKey key;
};

struct Surface {
    Vec3 p;
    Vec3 n;       // geometric normal facing against the ray
    Vec3 n_out;   // outward normal (sphere radial, triangle winding)
    int inside;   // 1 if the ray arrives from behind n_out
    int material;
};

enum LightKind : int { kLightArea = 0, kLightSky = 1, kLightSun = 2 };

struct LightSample {
    Vec3 p;          // area: point on the light
    Vec3 n;          // area: light normal; sky/sun: direction toward the light
    float pdf_solid; // includes the 1/n_sel selection probability; 0 if invalid
    Vec3 Le;
    int kind;
};

WF_HD float power_heuristic(float a, float b) {
    float a2 = a * a, b2 = b * b;
    float s = a2 + b2;
    return s > 0.0f ? a2 / s : 0.0f;
}

WF_HD float prim_area(const SceneView& s, const Prim& p) {
    if (prim_is_sphere(s, p)) {
        float r = p.radius;
        return 4.0f * kPi * r * r;
    }
    Vec3 e1 = p.p1 - p.p0;
    Vec3 e2 = p.p2 - p.p0;
    return 0.5f * length(cross(e1, e2));
}

WF_HD Vec3 sky_radiance(const EnvParams& e, Vec3 d) {
    if (e.sky_mode == 0) return e.env;
    float t = 0.5f * (d.y + 1.0f);
    return e.sky_lo * (1.0f - t) + e.sky_hi * t;
}

// Uniform direction on the unit sphere from two uniforms (z = 1 - 2 u1, phi = 2 pi u2), the
// parameterisation of the reference's sphere-light sampling (include/wf/path_core.hpp:115-119).
WF_HD Vec3 sphere_dir(float u1, float u2) {
    float z = 1.0f - 2.0f * u1;
    float r = sqrtf(max2(0.0f, 1.0f - z * z));
    float phi = 2.0f * kPi * u2;
    float sp, cp;
    sincos_f32(phi, &sp, &cp);
    return Vec3{r * cp, r * sp, z};
}

// Camera ray through pixel (px, py) of a width x height image seen by rig `cam`. A job of the
// batched engines can carry its own view (rig and size) over a shared scene.
WF_HD Ray camera_ray_view(const CameraRig& cam, int width, int height, int px, int py, PathKey key) {
    float jx = rng(key, 0, kDimPixelX), jy = rng(key, 0, kDimPixelY);
    float u = ((float)px + jx) / (float)width;
    float v = ((float)(height - 1 - py) + jy) / (float)height;
    Vec3 target = cam.lower_left + cam.horizontal * u + cam.vertical * v;
    Ray r;
    r.o = cam.origin;
    r.d = normalize(target - cam.origin);
    r.tmin = kRayEps;
    r.tmax = kInf;
    return r;
}

WF_HD Ray camera_ray(const SceneView& s, int px, int py, PathKey key) {
    return camera_ray_view(s.camera, s.width, s.height, px, py, key);
}

WF_HD PathState init_path_keyed(const CameraRig& cam, int width, int height, int px, int py, PathKey key) {
    PathState ps;
    ps.key = key;
    ps.ray = camera_ray_view(cam, width, height, px, py, ps.key);
    ps.throughput = Vec3{1.0f, 1.0f, 1.0f};
    ps.L = Vec3{0.0f, 0.0f, 0.0f};
    ps.depth = 0;
    ps.specular = 1;
    ps.prev_pdf = 0.0f;
    return ps;
}

WF_HD PathState init_path_view(const CameraRig& cam, int width, int height, int px, int py, uint32_t sample,
                               uint32_t seed) {
    return init_path_keyed(cam, width, height, px, py, make_path_key(seed, (uint32_t)(py * width + px), sample));
}

WF_HD PathState init_path(const SceneView& s, int px, int py, uint32_t sample, uint32_t seed) {
    return init_path_view(s.camera, s.width, s.height, px, py, sample, seed);
}

// The geometric normal is v / l: v = p - center and l = radius for a sphere, v = e1 x e2 and
// l = |v| for a triangle (normalize: v itself when |v| = 0). One division serves both kinds, so a
// wave whose lanes hit spheres and triangles divides once.
WF_HD Surface make_surface(const SceneView& s, const Ray& ray, const Hit& h) {
    Surface f;
    const Prim& pr = s.prims[h.prim];
    f.p = ray.o + ray.d * h.t;
    const bool sph = prim_is_sphere(s, pr);
    Vec3 v;
    float l;
    if (sph) {
        v = f.p - pr.p0;
        l = pr.radius;
    } else {
        Vec3 e1 = pr.p1 - pr.p0;
        Vec3 e2 = pr.p2 - pr.p0;
        v = cross(e1, e2);
        l = length(v);
    }
    f.n_out = (sph || l > 0.0f) ? v / l : v;
    f.material = pr.material;
    f.inside = dot(f.n_out, ray.d) > 0.0f ? 1 : 0;
    f.n = f.inside ? -f.n_out : f.n_out;
    return f;
}

// Light selection over the pool (area lights, then the sky, then the sun), then a sample of the
// chosen light converted to a solid-angle pdf from `from`. The area branch is the reference's
// sample_light (include/wf/path_core.hpp:104-140) with n_sel in place of n_lights.
WF_HD LightSample sample_light(const SceneView& s, Vec3 from, PathKey key, uint32_t depth) {
    LightSample ls;
    ls.pdf_solid = 0.0f;
    ls.p = from;
    ls.n = Vec3{0, 0, 1};
    ls.Le = Vec3{0, 0, 0};
    ls.kind = kLightArea;
    if (s.n_sel <= 0) return ls;
    // With one light the pick is 0 for every draw in [0, 1), so the draw is skipped.
    int idx = s.n_sel == 1 ? 0 : (int)(rng(key, depth, kDimLightPick) * (float)s.n_sel);
    if (idx >= s.n_sel) idx = s.n_sel - 1;
    float u1 = rng(key, depth, kDimLightU), u2 = rng(key, depth, kDimLightV);
    if (idx >= s.n_lights) {
        if (s.env.sky_light && idx == s.n_lights) {
            ls.kind = kLightSky;
            ls.n = sphere_dir(u1, u2);
            ls.Le = sky_radiance(s.env, ls.n);
            ls.pdf_solid = s.sky_pdf;
        } else {
            ls.kind = kLightSun;
            ls.n = s.env.sun_dir;
            ls.Le = s.env.sun_E;
            ls.pdf_solid = 1.0f / (float)s.n_sel;   // selection probability of a delta light
        }
        return ls;
    }
    const Prim& pr = s.prims[s.lights[idx].slot];
    float area = prim_area(s, pr);
    if (prim_is_sphere(s, pr)) {
        Vec3 dir = sphere_dir(u1, u2);
        ls.n = dir;
        ls.p = pr.p0 + dir * pr.radius;
    } else {
        float su = sqrtf(u1);
        float b0 = 1.0f - su, b1 = u2 * su;
        Vec3 p0 = pr.p0, p1 = pr.p1, p2 = pr.p2;
        ls.p = p0 * (1.0f - b0 - b1) + p1 * b0 + p2 * b1;
        ls.n = normalize(cross(p1 - p0, p2 - p0));
    }
    ls.Le = s.materials[pr.material].emission;
    Vec3 d = ls.p - from;
    float dist2 = length2(d);
    float dist = sqrtf(dist2);
    if (dist <= 0.0f) return ls;
    float cosl = fabsf(dot(ls.n, d)) / dist;
    if (cosl <= 1e-6f) return ls;
    ls.pdf_solid = (dist2 / (cosl * area)) / (float)s.n_sel;
    return ls;
}

// Solid-angle pdf (with selection probability) of hitting light primitive `slot` by BSDF sampling.
WF_HD float light_pdf_at_hit(const SceneView& s, int slot, float t, Vec3 dir, Vec3 n_out) {
    float cosl = fabsf(dot(n_out, dir));
    if (cosl <= 1e-6f) return 0.0f;
    return ((t * t) / (cosl * prim_area(s, s.prims[slot]))) / (float)s.n_sel;
}

// Bounce steps 1-3 (include/wf/path_core.hpp:151-170): escaped rays gather the sky, hits gather
// emission with the MIS weight; returns the material type to shade, or -1 if the path ends.
WF_HD int classify(const SceneView& s, PathState& ps, const Hit& h, int max_depth, Surface* surf) {
    if (h.prim < 0) {
        if (s.env.sky_light && !ps.specular && ps.prev_pdf > 0.0f) {
            Vec3 Le = sky_radiance(s.env, ps.ray.d);
            float w = power_heuristic(ps.prev_pdf, s.sky_pdf);
            add_to(ps.L, ps.throughput * Le * w);
        } else {
            add_to(ps.L, ps.throughput * sky_radiance(s.env, ps.ray.d));
        }
        return -1;
    }
    *surf = make_surface(s, ps.ray, h);
    const Material& m = s.materials[surf->material];
    if (m.emission.x > 0.0f || m.emission.y > 0.0f || m.emission.z > 0.0f) {
        float w = 1.0f;
        if (!ps.specular && ps.prev_pdf > 0.0f) {
            float pl = light_pdf_at_hit(s, h.prim, h.t, ps.ray.d, surf->n_out);
            w = power_heuristic(ps.prev_pdf, pl);
        }
        add_to(ps.L, ps.throughput * m.emission * w);
    }
    if (m.type == kEmissive) return -1;
    if (ps.depth >= max_depth) return -1;
    return m.type;
}

// Russian roulette from depth 3 (include/wf/path_core.hpp:173-179).
WF_HD bool russian_roulette(PathState& ps) {
    if (ps.depth < 3) return true;
    float q = clampf(max_component(ps.throughput), 0.05f, 0.95f);
    if (rng(ps.key, (uint32_t)ps.depth, kDimRoulette) >= q) return false;
    ps.throughput = ps.throughput / q;
    return true;
}

WF_HD void set_next_ray(PathState& ps, Vec3 origin, Vec3 dir) {
    ps.ray.o = origin;
    ps.ray.d = dir;
    ps.ray.tmin = kRayEps;
    ps.ray.tmax = kInf;
}

// Lambertian (include/wf/path_core.hpp:192-228): next-event estimation (shadow ray and its
// contribution, added by the caller if unoccluded, before the next classify), cosine-sampled
// scatter, roulette. Returns true if the path continues.
// Next-event estimation of a Lambertian hit from `org`: the shadow ray and its contribution
// (shade_lambert's first half). Sets *has_shadow to 0 when there is none.
WF_HD void lambert_nee(const SceneView& s, const PathState& ps, const Surface& f, const Material& m, Vec3 org,
                       Ray* shadow, Vec3* contrib, int* has_shadow) {
    *has_shadow = 0;
    if (s.n_sel <= 0) return;
    LightSample ls = sample_light(s, org, ps.key, (uint32_t)ps.depth);
    if (!(ls.pdf_solid > 0.0f)) return;
    Vec3 wi;
    float tmax;
    if (ls.kind == kLightArea) {
        Vec3 to = ls.p - org;
        float dist = length(to);
        wi = to / dist;
        tmax = dist * (1.0f - 1e-3f);
    } else {
        wi = ls.n;
        tmax = kInf;
    }
    float cos_s = dot(f.n, wi);
    if (cos_s > 0.0f) {
        float pb = cos_s * kInvPi;
        float w = ls.kind == kLightSun ? 1.0f : power_heuristic(ls.pdf_solid, pb);
        *contrib = ps.throughput * m.albedo * ls.Le * (cos_s * kInvPi * w / ls.pdf_solid);
        shadow->o = org;
        shadow->d = wi;
        shadow->tmin = kRayEps;
        shadow->tmax = tmax;
        *has_shadow = 1;
    }
}

WF_HD bool shade_lambert(const SceneView& s, PathState& ps, const Surface& f, Ray* shadow,
                         Vec3* contrib, int* has_shadow) {
    const Material& m = s.materials[f.material];
    uint32_t d = (uint32_t)ps.depth;
    Vec3 org = f.p + f.n * kRayEps;
    lambert_nee(s, ps, f, m, org, shadow, contrib, has_shadow);
    float u1 = rng(ps.key, d, kDimBsdfU), u2 = rng(ps.key, d, kDimBsdfV);
    float r = sqrtf(u1), phi = 2.0f * kPi * u2;
    float sp, cp;
    sincos_f32(phi, &sp, &cp);
    Vec3 t, b;
    make_basis(f.n, &t, &b);
    float z = sqrtf(max2(0.0f, 1.0f - u1));
    Vec3 dir = normalize(t * (r * cp) + b * (r * sp) + f.n * z);
    ps.throughput = ps.throughput * m.albedo;
    ps.specular = 0;
    ps.prev_pdf = dot(dir, f.n) * kInvPi;
    ps.depth += 1;
    set_next_ray(ps, org, dir);
    return russian_roulette(ps);
}

// Mirror about the facing normal (include/wf/path_core.hpp:230-240); with fuzz > 0 the mirror
// direction is perturbed by fuzz times a uniform unit vector (dims 9-10). Dies below the surface.
WF_HD bool shade_metal(const SceneView& s, PathState& ps, const Surface& f) {
    const Material& m = s.materials[f.material];
    Vec3 dir = reflect(ps.ray.d, f.n);
    if (m.fuzz > 0.0f) {
        uint32_t d = (uint32_t)ps.depth;
        Vec3 jit = sphere_dir(rng(ps.key, d, kDimFuzzU), rng(ps.key, d, kDimFuzzV));
        dir = dir + jit * m.fuzz;
    }
    if (dot(dir, f.n) <= 0.0f) return false;
    ps.throughput = ps.throughput * m.albedo;
    ps.specular = 1;
    ps.prev_pdf = 0.0f;
    ps.depth += 1;
    set_next_ray(ps, f.p + f.n * kRayEps, normalize(dir));
    return russian_roulette(ps);
}

// Exact unpolarised Fresnel with total internal reflection (include/wf/path_core.hpp:243-251).
WF_HD float fresnel_dielectric(float cos_i, float eta_i, float eta_t) {
    float sin_t = eta_i / eta_t * sqrtf(max2(0.0f, 1.0f - cos_i * cos_i));
    if (sin_t >= 1.0f) return 1.0f;
    float cos_t = sqrtf(max2(0.0f, 1.0f - sin_t * sin_t));
    float rpar = (eta_t * cos_i - eta_i * cos_t) / (eta_t * cos_i + eta_i * cos_t);
    float rperp = (eta_i * cos_i - eta_t * cos_t) / (eta_i * cos_i + eta_t * cos_t);
    return 0.5f * (rpar * rpar + rperp * rperp);
}

WF_HD bool shade_dielectric(const SceneView& s, PathState& ps, const Surface& f) {
    const Material& m = s.materials[f.material];
    float eta_i = f.inside ? m.ior : 1.0f;
    float eta_t = f.inside ? 1.0f : m.ior;
    float cos_i = clampf(-dot(ps.ray.d, f.n), 0.0f, 1.0f);
    float F = fresnel_dielectric(cos_i, eta_i, eta_t);
    float u = rng(ps.key, (uint32_t)ps.depth, kDimDielectric);
    Vec3 org, dir;
    if (u < F) {
        dir = reflect(ps.ray.d, f.n);
        org = f.p + f.n * kRayEps;
    } else {
        float eta = eta_i / eta_t;
        float sin2_t = eta * eta * (1.0f - cos_i * cos_i);
        float cos_t = sqrtf(max2(0.0f, 1.0f - sin2_t));
        dir = ps.ray.d * eta + f.n * (eta * cos_i - cos_t);
        org = f.p - f.n * kRayEps;
    }
    ps.throughput = ps.throughput * m.albedo;
    ps.specular = 1;
    ps.prev_pdf = 0.0f;
    ps.depth += 1;
    set_next_ray(ps, org, normalize(dir));
    return russian_roulette(ps);
}

// shade_lambert, shade_metal or shade_dielectric as `mat` (a type classify returned) says, in one
// pass: each step that two materials share is written once, with its operands chosen per material
// (the first scatter draw; the second draw, square root and sin/cos behind the Lambertian and the
// metal-fuzz directions; the final normalisation, state update and roulette). A wave whose lanes
// hold different materials so runs each such step once instead of once per material. Every lane
// performs the operations of its own material's function, in the same order on the same values,
// so the results are the same bits.
WF_HD bool shade_surface(const SceneView& s, PathState& ps, const Surface& f, int mat, Ray* shadow, Vec3* contrib,
                         int* has_shadow) {
    const Material& m = s.materials[f.material];
    const uint32_t d = (uint32_t)ps.depth;
    const bool lam = mat == kLambert, met = mat == kMetal;   // otherwise dielectric
    const Vec3 off = f.n * kRayEps;
    Vec3 org = f.p + off;
    *has_shadow = 0;
    if (lam) lambert_nee(s, ps, f, m, org, shadow, contrib, has_shadow);
    // Lambert's u1, the fuzz direction's u1, or the dielectric's reflect-or-refract draw.
    const float u1 = rng(ps.key, d, lam ? kDimBsdfU : (met ? kDimFuzzU : kDimDielectric));
    Vec3 raw;
    if (lam || (met && m.fuzz > 0.0f)) {
        // Lambert: r = sqrt(u1), z = sqrt(max(0, 1 - u1)); the fuzz direction (sphere_dir):
        // z = 1 - 2 u1, r = sqrt(max(0, 1 - z z)); both at the angle 2 pi u2.
        const float u2 = rng(ps.key, d, lam ? kDimBsdfV : kDimFuzzV);
        const float zs = 1.0f - 2.0f * u1;
        const float r = sqrtf(lam ? u1 : max2(0.0f, 1.0f - zs * zs));
        const float phi = 2.0f * kPi * u2;
        float sp, cp;
        sincos_f32(phi, &sp, &cp);
        if (lam) {
            Vec3 t, b;
            make_basis(f.n, &t, &b);
            const float z = sqrtf(max2(0.0f, 1.0f - u1));
            raw = t * (r * cp) + b * (r * sp) + f.n * z;
        } else {
            raw = reflect(ps.ray.d, f.n) + Vec3{r * cp, r * sp, zs} * m.fuzz;
        }
    } else if (met) {
        raw = reflect(ps.ray.d, f.n);
    } else {
        const float eta_i = f.inside ? m.ior : 1.0f;
        const float eta_t = f.inside ? 1.0f : m.ior;
        const float cos_i = clampf(-dot(ps.ray.d, f.n), 0.0f, 1.0f);
        const float F = fresnel_dielectric(cos_i, eta_i, eta_t);
        if (u1 < F) {
            raw = reflect(ps.ray.d, f.n);
        } else {
            const float eta = eta_i / eta_t;
            const float sin2_t = eta * eta * (1.0f - cos_i * cos_i);
            const float cos_t = sqrtf(max2(0.0f, 1.0f - sin2_t));
            raw = ps.ray.d * eta + f.n * (eta * cos_i - cos_t);
            org = f.p - off;
        }
    }
    if (met && dot(raw, f.n) <= 0.0f) return false;   // scattered below the surface
    const Vec3 dir = normalize(raw);
    ps.throughput = ps.throughput * m.albedo;
    ps.specular = lam ? 0 : 1;
    ps.prev_pdf = lam ? dot(dir, f.n) * kInvPi : 0.0f;
    ps.depth += 1;
    set_next_ray(ps, org, dir);
    return russian_roulette(ps);
}

// Per-path ray counts, kept by every arm so the integer work can be compared exactly.
struct RayCount {
    uint32_t extension;
    uint32_t shadow;
    uint32_t invalid;
};

// One bounce after the extension hit is known: classify, shade, trace the shadow ray.
// Returns true if the path continues. Used by the CPU arm and the megakernel.
template <class NodeSrc>
WF_HD bool bounce_after_hit(const SceneView& s, const NodeSrc& src, PathState& ps, Hit h,
                            int max_depth, RayCount* rc) {
    if (!hit_is_valid(s, h)) {
        rc->invalid += 1;
        h.prim = -1;
    }
    Surface f;
    int mat = classify(s, ps, h, max_depth, &f);
    if (mat < 0) return false;
    bool alive;
    if (mat == kLambert) {
        Ray sh;
        Vec3 c;
        int has;
        alive = shade_lambert(s, ps, f, &sh, &c, &has);
        if (has) {
            Hit sh_hit;
            rc->shadow += 1;
            if (!intersect_with(s, src, sh, &sh_hit, true)) add_to(ps.L, c);
        }
    } else if (mat == kMetal) {
        alive = shade_metal(s, ps, f);
    } else {
        alive = shade_dielectric(s, ps, f);
    }
    return alive;
}

// Complete path for one (pixel, sample) (include/wf/path_core.hpp:279-304).
template <class NodeSrc>
WF_HD Vec3 trace_path_with(const SceneView& s, const NodeSrc& src, int px, int py, uint32_t sample,
                           uint32_t seed, int max_depth, RayCount* rc) {
    PathState ps = init_path(s, px, py, sample, seed);
    for (;;) {
        Hit h;
        intersect_with(s, src, ps.ray, &h, false);
        rc->extension += 1;
        if (!bounce_after_hit(s, src, ps, h, max_depth, rc)) break;
    }
    return ps.L;
}

WF_HD Vec3 trace_path(const SceneView& s, int px, int py, uint32_t sample, uint32_t seed,
                      int max_depth, RayCount* rc) {
    return trace_path_with(s, mem_nodes(s), px, py, sample, seed, max_depth, rc);
}

}  // namespace wf
