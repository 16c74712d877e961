/* Oracle: a CPU restatement of the reference path tracer, in plain C (C11), for checking the
 * product. TEST INFRASTRUCTURE ONLY: it is loaded by tests/, by __graft_entry__.smoke() and by
 * bench.py's cpu_baseline leg, never by the product library, which must fail loudly instead of
 * falling back to it.
 *
 * It follows the reference (paths relative to wavefront-vs-megakernel-path-tracing/) in the
 * reference's own data layout -- 48-byte preorder BvhNode with implicit left child N+1 and a
 * prim_order indirection -- not the product's device layout, so a bitwise match between the two
 * also checks the product's layout conversion. Trigonometry uses the C library's binary64
 * sin/cos rounded to binary32; the product evaluates its own polynomial, and tests pin the two
 * against each other on all 2^24 angles 2*pi*u the estimator can produce. A second build
 * (liboracle_libm.so, OR_LIBM_F32_TRIG) uses the C library's binary32 sinf/cosf instead, as the
 * reference's host code does, and reproduces the reference's own renders bit for bit.
 *
 * Scene arrays use the product C ABI's layout (include/wfamd.h, wf_scene_desc).
 */
#ifndef WF_ORACLE_H
#define WF_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct or_scene or_scene;

/* Builds lights (src/scene.cpp:68-91) and the BVH (src/bvh_build.cpp:60-251). NULL on error. */
or_scene* or_scene_new(const float* spheres, int n_spheres, const float* vertices, int n_vertices,
                       const int* triangles, int n_triangles, const float* materials, int n_materials,
                       const float* env20, const float* camera10, int width, int height);
void or_scene_free(or_scene* s);
int or_scene_n_nodes(const or_scene* s);
int or_scene_n_lights(const or_scene* s);
/* nodes: 12 ints per node (lo xyz bits, skip, hi xyz bits, prim_start, prim_count, 0, 0, 0). */
void or_scene_bvh(const or_scene* s, int* nodes, int* prim_order);
void or_camera_rig(const or_scene* s, float* rig12);
/* Another view of the same geometry (the reference builds a scene per camera, whose BVH depends on
 * the geometry only): camera (10 floats, as or_scene_new) and image size; the BVH is kept. */
void or_scene_set_view(or_scene* s, const float* camera10, int width, int height);
uint64_t or_scene_hash(const or_scene* s);

/* RNG (include/wf/rng.hpp:8-42). */
uint32_t or_hash_u32(uint32_t x);
void or_path_key(uint32_t seed, uint32_t pixel, uint32_t sample, uint32_t* k2);
float or_rng(uint32_t k0, uint32_t k1, uint32_t depth, uint32_t dim);
void or_sincos(const float* x, int n, float* s, float* c);

/* Scattered direction of a metal surface (include/wf/path_core.hpp:230-240, plus the fuzz
 * extension): mirror of d about the facing normal n, plus fuzz times a uniform unit vector drawn
 * from dims 9-10 at `depth`, normalized; returns 0 (out3 zero) if it points below the surface. */
int or_metal_dir(const float* d3, const float* n3, float fuzz, uint32_t k0, uint32_t k1, uint32_t depth, float* out3);

/* Closest or any hit for n rays (8 floats each: o xyz, tmin, d xyz, tmax); prim ids (-1 miss). */
void or_intersect(const or_scene* s, const float* rays, int n, int anyhit, float* t, int* prim);
void or_intersect_brute(const or_scene* s, const float* rays, int n, float* t, int* prim);

/* Film sums of samples [first, first+count) added in sample order (src/cpu_renderer.cpp:55-74);
 * counts[0..2] += extension rays, shadow rays, invalid hits. threads <= 0 -> 1. */
void or_render(const or_scene* s, int first_sample, int count, uint32_t seed, int max_depth, int threads,
               float* rgb_sum, float* lum_sum, float* lum2_sum, uint64_t* counts);

/* Radiance of single paths (px, py, sample triples), trace_path (include/wf/path_core.hpp:279-304);
 * counts[0..2] += extension rays, shadow rays, invalid hits. */
void or_trace_paths(const or_scene* s, const int* pxys, int n, uint32_t seed, int max_depth, float* L, uint64_t* counts);

/* One wavefront pass (src/gpu/wavefront.cu:433-601) restated on the CPU with material queues in
 * path order and a stable compaction. Writes up to max_records records of 6 words (active,
 * lambert, metal, dielectric, shadow, next_active) and the FNV-1a 64 hash of each compacted list;
 * adds the pass to the film sums; returns the number of bounces. queue_capacity 0 = no limit. */
int or_wavefront_pass(const or_scene* s, uint32_t sample, uint32_t seed, int max_depth, int queue_capacity,
                      uint32_t* records, uint64_t* hashes, int max_records, float* rgb_sum, float* lum_sum,
                      float* lum2_sum, uint64_t* counts);

/* mean = sum / spp; unbiased luminance variance in binary64. */
void or_finalize(const float* rgb_sum, const float* lum_sum, const float* lum2_sum, int pixels, int spp, float* rgb,
                 float* lum_var);
/* compare_results (src/film.cpp:45-113): out[0] max_abs_rel, out[1] mean_rel, out[2] frac within
 * 3 SE, out[3] non-finite channels. */
void or_compare(const float* rgb_a, const float* var_a, int spp_a, const float* rgb_b, const float* var_b, int spp_b,
                int pixels, double* out4);

#ifdef __cplusplus
}
#endif
#endif
