/* C ABI of the MI355X path-tracing engine (libwfamd.so).
 *
 * The reference exposes its engine as C++ only: wf::GpuRenderer, wf::render_cpu and the queue
 * primitives (no plugin registry, no C ABI). Each entry point below names the reference interface
 * it replaces (paths relative to wavefront-vs-megakernel-path-tracing/). Every function returns
 * WF_OK or an error code; wf_last_error() gives the message of the last failure on the calling
 * thread. Invalid data met while rendering (queue overflow, invalid hits) is counted in
 * wf_render_stats, not reported as an error. Handles are owned by the caller and freed with the
 * matching destroy function; host buffers are caller-owned (pointer + size).
 */
#ifndef WFAMD_H
#define WFAMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum wf_status {
    WF_OK = 0,
    WF_ERR_INVALID = 1,   /* bad argument (the reference's std::invalid_argument / cudaErrorInvalidValue) */
    WF_ERR_RUNTIME = 2,   /* runtime or device failure (the reference's std::runtime_error) */
    WF_ERR_BUDGET = 3     /* serving: the request does not fit the device-memory budget */
};

enum wf_mode {
    WF_MODE_CPU = 0,         /* render_cpu, include/wf/renderer.hpp:39 */
    WF_MODE_MEGAKERNEL = 1,  /* GpuMode::Megakernel, include/wf/renderer.hpp:41 */
    WF_MODE_WAVEFRONT = 2,   /* GpuMode::Wavefront, include/wf/renderer.hpp:41 */
    WF_MODE_PERSISTENT = 3   /* persistent batched wavefront (no reference counterpart) */
};

typedef struct wf_scene wf_scene;
typedef struct wf_renderer wf_renderer;
typedef struct wf_scan_ws wf_scan_ws;
typedef struct wf_server wf_server;

/* RenderParams, include/wf/renderer.hpp:11-21, plus persistent-engine tuning (0 = default). */
typedef struct wf_render_params {
    int spp;
    uint32_t seed;
    int max_depth;
    int threads;          /* CPU arm: worker threads, 0 = all cores */
    int count_rays;       /* accepted for compatibility; rays are always counted */
    int stage_timing;     /* wavefront: per-stage event timing (synchronising) */
    int queue_capacity;   /* wavefront test hook: 0 = path count; smaller -> counted overflows */
    int trace_bounces;    /* wavefront: record per-bounce integer work, see wf_renderer_trace */
    int block_size;       /* persistent: threads per workgroup (256, 512, 1024; 0 = 1024 for one job, 512 for several) */
    int pool_paths;       /* persistent: path slots (= queue capacity) */
    int scan_tile;        /* plain wavefront: compaction tile (elements) */
    int lds_nodes;        /* megakernel / plain wavefront: BVH nodes staged in LDS, -1 = default */
    int lds_kib;          /* persistent: LDS per workgroup for BVH nodes + primitive records (KiB; 0 = 64 for one job, 32 for several) */
    int refill;           /* persistent: a wave refetches rays when fewer lanes than this traverse */
    int leaf_batch;       /* persistent: parked leaves that start a wave's leaf-test round */
    int dir_bins;         /* persistent: survivor grouping, 1 none or 8 direction octants */
    int shade_block;      /* persistent: threads per shade workgroup = compaction round (128..1024) */
    int graph_iters;      /* persistent: pool iterations per captured graph launch (0 = max_depth + 2
                           * when every sample fits the pool, else 16; 8 in the serving loop;
                           * rounded up to even) */
} wf_render_params;

/* RenderStats, include/wf/renderer.hpp:23-37. stage_ms: extend, classify, shade_lambert,
 * shade_metal, shade_dielectric, shadow, compact, other. With stage_timing set, stage_items and
 * stage_launches give the items each stage processed (paths, queue entries or shadow rays) and
 * its kernel launches. The persistent engine fuses classify into extend and shades the three
 * materials in one kernel, reported under extend and shade_lambert. */
typedef struct wf_render_stats {
    double render_ms;
    double total_ms;
    uint64_t extension_rays;
    uint64_t shadow_rays;
    int64_t queue_high_water[5]; /* lambert, metal, dielectric, shadow, active */
    int64_t queue_overflows;
    int64_t invalid_hits;
    uint64_t peak_vram_bytes;
    double stage_ms[8];
    int64_t iterations;
    int64_t stage_items[8];
    int64_t stage_launches[8];
    int64_t stage_bytes[8];      /* persistent engine: algorithmic device-memory bytes per stage */
    int64_t stage_threads[8];    /* persistent engine: threads per launch of each stage (persistent
                                  * grid x workgroup size: the occupancy the kernel runs at) */
} wf_render_stats;

typedef struct wf_scene_info {
    int width, height;
    int n_spheres, n_triangles, n_vertices, n_materials, n_lights, n_nodes;
    int bvh_leaves, bvh_internal, bvh_max_depth;
    double bvh_sah_cost;
    uint64_t hash;
} wf_scene_info;

/* Scene arrays in the reference layout (include/wf/scene.hpp:12-32), for building a scene from
 * the caller's data and for exporting one. spheres: 8 words each (cx, cy, cz, radius, material
 * as int bits, 3 pad); vertices: 3 floats; triangles: 4 ints (i0, i1, i2, material); materials:
 * 12 words (type int, ior, fuzz, pad, albedo xyz, pad, emission xyz, pad). env: 20 floats
 * (env xyz, sky_mode int, sky_lo xyz, sky_light int, sky_hi xyz, sun_light int, sun_dir xyz, pad,
 * sun_E xyz, pad). camera: 10 floats (position, look_at, up, vfov_deg). */
typedef struct wf_scene_desc {
    const float* spheres; int n_spheres;
    const float* vertices; int n_vertices;
    const int* triangles; int n_triangles;
    const float* materials; int n_materials;
    const float* env;
    const float* camera;
    int width, height;
} wf_scene_desc;

typedef struct wf_device_facts {
    char name[256];
    char arch[64];
    int compute_units;
    int runtime_version;
    int driver_version;
    uint64_t total_memory_bytes;
    uint64_t free_memory_bytes;
} wf_device_facts;

/* One bounce of the plain wavefront: counts and the FNV-1a 64 hash of the compacted list. */
typedef struct wf_bounce_trace {
    uint32_t active, lambert, metal, dielectric, shadow, next_active;
    uint64_t list_hash;
} wf_bounce_trace;

const char* wf_last_error(void);
const char* wf_version(void);

/* ---- scenes: make_scene src/scenes.cpp:29, finalize_scene src/scene.cpp:58,
 *      scene_hash src/scene.cpp:107, load from arrays (the reference's Scene aggregate) ---- */
int wf_scene_create(const char* name, int width, int height, wf_scene** out);
int wf_scene_from_arrays(const wf_scene_desc* desc, wf_scene** out);
int wf_scene_orbit(wf_scene* scene, float yaw_deg);  /* viewer orbit, src/viewer/viewer.cpp:86-90 */
/* A copy of `scene` seen through another view: image size width x height and, with orbit = 1, the
 * camera orbited to yaw_deg. Geometry and BVH are copied, not rebuilt. This is the standalone
 * scene of one wf_job view (the reference builds a new scene per camera, src/viewer/viewer.cpp:114-120). */
int wf_scene_view(const wf_scene* scene, int width, int height, float yaw_deg, int orbit, wf_scene** out);
void wf_scene_destroy(wf_scene* scene);
int wf_scene_info_get(const wf_scene* scene, wf_scene_info* out);
/* Exports (counts in elements; pass NULL to skip an array). nodes: reference BvhNode layout,
 * 12 words each (lo xyz, skip, hi xyz, prim_start, prim_count, 3 pad), src/bvh_build.cpp:191-226. */
int wf_scene_export(const wf_scene* scene, float* spheres, float* vertices, int* triangles, float* materials,
                    float* env, float* camera);
int wf_scene_bvh(const wf_scene* scene, int* nodes, int* prim_order);
/* Camera rig (origin, lower_left, horizontal, vertical: 12 floats), include/wf/scene.hpp:129-141. */
int wf_scene_camera_rig(const wf_scene* scene, float* rig12);

/* ---- renderer: wf::GpuRenderer, include/wf/renderer.hpp:44-67 ---- */
int wf_renderer_create(const wf_scene* scene, int mode, wf_renderer** out);
void wf_renderer_destroy(wf_renderer* r);
/* reset(); run_passes(0, spp); download(spp). rgb: 3*W*H floats, lum_var: W*H (either may be NULL). */
int wf_renderer_render(wf_renderer* r, const wf_render_params* p, wf_render_stats* stats, float* rgb, float* lum_var);
int wf_renderer_reset(wf_renderer* r);
int wf_renderer_run_passes(wf_renderer* r, int first_sample, int count, const wf_render_params* p,
                           wf_render_stats* stats);
int wf_renderer_download(wf_renderer* r, int spp_done, float* rgb, float* lum_var);
int wf_renderer_download_mean(wf_renderer* r, int spp_done, float* rgb);
/* Raw film sums (rgb 3*W*H, lum W*H, lum2 W*H); any may be NULL. */
int wf_renderer_download_sums(wf_renderer* r, float* rgb, float* lum, float* lum2);
uint64_t wf_renderer_device_bytes(const wf_renderer* r);
const char* wf_renderer_device_name(const wf_renderer* r);
/* Per-bounce trace of the last run_passes with trace_bounces set; returns the record count. */
int wf_renderer_trace(const wf_renderer* r, wf_bounce_trace* out, int capacity);

/* ---- CPU arm: render_cpu, include/wf/renderer.hpp:39 ---- */
int wf_render_cpu(const wf_scene* scene, const wf_render_params* p, wf_render_stats* stats, float* rgb,
                  float* lum_var);
/* CPU arm, raw film sums of samples [first_sample, first_sample + count) (added to the buffers). */
int wf_render_cpu_sums(const wf_scene* scene, int first_sample, int count, const wf_render_params* p,
                       wf_render_stats* stats, float* rgb, float* lum, float* lum2);
/* Shared finalize: mean = sum / spp, unbiased luminance variance (binary64). */
int wf_finalize(const float* rgb_sum, const float* lum_sum, const float* lum2_sum, int pixels, int spp, float* rgb,
                float* lum_var);

/* ---- device facts: query_cuda_device, include/wf/renderer.hpp:70-74 ---- */
int wf_query_device(wf_device_facts* out);

/* ---- primitives: src/gpu/queues.cuh:29-54 (device pointers, stream = hipStream_t or NULL) ---- */
int wf_scan_workspace_create(int capacity, wf_scan_ws** out);
void wf_scan_workspace_destroy(wf_scan_ws* ws);
int wf_scan_exclusive_u32(const uint32_t* d_in, uint32_t* d_out, int n, uint32_t* d_total, wf_scan_ws* ws,
                          void* stream, int tile);
int wf_compact_u32(const uint32_t* d_values, const uint32_t* d_flags, uint32_t* d_out, uint32_t* d_count,
                   uint32_t* d_scratch, int n, wf_scan_ws* ws, void* stream, int tile);
int wf_test_queue_push(const uint32_t* d_flags, int n, uint32_t* d_items, uint32_t capacity, uint32_t* d_counter,
                       uint32_t* d_overflow, void* stream);
/* Stable compaction that allocates nothing: the result of wf_compact_u32, with the per-tile partial
 * sums in a static array of the library (WF_NOALLOC_MAX_N / 512 + 1 words = 512 KiB on each GPU it
 * runs on, held until the process ends) instead of a workspace. No counterpart in the reference:
 * it is for callers that cannot allocate device memory (code under HIP graph capture, serving
 * loops, harnesses that forbid hipMalloc). n above WF_NOALLOC_MAX_N -> WF_ERR_INVALID with no side
 * effects. The array is shared by every call on one GPU: calls that overlap on one GPU (different
 * streams or threads) race; calls in one stream order are safe. tile: 0 (default 2048), 512, 1024,
 * 2048 or 4096. */
#define WF_NOALLOC_MAX_N (1 << 26)
int wf_compact_u32_noalloc(const uint32_t* d_values, const uint32_t* d_flags, uint32_t* d_out, uint32_t* d_count, int n,
                           void* stream, int tile);

/* ---- estimator test hooks (host buffers). device = 0 evaluates on the host, 1 on the GPU.
 *      The reference tests the same functions directly from C++ (tests/test_math_rng.cpp,
 *      tests/test_intersect.cpp, tests/test_bvh.cpp). ---- */
/* sin and cos of n binary32 angles with the estimator's own evaluation. */
int wf_test_sincos(const float* x, int n, float* s, float* c, int device);
/* sin/cos of 2*pi*u for every u = k * 2^-24, k in [k0, k1): writes FNV-1a 64 of the (s, c) bit
 * stream, so host and device can be compared over the whole 2^24 grid without moving 128 MiB. */
int wf_test_sincos_grid_hash(uint32_t k0, uint32_t k1, int device, uint64_t* hash);
/* Closest (anyhit = 0) or any hit for n rays (8 floats: o xyz, tmin, d xyz, tmax); returns the
 * hit distance and the primitive id in the reference's id space (-1 for a miss). device: 0 host,
 * 1 one GPU thread per ray, 2 the engines' wave-scheduled traversal with the tree top (up to 1024
 * nodes, and the primitive records when the whole tree fits) in LDS, 3 the same with only 8 nodes
 * in LDS (the rest read from memory). */
int wf_test_intersect(const wf_scene* scene, const float* rays, int n, int anyhit, float* t, int* prim, int device);
/* Slab tests of n cases (14 floats: box lo xyz, hi xyz, ray o xyz, d xyz, tmin, tmax), 1/d taken
 * as the traversal does: out[2i] = the reference-order slab test (include/wf/bvh.hpp:64-78),
 * out[2i+1] = the per-ray-signed form the GPU traversal uses, or -1 where that form does not apply
 * (an infinite or zero 1/d component, a non-finite origin, a NaN tmin or tmax). */
int wf_test_slab(const float* cases, int n, int* out, int device);
/* Quotients n[i] / d by the multiply-high division by an invariant divisor that the persistent
 * engine uses for new-path indexing (no reference counterpart; checked against plain division). */
int wf_test_divu32(const uint32_t* n, int count, uint32_t d, uint32_t* q, int device);
/* The persistent engine's LDS plan for n (<= 8) scenes of n_nodes[k] nodes and n_prims[k] primitives
 * within budget_bytes: nodes and primitive records each scene keeps in LDS (no reference
 * counterpart; host only). */
int wf_test_lds_plan(const int* n_nodes, const int* n_prims, int n, int64_t budget_bytes, int* node_cnt,
                     int* prim_cnt);
/* Radiance of single paths (px, py, sample triples) with the shared estimator, and per-path ray
 * counts (extension, shadow). */
int wf_test_trace_paths(const wf_scene* scene, const int* pxys, int n, uint32_t seed, int max_depth, float* L,
                        uint32_t* rays, int device);
/* One classify + shade step (include/wf/path_core.hpp:151-276) per case of 20 words (floats and
 * ints as their bits): ray o xyz, d xyz, hit t, hit primitive (device order; -1 a miss; -2: the
 * closest hit is found on the host), throughput xyz, L xyz, prev_pdf, depth, specular, key k0, k1,
 * max_depth. form 0: the per-material shading functions on the host, 1: the one-pass form the
 * persistent engine runs (no reference counterpart), 2 and 3: the same two on the GPU. out: 34 words
 * per case: alive, has_shadow, material type (-1 ended), ray o xyz, d xyz, tmin, tmax, throughput
 * xyz, L xyz, prev_pdf, depth, specular, shadow ray o xyz, d xyz, tmin, tmax, its contribution xyz,
 * hit primitive, hit t, 0. */
int wf_test_shade(const wf_scene* scene, const uint32_t* cases, int n, int form, uint32_t* out);

/* ---- memory (no reference counterpart) ---- */
/* Persistent engines (renderers, batches, the serving loop) hand their large device blocks (path
 * state and sample arenas, 64 MiB and up) to a process-wide cache when they end, and the next
 * engine that needs a block of the same size takes it from there: allocating and freeing a 12 GB
 * work-sized pool for every batch render stalled the device for up to 2 s now and then. At most
 * 32 GiB is kept. wf_cached_memory_bytes: what the cache holds; wf_release_cached_memory frees it
 * all (no engine may be running; freed_bytes may be NULL). */
uint64_t wf_cached_memory_bytes(void);
int wf_release_cached_memory(uint64_t* freed_bytes);

/* ---- measurement (no reference counterpart) ---- */
/* Runs a plain 16-byte-per-lane device copy of `bytes` (multiple of 16) reps times and returns the
 * average duration per copy. It reads and writes exactly `bytes`, so hardware traffic counters
 * (FETCH_SIZE, WRITE_SIZE) read on it calibrate the counters read on the engine's kernels. */
int wf_copy_probe(uint64_t bytes, int reps, double* ms_per_copy);

/* ---- batch: many jobs in one persistent launch (no reference counterpart) ---- */
typedef struct wf_job {
    const wf_scene* scene;
    int spp;
    uint32_t seed;
    int max_depth;
    float* rgb;       /* 3*W*H mean, may be NULL */
    float* lum_var;   /* W*H, may be NULL */
    /* The job's view of the scene (many jobs can share one scene and its BVH):
     * width, height: image size, 0 = the scene's; orbit = 1: the camera orbited to yaw_deg
     * (degrees) about its look-at point (as wf_scene_orbit), 0: the scene's camera. */
    int width, height;
    float yaw_deg;
    int orbit;
} wf_job;

typedef struct wf_batch_stats {
    double render_ms;
    uint64_t extension_rays, shadow_rays;
    int64_t invalid_hits, queue_overflows;
    int64_t iterations, graph_launches;
    int block_size, pool_paths, lds_kib, refill, leaf_batch, dir_bins, shade_block;
    uint64_t device_bytes;
    /* With tuning->stage_timing: per-stage time, items, launches and algorithmic device bytes in
     * wf_render_stats' order (extend, -, shade, -, -, shadow, -, other); otherwise zero. */
    double stage_ms[8];
    int64_t stage_items[8];
    int64_t stage_launches[8];
    int64_t stage_bytes[8];
    int64_t stage_threads[8];
} wf_batch_stats;

int wf_batch_render(const wf_job* jobs, int n_jobs, const wf_render_params* tuning, wf_batch_stats* stats);
/* Autotuner: a coordinate search over short renders of the batch (probe_spp samples per job):
 * block size, refill threshold x leaf batch, shade block, survivor grouping, LDS budget. The three
 * fastest candidates and the default tuning are then timed on the whole batch; the fastest of them
 * is timed again at smaller queue capacities (2^24 and 2^22 path slots), and the fastest of all is
 * written into *tuning (pool_paths 0: sized to the work). Renders into the jobs' films. */
int wf_autotune(const wf_job* jobs, int n_jobs, int probe_spp, wf_render_params* tuning);
/* The same search with the path pool's device memory limited to state_budget_bytes (184 bytes a
 * slot; 0 = no limit): every candidate's queue capacity fits it, so pool_paths in *tuning does.
 * WF_ERR_INVALID_ARGUMENT when the budget holds fewer than 1024 slots. No reference counterpart
 * (the reference has no autotuner). */
int wf_autotune_budget(const wf_job* jobs, int n_jobs, int probe_spp, uint64_t state_budget_bytes,
                       wf_render_params* tuning);

/* ---- serving loop: continuous batching under a device-memory budget (no reference counterpart) ----
 * A request renders in consecutive sample ranges of about one path pool each, resolved into its
 * film in sample order (the same bits as one render of all its samples): the device memory it
 * holds is its film (20 B per pixel) plus at most two ranges' sample buffers (16 B per pixel and
 * sample of the range), whatever its spp. The server keeps requests' scenes uploaded (keyed by the
 * scene handle, not its address) while the budget allows. A scene handle must outlive every
 * request that names it (until wf_server_poll reports it done or failed). */
typedef struct wf_server_stats {
    int64_t submitted, admitted, completed, rejected;
    int64_t in_flight;
    uint64_t extension_rays, shadow_rays;
    uint64_t bytes_in_use, budget_bytes;
    double busy_ms;
} wf_server_stats;

int wf_server_create(uint64_t budget_bytes, const wf_render_params* tuning, wf_server** out);
void wf_server_destroy(wf_server* s);
/* Queues a request; returns its id in *id. WF_ERR_BUDGET if it can never fit the budget (its film,
 * one sample of sample buffer and its scene beside the path pool). */
int wf_server_submit(wf_server* s, const wf_job* job, int64_t* id);
/* Admits queued requests that fit, queues `launches` graph replays, and then finishes the requests
 * that the previous step's replays completed (resolves and downloads their results on a second
 * stream while the replays just queued run). A request is therefore reported done by the step
 * after the one whose replays completed it; step(s, 0) only collects. */
int wf_server_step(wf_server* s, int launches);
/* 1 if request `id` is done (its buffers are filled), 0 if not yet, -2 if it failed (a device error
 * while admitting it, which that wf_server_step call returned; what it held is released), -1 if
 * the id is unknown. */
int wf_server_poll(wf_server* s, int64_t id);
/* Ray counts are those of the last collected step while requests run, and exact when none do. */
int wf_server_stats_get(wf_server* s, wf_server_stats* out);

#ifdef __cplusplus
}
#endif

#endif /* WFAMD_H */
