// Persistent batched wavefront. A pool of path slots is kept full from a global work counter over
// all admitted jobs (work item = one (job, sample, pixel) path, sample-major inside a job, pixels
// in 8x8 tiles). The live paths of an iteration are the items [0, n_act): the survivors of the
// previous iteration, stored densely in item order, then the new paths, which are generated from
// their work index wherever they are needed instead of being stored. Per iteration:
//   plan    1 thread: fold the previous iteration's counts, size this one
//   extend  closest-hit traversal of every item (hit -> hit[item])
//   shade   classify + shade every item (all materials in one pass); a survivor's whole state is
//           written compacted, in item order, to the other buffer; a finished path writes its
//           radiance to its job's sample buffer; Lambert bounces append a shadow ray
//   shadow  any-hit traversal of the shadow rays; an unoccluded one adds its contribution to the
//           survivor's radiance (or to the sample of a path that ended with it)
// Survivor and shadow-ray slots are reserved once per workgroup round (the waves' ballot counts are
// combined in LDS, one atomic per counter per round), so every state read and write of extend and
// shade is a contiguous run. Traversal kernels run "while-while" with dynamic fetch: a workgroup
// owns every B-th slice of 128 rays and its waves take those slices in turn from an LDS counter; a
// lane whose ray has finished takes the next ray of its wave's slice as soon as fewer than `refill`
// lanes of the wave are still traversing, so neither a wave nor a workgroup idles on its longest
// ray. Inside the loop a lane that reaches a hit leaf parks it, and the wave tests parked
// leaves together once `leaf_batch` lanes have one; each ray still visits the nodes and tests the
// primitives of the shared traversal in the same order, so hits are the megakernel's bits.
// Every stage kernel reads its item count from the device control block, so an iteration needs
// no host round trip; K iterations are captured once into a hipGraph and replayed until the control
// block reports that all admitted work is done.
//
// A finished path writes its radiance to its job's sample buffer at (sample, pixel); when a job's
// paths have all finished, a resolve kernel adds the samples to the film in sample order, so the
// film holds exactly the bits that per-sample passes (megakernel, plain wavefront) produce.
#pragma once
#include <cstdint>
#include <memory>
#include <vector>

#include "wf_film.h"
#include "wf_gpu.h"
#include "wf_pathpool.h"
#include "wf_queues.h"
#include "wf_renderer.h"

namespace wf::gpu {

constexpr int kMaxScenes = 8;
// Job slots form a ring: a job is admitted with the next sequence number and lives in slot
// seq & kJobMask until it is finished and every older job has been finished too.
constexpr int kMaxJobs = 4096;
constexpr uint32_t kJobMask = kMaxJobs - 1;
// Deepest path the survivor state can carry (its depth field has 13 bits).
constexpr int kMaxDepth = 8191;

struct PersistentJob {
    const SceneView* scene = nullptr;   // view with device pointers (one per distinct scene)
    // The job's own view of the scene: image size and camera rig (has_view = false: the scene's).
    bool has_view = false;
    int width = 0, height = 0;
    CameraRig cam{};
    DeviceFilm film{};
    int first_sample = 0;
    int spp = 1;
    uint32_t seed = 1234;
    int max_depth = 8;
};

struct PersistentTuning {
    int block_size = 0;    // threads per workgroup: 256, 512 or 1024 (0 -> 1024 for one job, 512 for several)
    int pool_paths = 0;    // path slots = queue capacity (0 -> the render's work, up to 2^26)
    int lds_kib = 0;       // LDS per workgroup for BVH nodes and primitive records (0 -> 64 KiB for one job,
                           // 32 KiB for several; at most the device limit less the kernels' own LDS)
    int graph_iters = 0;   // iterations per captured graph (0 -> max_depth + 2 when every sample fits
                           // the pool, else 16; rounded up to even)
    int refill = 0;        // traversal refill threshold in lanes of 64 (0 -> 32)
    int leaf_batch = 0;    // parked leaves that trigger a wave's leaf-test round (0 -> 16)
    int dir_bins = 0;      // survivor grouping per block round: 1 none, 8 direction octants (0 -> 1)
    int shade_block = 0;   // threads per shade workgroup (= compaction round): 128..1024 (0 -> 512)
};

struct PersistentResult {
    int iterations = 0;
    int graph_launches = 0;
    int chunks = 0;
};

// Device job record.
struct JobDev {
    int scene;
    int width, height, max_depth;
    CameraRig cam;
    uint32_t first_sample, spp, key_a, key_b;   // path_key_a(seed), path_key_b(seed)
    DivU32 div_pixels, div_band;   // by width * height and by 8 * width (new-path indexing)
    unsigned long long work_begin, work_end;
    float4* sbuf;
    DeviceFilm film;
};

// Survivors and shadow rays are compacted per segment: shade round r (a workgroup's items) uses the
// counters of segment r mod kSegs, so concurrent rounds spread their atomics over kSegs addresses
// instead of serialising on one. Segment s compacts into its own window of the output (its size is
// the number of items of the rounds that map to s), and the next stage finds item i in segment s
// with seg_prefix[s] <= i < seg_prefix[s + 1] at position seg_base[s] + i - seg_prefix[s].
constexpr int kSegs = 16;

// Device control block. Jobs are numbered by admission (sequence numbers, modulo 2^32): job_hi is
// the next number, job_lo the oldest job whose paths are not all issued yet, so the new paths of
// an iteration belong to the jobs [job_lo, job_hi), whose work ranges follow each other.
struct PState {
    unsigned long long next_work, total_work, it_work_base, pad64;
    uint32_t n_live, have_prev, iterations, job_lo;
    uint32_t it_live, it_new, it_act, done;
    uint32_t job_hi, unocc, pad2, pad3;   // unocc: unoccluded shadow rays of this iteration (set by shadow)
    uint32_t n_next, n_shadow, shade_block, pad;   // n_next, n_shadow: totals of this iteration (set by shadow)
    uint32_t seg_cnt[kSegs];          // survivors per segment (this iteration)
    uint32_t sh_cnt[kSegs];           // shadow rays per segment (this iteration)
    uint32_t in_prefix[kSegs + 1];    // where this iteration's survivor items are (set by plan)
    uint32_t in_base[kSegs];
};

// One buffer of dense per-item path state (two of them alternate by iteration), 64 bytes a path:
//   ro : origin xyz, prev_pdf      rd : direction xyz, depth | specular << 13 | job slot << 14 |
//                                       scene slot << 26
//   T  : throughput xyz, key k0    L  : radiance xyz, sample-buffer index
// k1 of the path's key is rebuilt from the job's seed hash, the pixel and the sample, which the
// sample-buffer index gives (path_key_k1). The scene slot rides along so that a survivor's
// traversal needs no job lookup.
constexpr uint32_t kSpecShift = 13, kJobShift = 14, kSceneShift = 26;
static_assert(kMaxDepth < (1 << kSpecShift), "depth field");
static_assert(kMaxJobs == (1 << (kSceneShift - kJobShift)), "the job slot has 12 bits in a survivor's rd.w");
static_assert(kMaxScenes <= 8, "the scene slot has 3 bits in a survivor's rd.w");
__host__ __device__ inline uint32_t pack_rdw(uint32_t depth, uint32_t specular, uint32_t slot, uint32_t scene) {
    return depth | (specular << kSpecShift) | (slot << kJobShift) | (scene << kSceneShift);
}
// A shadow ray's tag: job slot | scene slot << 12 | final flag (bit 31: the path ended with it).
constexpr uint32_t kTagSceneShift = 12;
struct PathBuf {
    float4* ro = nullptr;
    float4* rd = nullptr;
    float4* T = nullptr;
    float4* L = nullptr;
};

// LDS layout of a workgroup: nodes of each scene, then primitive records of each scene.
struct LdsPlan {
    int n;
    int node_off[kMaxScenes];
    int node_cnt[kMaxScenes];
    int prim_off[kMaxScenes];
    int prim_cnt[kMaxScenes];   // 0 or the scene's primitive count
    int total_nodes, total_prims;
};

// Host mirror of the control block fields the serving loop polls.
struct PersistentProgress {
    unsigned long long next_work = 0, total_work = 0;
    uint32_t n_live = 0, done = 0, iterations = 0;
};

class PersistentEngine {
public:
    PersistentEngine();
    ~PersistentEngine();
    PersistentEngine(const PersistentEngine&) = delete;
    PersistentEngine& operator=(const PersistentEngine&) = delete;

    // Renders every job to completion, adding samples [first_sample, first_sample + spp) to each
    // job's film. Jobs whose sample buffers would exceed the buffer budget are split into
    // consecutive sample chunks. With `stages`, the iterations run eagerly with HIP events around
    // each stage (profiling) instead of graph replays.
    void render(const PersistentJob* jobs, int n_jobs, const PersistentTuning& t, DevCounters* ctr, hipStream_t stream,
                PersistentResult* out, StageTimes* stages = nullptr);

    // Session API (continuous batching). open() sizes the pool; admit() appends a job (its
    // sample buffer is allocated here) and returns its sequence number (the job's id); run()
    // replays the captured iteration graph `launches` times; progress() reads the control block
    // and the per-job finished-path counts; finish_job() resolves a complete job into its film and
    // frees its sample buffer (on `stream` when given, else the engine's); a finished job's slot
    // is reused once every older job is finished too. can_admit() tells whether a job on `scene`
    // would find a job slot and a scene slot (a scene slot frees when no unfinished job uses it);
    // forget_scene() drops a scene no unfinished job uses from the scene table (before its device
    // memory is freed); reset_jobs() clears the job table once nothing is in flight.
    // For a caller that keeps a launch queued while it works on the host: snapshot() enqueues,
    // after everything enqueued so far, copies of the control block, the per-job finished counts
    // and *ctr (optional) into one of two pinned slots, records an event and returns the slot;
    // collect() waits for that event only (not for later launches) and applies the finished
    // counts of the jobs that existed at the snapshot. At most two snapshots are outstanding.
    void open(const PersistentTuning& t, DevCounters* ctr, hipStream_t stream);
    int64_t admit(const PersistentJob& job);
    bool can_admit(const SceneView* scene) const;
    void forget_scene(const SceneView* scene);
    int64_t live_jobs() const;
    // Stream-ordered allocations of sample buffers come from `pool` (null: the device's default).
    void set_mem_pool(hipMemPool_t pool);
    void run(int launches);
    void run_profiled(StageTimes* out);
    PersistentProgress progress();
    int snapshot(const DevCounters* ctr);
    PersistentProgress collect(int slot, DevCounters* ctr_out);
    hipEvent_t snapshot_event(int slot) const;
    bool job_complete(int64_t job) const;
    void finish_job(int64_t job, hipStream_t stream = nullptr);
    void reset_jobs();
    uint64_t pool_bytes() const { return pool_bytes_; }
    // Device bytes of one path slot: two 64-byte survivor states, the hit and the shadow-queue entry.
    static constexpr uint64_t kBytesPerSlot = 2 * 64 + 8 + 48;
    uint64_t bytes() const;
    int pool_paths() const { return cap_; }

    struct Impl;

private:
    struct LaunchPlan;
    LaunchPlan make_plan(bool single);
    void launch_iteration(const LaunchPlan& L, int par, hipEvent_t* ev);
    void ensure_pool(const PersistentTuning& t);
    void capture(bool single);
    void update_lds_plan();
    JobDev prepare_job(const PersistentJob& job, float4* sbuf, int64_t* seq);
    int scene_slot(const SceneView* scene) const;
    void retire_jobs();
    void admit_chunk(const std::vector<PersistentJob>& chunk);
    uint64_t sbuf_budget_ = 4ull << 30;
    int cap_ = 0;
    uint64_t pool_bytes_ = 0;
    std::unique_ptr<Impl> m_;
};

// LDS layout for scenes of n_nodes[k] nodes and n_prims[k] primitives within `budget` bytes (the
// plan of a batch's workgroups).
LdsPlan plan_lds(const int* n_nodes, const int* n_prims, int n, long long budget);

// Picks block size, traversal scheduling, shade block, survivor grouping, LDS budget and queue
// capacity (path slots) for a workload by timing renders of it. state_budget: the most device bytes
// of path state the pool may take (kBytesPerSlot a slot; 0 = no limit).
PersistentTuning autotune_persistent(const PersistentJob* jobs, int n_jobs, hipStream_t stream, int probe_spp,
                                     std::vector<std::pair<PersistentTuning, double>>* trials,
                                     uint64_t state_budget = 0);

// Large device blocks of ended engines, kept for the next engine of the same block size: the bytes
// kept now, and freeing them all (returns the bytes freed).
uint64_t cached_block_bytes();
uint64_t release_cached_blocks();

}  // namespace wf::gpu
