// Persistent batched wavefront. A fixed pool of path slots is kept full: each iteration
// regenerates freed slots from a global work counter over all admitted jobs (work item = one
// (job, sample, pixel) path, sample-major inside a job), traces and classifies every live path,
// shades the three material queues, traces the shadow queue and retires finished paths with a
// stable two-way partition (survivors -> next active list, finished -> free-slot stack). Every
// stage kernel runs on a persistent grid and reads its item count from a device control block,
// so an iteration needs no host round trip; K iterations are captured once into a hipGraph and
// replayed until the control block reports that all work is done.
//
// A finished path writes its radiance to its job's sample buffer at (sample, pixel); when a job
// has no work left, a resolve kernel adds the samples to the film in sample order, so the film
// holds exactly the bits that per-sample passes (megakernel, plain wavefront) produce.
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

struct PersistentJob {
    const SceneView* scene = nullptr;   // view with device pointers (one per distinct scene)
    DeviceFilm film{};
    int first_sample = 0;
    int spp = 1;
    uint32_t seed = 1234;
    int max_depth = 8;
};

struct PersistentTuning {
    int block_size = 0;    // threads per workgroup (0 -> 256)
    int pool_paths = 0;    // path slots = queue capacity (0 -> 2^22)
    int scan_tile = 0;     // retire partition tile (0 -> 2048)
    int lds_nodes = -1;    // BVH nodes staged in LDS per workgroup (-1 -> 1024)
    int graph_iters = 0;   // iterations per captured graph (0 -> 16; rounded up to even)
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
    uint32_t first_sample, spp, seed, pad;
    unsigned long long work_begin, work_end;
    float4* sbuf;
    DeviceFilm film;
};

// Device control block.
struct PState {
    unsigned long long next_work, total_work, it_work_base, min_live_work;
    uint32_t n_live, n_free, n_jobs, have_prev;
    uint32_t it_live_old, it_new, it_act, it_free_after;
    uint32_t q[4];   // lambert, metal, dielectric, shadow
    uint32_t n_alive, iterations, done, overflow;
};

struct LdsPlan {
    int n;
    int off[kMaxScenes];
    int cnt[kMaxScenes];
};

// Host mirror of the control block fields the serving loop polls.
struct PersistentProgress {
    unsigned long long next_work = 0, total_work = 0, min_live_work = 0;
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
    // consecutive sample chunks.
    // With `stages`, the iterations run eagerly with HIP events around each stage (profiling).
    void render(const PersistentJob* jobs, int n_jobs, const PersistentTuning& t, DevCounters* ctr, hipStream_t stream,
                PersistentResult* out, StageTimes* stages = nullptr);

    // Session API (continuous batching). open() sizes the pool; admit() appends a job (its
    // sample buffer is allocated here) and returns its index; run() replays the captured
    // iteration graph `launches` times; progress() reads the control block; finish_job()
    // resolves a job whose work is complete (work_end <= min(next_work, min_live_work)) into
    // its film and frees its sample buffer. reset_jobs() clears the job table once all are done.
    void open(const PersistentTuning& t, DevCounters* ctr, hipStream_t stream);
    int admit(const PersistentJob& job);
    void run(int launches);
    PersistentProgress progress();
    bool job_complete(int job, const PersistentProgress& pr) const;
    void finish_job(int job);
    void reset_jobs();
    uint64_t job_buffer_bytes(const PersistentJob& job) const;
    uint64_t pool_bytes() const { return pool_bytes_; }
    uint64_t bytes() const;
    int pool_paths() const { return cap_; }

    struct Impl;
    void run_profiled(StageTimes* out);

private:
    struct LaunchPlan;
    LaunchPlan make_plan(bool single);
    void launch_iteration(const LaunchPlan& L, int par, hipEvent_t* ev);
    void ensure_pool(const PersistentTuning& t);
    void capture(bool single);
    uint64_t sbuf_budget_ = 4ull << 30;
    int cap_ = 0;
    uint64_t pool_bytes_ = 0;
    std::unique_ptr<Impl> m_;
};

// Picks block size, pool size and scan tile for a workload by timing short renders of it.
PersistentTuning autotune_persistent(const PersistentJob* jobs, int n_jobs, hipStream_t stream, int probe_spp,
                                     std::vector<std::pair<PersistentTuning, double>>* trials);

}  // namespace wf::gpu
