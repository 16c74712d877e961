#pragma once
#include <vector>

#include "wf_film.h"
#include "wf_gpu.h"
#include "wf_pathpool.h"
#include "wf_queues.h"
#include "wf_renderer.h"

namespace wf::gpu {

constexpr int kWavefrontLdsNodes = 1024;
constexpr int kMaxTraceBounces = 80;

struct PassOptions {
    int queue_capacity = 0;
    int lds_nodes = -1;
    int scan_tile = 0;
    bool stage_timing = false;
};

// Plain wavefront: path pool of one path per pixel (src/gpu/engine.cuh:51-71).
class WavefrontEngine {
public:
    WavefrontEngine(const SceneView& scene, int width, int height);
    ~WavefrontEngine();
    WavefrontEngine(const WavefrontEngine&) = delete;
    WavefrontEngine& operator=(const WavefrontEngine&) = delete;

    // Renders sample `sample` of every pixel into the film; synchronises once per bounce.
    void run_pass(const DeviceFilm& film, DevCounters* ctr, uint32_t sample, uint32_t seed, int max_depth,
                  const PassOptions& opt, StageTimes* stages, std::vector<BounceTrace>* trace, int* bounces,
                  hipStream_t stream);
    uint64_t bytes() const { return bytes_; }

private:
    template <class T> void alloc(T** p, size_t bytes);
    void release();

    SceneView scene_;
    int pixels_;
    PathPool pool_;
    ShadowQueue sq_;
    uint32_t* alive_ = nullptr;
    uint32_t* queues_[3] = {nullptr, nullptr, nullptr};
    uint32_t* active_[2] = {nullptr, nullptr};
    uint32_t* flags_ = nullptr;
    uint32_t* counters_ = nullptr;
    uint32_t* trace_dev_ = nullptr;
    ScanWorkspace ws_;
    hipEvent_t events_[2] = {nullptr, nullptr};
    // Persistent grid sizes of the traversal kernels for the LDS size they were computed for.
    size_t trace_lds_ = (size_t)-1;
    int extend_grid_ = 0, shadow_grid_ = 0, cus_ = 0;
    std::vector<void*> allocs_;
    uint64_t bytes_ = 0;
};

}  // namespace wf::gpu
