// GpuRenderer: owns the device copy of one scene, the film, the counters and (for the wavefront
// modes) the path pool. Same API and ownership as include/wf/renderer.hpp:44-67.
#pragma once
#include <cstdint>
#include <memory>
#include <string>
#include <vector>

#include "wf_film.h"
#include "wf_scene.h"

namespace wf {

struct RenderParams {
    int spp = 64;
    uint32_t seed = 1234;
    int max_depth = 8;
    int threads = 0;             // CPU arm only
    bool count_rays = false;     // kept for the reference API; rays are always counted (per wave)
    bool stage_timing = false;   // wavefront: per-stage events (synchronising)
    int queue_capacity = 0;      // wavefront test hook: 0 = path count; smaller -> counted overflows
    bool trace_bounces = false;  // wavefront: record per-bounce integer work (and list hashes)
    // Persistent engine tuning (0 = default / autotuned).
    int block_size = 0;
    int pool_paths = 0;
    int scan_tile = 0;           // plain wavefront compaction tile
    int lds_kib = 0;             // persistent: LDS per workgroup (KiB)
    int refill = 0;              // persistent: traversal refill threshold (lanes)
    int leaf_batch = 0;          // persistent: parked leaves per leaf-test round
    int dir_bins = 0;            // persistent: survivor grouping (1 or 8 direction octants)
    int shade_block = 0;         // persistent: threads per shade workgroup
    int graph_iters = 0;         // persistent: pool iterations per captured graph launch
    int lds_nodes = -1;          // nodes staged in LDS per workgroup (-1 = default)
};

struct StageTimes {   // milliseconds summed over a render
    double extend = 0, classify = 0, shade_lambert = 0, shade_metal = 0, shade_dielectric = 0, shadow = 0, compact = 0,
           other = 0;
    // Items processed and kernel launches per stage, same order as the fields above.
    long long items[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    long long launches[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    // Algorithmic device-memory bytes per stage (persistent engine; 0 where not computed).
    long long bytes[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    // Threads per launch of each stage (persistent engine: persistent grid x workgroup size).
    long long threads[8] = {0, 0, 0, 0, 0, 0, 0, 0};
};

struct RenderStats {
    double render_ms = 0.0;
    double total_ms = 0.0;
    uint64_t extension_rays = 0, shadow_rays = 0;
    long long queue_high_water[5] = {0, 0, 0, 0, 0};
    long long queue_overflows = 0;
    long long invalid_hits = 0;
    uint64_t peak_vram_bytes = 0;
    int iterations = 0;   // wavefront bounce steps (persistent: pool iterations)
    StageTimes stages;
};

// One bounce of the plain wavefront: counts and an FNV-1a hash of the compacted active list.
struct BounceTrace {
    uint32_t active, lambert, metal, dielectric, shadow, next_active;
    uint64_t list_hash;
};

enum class GpuMode { Megakernel = 1, Wavefront = 2, Persistent = 3 };

class GpuRenderer {
public:
    GpuRenderer(const Scene& scene, GpuMode mode);
    ~GpuRenderer();
    GpuRenderer(const GpuRenderer&) = delete;
    GpuRenderer& operator=(const GpuRenderer&) = delete;

    RenderResult render(const RenderParams& p, RenderStats* stats);
    void reset();
    void run_passes(int first_sample, int count, const RenderParams& p, RenderStats* stats);
    RenderResult download(int spp_done);
    // download() into caller-owned buffers (3 * pixels and pixels floats; either may be null).
    void download_to(int spp_done, float* rgb, float* lum_var);
    void download_mean(int spp_done, std::vector<float>* rgb);
    void download_sums(std::vector<float>* rgb, std::vector<float>* lum, std::vector<float>* lum2);
    uint64_t device_bytes() const;
    const std::string& device_name() const;
    const std::vector<BounceTrace>& trace() const;

    struct Impl;
private:
    std::unique_ptr<Impl> impl_;
};

struct DeviceFacts {
    std::string name, arch, driver_version, runtime_version;
    int compute_units = 0;
    uint64_t total_memory_bytes = 0, free_memory_bytes = 0;
};
DeviceFacts query_device();

}  // namespace wf
