// CPU arm (the reference's render_cpu, src/cpu_renderer.cpp:13-137).
#pragma once
#include <cstdint>

#include "wf_film.h"
#include "wf_scene.h"

namespace wf {

struct CpuCounts {
    uint64_t extension_rays = 0, shadow_rays = 0, invalid_hits = 0;
};

// Adds samples [first_sample, first_sample + count) of every pixel to the film sums.
void cpu_render_passes(const Scene& scene, int first_sample, int count, uint32_t seed, int max_depth, int threads,
                       FilmSums* film, CpuCounts* counts);

// Full render: sums over spp samples, then the shared finalize. render_ms is the wall clock
// around the path tracing only (src/cpu_renderer.cpp:51,96).
RenderResult render_cpu(const Scene& scene, int spp, uint32_t seed, int max_depth, int threads, CpuCounts* counts,
                        double* render_ms);

}  // namespace wf
