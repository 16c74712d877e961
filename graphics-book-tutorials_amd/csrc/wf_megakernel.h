#pragma once
#include "wf_film.h"
#include "wf_gpu.h"

namespace wf::gpu {

// Default number of top BVH nodes staged in LDS per megakernel workgroup (32 KiB).
constexpr int kMegakernelLdsNodes = 1024;

// Adds samples [first_sample, first_sample + count) of every pixel to the film (stream-ordered).
void run_megakernel(const SceneView& s, const DeviceFilm& film, DevCounters* ctr, uint32_t first_sample, int count,
                    uint32_t seed, int max_depth, int lds_nodes, hipStream_t stream);

}  // namespace wf::gpu
