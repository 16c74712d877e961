// Megakernel (src/gpu/megakernel.cu:11-104): one thread per pixel traces complete paths. Each
// thread runs the requested samples in order and keeps its pixel's film sums in registers, so
// the film is read and written once per launch and the sums are added in sample order, exactly as
// successive one-sample passes would. A block covers a 16x16 pixel tile with one 8x8 tile per
// wave64, blocks are spread over the XCDs in contiguous ranges, and the breadth-first top of the
// BVH is staged in LDS. Ray counts are reduced per wave before one atomic per wave.
#include "wf_megakernel.h"

namespace wf::gpu {

__global__ __launch_bounds__(256) void megakernel_kernel(SceneView s, DeviceFilm film, DevCounters* ctr,
                                                         uint32_t first_sample, int count, uint32_t seed,
                                                         int max_depth, int n_lds, int tiles_x, int nblocks) {
    extern __shared__ Node lds_nodes[];
    stage_nodes(lds_nodes, s.nodes, n_lds);
    const int blk = xcd_remap(blockIdx.x, nblocks);
    const int tx = blk % tiles_x, ty = blk / tiles_x;
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    const int px = tx * 16 + (w & 1) * 8 + (l & 7);
    const int py = ty * 16 + (w >> 1) * 8 + (l >> 3);
    uint32_t ext = 0, sh = 0, inv = 0;
    if (px < s.width && py < s.height) {
        const int p = py * s.width + px;
        float rgb[3] = {film.rgb[3 * p], film.rgb[3 * p + 1], film.rgb[3 * p + 2]};
        float lum = film.lum[p], lum2 = film.lum2[p];
        const LdsNodes src = lds_nodes_of(s, lds_nodes, n_lds);
        for (int k = 0; k < count; ++k) {
            RayCount rc{0, 0, 0};
            Vec3 L = trace_path_with(s, src, px, py, first_sample + (uint32_t)k, seed, max_depth, &rc);
            film_add3(rgb, &lum, &lum2, L);
            ext += rc.extension;
            sh += rc.shadow;
            inv += rc.invalid;
        }
        film.rgb[3 * p] = rgb[0];
        film.rgb[3 * p + 1] = rgb[1];
        film.rgb[3 * p + 2] = rgb[2];
        film.lum[p] = lum;
        film.lum2[p] = lum2;
    }
    wave_add_u64(&ctr->extension_rays, ext);
    wave_add_u64(&ctr->shadow_rays, sh);
    wave_add_u32(&ctr->invalid_hits, inv);
}

void run_megakernel(const SceneView& s, const DeviceFilm& film, DevCounters* ctr, uint32_t first_sample, int count,
                    uint32_t seed, int max_depth, int lds_nodes, hipStream_t stream) {
    const int n_lds = lds_nodes < 0 ? (s.n_nodes < kMegakernelLdsNodes ? s.n_nodes : kMegakernelLdsNodes)
                                    : (lds_nodes < s.n_nodes ? lds_nodes : s.n_nodes);
    const int tiles_x = (s.width + 15) / 16, tiles_y = (s.height + 15) / 16;
    const int nblocks = tiles_x * tiles_y;
    megakernel_kernel<<<nblocks, 256, (size_t)n_lds * sizeof(Node), stream>>>(s, film, ctr, first_sample, count, seed,
                                                                              max_depth, n_lds, tiles_x, nblocks);
    check(hipGetLastError(), "megakernel launch");
}

}  // namespace wf::gpu
