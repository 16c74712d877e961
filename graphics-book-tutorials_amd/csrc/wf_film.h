// Film sums, the one finalize shared by every arm, and image comparison.
//   film_add      src/gpu/engine.cuh:25-32, src/cpu_renderer.cpp:62-70
//   finalize      src/cpu_renderer.cpp:106-120 (mean by division) with the variance in binary64
//   compare       src/film.cpp:45-113
#pragma once
#include <cstdint>
#include <string>
#include <vector>

#include "wf_core.h"

namespace wf {

// Per-pixel float sums over samples, accumulated in sample order (no atomics, one path per
// pixel per pass), so every arm that adds the same per-path L in the same order holds the
// same bits.
struct FilmSums {
    int width = 0, height = 0;
    std::vector<float> rgb, lum, lum2;
    void reset(int w, int h) {
        width = w;
        height = h;
        rgb.assign((size_t)3 * w * h, 0.0f);
        lum.assign((size_t)w * h, 0.0f);
        lum2.assign((size_t)w * h, 0.0f);
    }
};

WF_HD void film_add3(float* rgb, float* lum, float* lum2, Vec3 L) {
    rgb[0] += L.x;
    rgb[1] += L.y;
    rgb[2] += L.z;
    float l = luminance(L);
    *lum += l;
    *lum2 += l * l;
}

struct RenderResult {
    int width = 0, height = 0, spp = 0;
    std::vector<float> rgb;       // 3*w*h mean radiance
    std::vector<float> lum_var;   // w*h unbiased variance of per-sample luminance
};

// mean = sum / spp (binary32 division); lum_var = max(0, (S2 - S1^2/n) / (n - 1)) in binary64,
// 0 for n = 1. Either output pointer may be null.
void finalize_film(const float* rgb_sum, const float* lum_sum, const float* lum2_sum, int pixels, int spp,
                   float* rgb_out, float* lum_var_out);

struct ImageDiff {
    double max_abs_rel = 0.0;
    double mean_rel = 0.0;
    double frac_within_3se = 0.0;
    int nonfinite = 0;
};
ImageDiff compare_results(const RenderResult& a, const RenderResult& b);
int count_nonfinite(const std::vector<float>& rgb);   // pixels with any non-finite channel

}  // namespace wf
