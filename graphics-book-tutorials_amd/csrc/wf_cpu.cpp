// CPU arm (src/cpu_renderer.cpp:13-137): every pixel traces its samples in order with the shared
// estimator and adds them to float sums. Rows are handed out dynamically (an atomic row counter)
// instead of the reference's static bands; the image does not depend on the thread count because
// every pixel is owned by exactly one thread.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <stdexcept>
#include <thread>
#include <vector>

#include "wf_cpu.h"

namespace wf {

void finalize_film(const float* rgb_sum, const float* lum_sum, const float* lum2_sum, int pixels, int spp,
                   float* rgb_out, float* lum_var_out) {
    for (int i = 0; i < pixels; ++i) {
        if (rgb_out) {
            rgb_out[3 * i + 0] = rgb_sum[3 * i + 0] / (float)spp;
            rgb_out[3 * i + 1] = rgb_sum[3 * i + 1] / (float)spp;
            rgb_out[3 * i + 2] = rgb_sum[3 * i + 2] / (float)spp;
        }
        if (lum_var_out) {
            if (spp > 1) {
                double s1 = lum_sum[i], s2 = lum2_sum[i], n = spp;
                double v = (s2 - s1 * s1 / n) / (n - 1.0);
                lum_var_out[i] = (float)std::max(0.0, v);
            } else {
                lum_var_out[i] = 0.0f;
            }
        }
    }
}

int count_nonfinite(const std::vector<float>& rgb) {
    int c = 0;
    for (size_t i = 0; i + 2 < rgb.size(); i += 3)
        if (!std::isfinite(rgb[i]) || !std::isfinite(rgb[i + 1]) || !std::isfinite(rgb[i + 2])) ++c;
    return c;
}

ImageDiff compare_results(const RenderResult& a, const RenderResult& b) {
    if (a.width != b.width || a.height != b.height) throw std::invalid_argument("compare_results: image dimensions mismatch");
    ImageDiff d;
    const int n = a.width * a.height;
    if (n == 0) return d;
    for (size_t i = 0; i < a.rgb.size(); ++i)
        if (!std::isfinite(a.rgb[i]) || !std::isfinite(b.rgb[i])) ++d.nonfinite;
    double ma = 0.0, mb = 0.0;
    int within = 0;
    for (int i = 0; i < n; ++i) {
        float la = luminance(Vec3{a.rgb[3 * i], a.rgb[3 * i + 1], a.rgb[3 * i + 2]});
        float lb = luminance(Vec3{b.rgb[3 * i], b.rgb[3 * i + 1], b.rgb[3 * i + 2]});
        ma += la;
        mb += lb;
        for (int c = 0; c < 3; ++c) {
            float va = a.rgb[3 * i + c], vb = b.rgb[3 * i + c];
            if (std::isfinite(va) && std::isfinite(vb)) {
                float m = std::max(std::max(std::fabs(va), std::fabs(vb)), 1e-6f);
                d.max_abs_rel = std::max(d.max_abs_rel, (double)(std::fabs(va - vb) / m));
            }
        }
        if (std::isfinite(la) && std::isfinite(lb)) {
            float sea = std::sqrt(a.lum_var[i] / a.spp), seb = std::sqrt(b.lum_var[i] / b.spp);
            if (std::fabs(la - lb) <= 3.0f * std::sqrt(sea * sea + seb * seb)) ++within;
        }
    }
    ma /= n;
    mb /= n;
    d.mean_rel = ma > 1e-12 ? std::fabs(ma - mb) / ma : std::fabs(ma - mb);
    d.frac_within_3se = (double)within / n;
    return d;
}

void cpu_render_passes(const Scene& scene, int first_sample, int count, uint32_t seed, int max_depth, int threads,
                       FilmSums* film, CpuCounts* counts) {
    if (!scene.finalized()) throw std::runtime_error("Scene not finalized (no BVH nodes)");
    if (film->width != scene.width || film->height != scene.height) film->reset(scene.width, scene.height);
    const SceneView v = host_view(scene);
    const int W = scene.width, H = scene.height;
    if (threads <= 0) threads = (int)std::max(1u, std::thread::hardware_concurrency());
    threads = std::min(threads, H);
    std::atomic<int> next_row{0};
    std::vector<CpuCounts> per(threads);
    auto work = [&](int t) {
        RayCount rc{0, 0, 0};
        uint64_t ext = 0, sh = 0, inv = 0;
        for (int py; (py = next_row.fetch_add(1)) < H;) {
            for (int px = 0; px < W; ++px) {
                const int p = py * W + px;
                for (int s = first_sample; s < first_sample + count; ++s) {
                    rc = RayCount{0, 0, 0};
                    Vec3 L = trace_path(v, px, py, (uint32_t)s, seed, max_depth, &rc);
                    film_add3(&film->rgb[3 * (size_t)p], &film->lum[p], &film->lum2[p], L);
                    ext += rc.extension;
                    sh += rc.shadow;
                    inv += rc.invalid;
                }
            }
        }
        per[t].extension_rays = ext;
        per[t].shadow_rays = sh;
        per[t].invalid_hits = inv;
    };
    std::vector<std::thread> pool;
    for (int t = 1; t < threads; ++t) pool.emplace_back(work, t);
    work(0);
    for (auto& th : pool) th.join();
    for (const CpuCounts& c : per) {
        counts->extension_rays += c.extension_rays;
        counts->shadow_rays += c.shadow_rays;
        counts->invalid_hits += c.invalid_hits;
    }
}

RenderResult render_cpu(const Scene& scene, int spp, uint32_t seed, int max_depth, int threads, CpuCounts* counts,
                        double* render_ms) {
    FilmSums film;
    film.reset(scene.width, scene.height);
    CpuCounts local;
    auto t0 = std::chrono::steady_clock::now();
    cpu_render_passes(scene, 0, spp, seed, max_depth, threads, &film, &local);
    auto t1 = std::chrono::steady_clock::now();
    if (render_ms) *render_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
    if (counts) *counts = local;
    RenderResult r;
    r.width = scene.width;
    r.height = scene.height;
    r.spp = spp;
    r.rgb.resize((size_t)3 * scene.width * scene.height);
    r.lum_var.resize((size_t)scene.width * scene.height);
    finalize_film(film.rgb.data(), film.lum.data(), film.lum2.data(), scene.width * scene.height, spp, r.rgb.data(),
                  r.lum_var.data());
    return r;
}

}  // namespace wf
