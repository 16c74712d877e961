// `wf render` (src/main.cpp:35-203): warm-up and timed renders with any arm, optional PNG/PFM
// output and a report.json in the reference's schema v1 plus an `mi355x` extension object.
// C++20 host code over the engine's C ABI (include/wfamd.h).
//
// Differences from the reference, on purpose: rays are counted in every render (one atomic per
// wave), so no separate counting run is needed; rays_per_second counts extension + shadow rays
// over the median render time (src/main.cpp:149-158); nonfinite_pixels counts pixels (the
// reference divides a pixel count by 3, src/main.cpp:131).
#include <cstdio>
#include <iostream>
#include <memory>

#include "wf_cli.h"
#include "wf_gif.h"
#include "wf_json.h"
#include "wfamd.h"

using namespace wfcli;

namespace {

std::string git_revision() {
    FILE* p = popen("git rev-parse HEAD 2>/dev/null", "r");
    if (!p) return "";
    char buf[64] = {0};
    size_t n = fread(buf, 1, sizeof(buf) - 1, p);
    pclose(p);
    std::string s(buf, n);
    while (!s.empty() && (s.back() == '\n' || s.back() == '\r')) s.pop_back();
    return s;
}

void check(int rc) {
    if (rc != WF_OK) throw std::runtime_error(wf_last_error());
}

int mode_id(const std::string& m) {
    if (m == "cpu") return WF_MODE_CPU;
    if (m == "megakernel") return WF_MODE_MEGAKERNEL;
    if (m == "wavefront") return WF_MODE_WAVEFRONT;
    return WF_MODE_PERSISTENT;
}

// `wf validate-report FILE`: checks a report.json against schema version 1 (src/report.cpp:348-475).
// Exit 0 and "valid", or exit 1 and the first problem; 2 when the file cannot be read.
int validate_report_command(int argc, const char* argv[]) {
    if (argc != 3) {
        std::cerr << "Usage: wf validate-report REPORT.json\n";
        return 2;
    }
    std::ifstream f(argv[2], std::ios::binary);
    if (!f) {
        std::cerr << "Error: cannot open " << argv[2] << "\n";
        return 2;
    }
    std::ostringstream text;
    text << f.rdbuf();
    const std::string problem = validate_report_json(text.str());
    if (problem.empty()) {
        std::printf("valid\n");
        return 0;
    }
    std::printf("invalid: %s\n", problem.c_str());
    return 1;
}

// `wf imgdiff A.pfm B.pfm [--tol T]` (tools/imgdiff.cpp): prints max_abs_rel and mean_rel; exit 0
// when max_abs_rel <= T (default 1e-5), 1 when above, 2 on an error (unreadable file, different
// sizes, a non-finite value without an equal counterpart).
int imgdiff_command(int argc, const char* argv[]) {
    if (argc < 4) {
        std::cerr << "Usage: wf imgdiff A.pfm B.pfm [--tol T]\n";
        return 2;
    }
    double tol = 1e-5;
    for (int i = 4; i < argc; ++i) {
        const std::string a = argv[i];
        if (a == "--tol" && i + 1 < argc) {
            char* end = nullptr;
            tol = std::strtod(argv[++i], &end);
            if (!end || *end != '\0' || !(tol >= 0.0)) {
                std::cerr << "Error: --tol: not a non-negative number\n";
                return 2;
            }
        } else {
            std::cerr << "Error: unknown option " << a << "\n";
            return 2;
        }
    }
    int w1 = 0, h1 = 0, w2 = 0, h2 = 0;
    const std::vector<float> a = read_pfm(argv[2], &w1, &h1);
    const std::vector<float> b = read_pfm(argv[3], &w2, &h2);
    if (w1 != w2 || h1 != h2) {
        std::cerr << "Error: image sizes differ (" << w1 << "x" << h1 << " and " << w2 << "x" << h2 << ")\n";
        return 2;
    }
    const ImageDiff d = image_diff(a, b);
    if (d.nonfinite_mismatch) {
        std::cerr << "Error: a non-finite value differs between the images\n";
        return 2;
    }
    std::printf("max_abs_rel=%.17g\nmean_rel=%.17g\n", d.max_abs_rel, d.mean_rel);
    return d.max_abs_rel <= tol ? 0 : 1;
}

// `wf gif OUT.gif [options]` (tools/wfgif.cpp; options in wf_gif.h): progressive-refinement or orbit
// frames of GPU renders as one looping GIF. Exit 0 when written, 2 on a usage or render error.
int gif_command(int argc, const char* argv[]) {
    GifConfig c;
    try {
        c = parse_gif_args(argc, argv);
    } catch (const ConfigError& e) {
        std::cerr << "Error: " << e.what() << "\n" << gif_usage();
        return 2;
    }
    const int W = c.width, H = c.height, scale = std::max(1, H / 126);
    const int mode = mode_id(c.mode);
    wf_render_params p{};
    p.seed = c.seed;
    p.max_depth = c.max_depth;
    p.lds_nodes = -1;
    GifWriter gif(c.out, W, H);
    std::vector<float> mean((size_t)3 * W * H);
    auto scene_upper = [](std::string s) {
        for (char& ch : s) ch = (char)std::toupper((unsigned char)ch);
        return s;
    };
    for (const std::string& name : c.scenes) {
        wf_scene* scene = nullptr;
        check(wf_scene_create(name.c_str(), W, H, &scene));
        std::unique_ptr<wf_scene, void (*)(wf_scene*)> scene_guard(scene, wf_scene_destroy);
        std::vector<std::vector<float>> frames;
        std::vector<std::string> captions;
        if (c.orbit == 0) {
            wf_renderer* r = nullptr;
            check(wf_renderer_create(scene, mode, &r));
            std::unique_ptr<wf_renderer, void (*)(wf_renderer*)> r_guard(r, wf_renderer_destroy);
            check(wf_renderer_reset(r));
            int done = 0;
            for (int target : c.steps) {
                wf_render_stats st{};
                p.spp = target - done;
                check(wf_renderer_run_passes(r, done, target - done, &p, &st));
                done = target;
                check(wf_renderer_download_mean(r, done, mean.data()));
                frames.push_back(mean);
                captions.push_back(scene_upper(name) + "  " + std::to_string(done) + " SPP");
            }
        } else {
            for (int k = 0; k < c.orbit; ++k) {
                const float yaw = 360.0f * (float)k / (float)c.orbit;
                wf_scene* view = nullptr;
                check(wf_scene_view(scene, W, H, yaw, 1, &view));
                std::unique_ptr<wf_scene, void (*)(wf_scene*)> view_guard(view, wf_scene_destroy);
                wf_renderer* r = nullptr;
                check(wf_renderer_create(view, mode, &r));
                std::unique_ptr<wf_renderer, void (*)(wf_renderer*)> r_guard(r, wf_renderer_destroy);
                wf_render_stats st{};
                p.spp = c.spp;
                check(wf_renderer_render(r, &p, &st, mean.data(), nullptr));
                frames.push_back(mean);
                captions.push_back(scene_upper(name) + "  " + std::to_string((int)std::lround(yaw)) + " DEG  " +
                                   std::to_string(c.spp) + " SPP");
            }
        }
        const float exposure = c.exposure > 0.0f ? c.exposure
                                                 : gif_auto_exposure(c.orbit == 0 ? frames.back() : frames.front());
        for (size_t k = 0; k < frames.size(); ++k) {
            std::vector<uint8_t> rgb8 = tonemap_srgb8(frames[k], exposure);
            const bool last = k + 1 == frames.size();
            if (last && !c.png_prefix.empty()) write_png(c.png_prefix + name + ".png", W, H, rgb8);
            gif_caption(rgb8, W, H, 3 * scale, 3 * scale, captions[k], scale);
            gif.frame(gif_quantize(rgb8, W, H), last && c.orbit == 0 ? c.hold : c.delay);
        }
    }
    const int n = gif.frames();
    gif.close();
    std::printf("wrote %s (%d frames, %dx%d)\n", c.out.c_str(), n, W, H);
    return 0;
}

}  // namespace

int main(int argc, const char* argv[]) {
    if (argc >= 2 && std::string(argv[1]) == "gif") {
        try {
            return gif_command(argc, argv);
        } catch (const std::exception& e) {
            std::cerr << "Error: " << e.what() << std::endl;
            return 2;
        }
    }
    if (argc >= 2 && std::string(argv[1]) == "validate-report") return validate_report_command(argc, argv);
    if (argc >= 2 && std::string(argv[1]) == "imgdiff") {
        try {
            return imgdiff_command(argc, argv);
        } catch (const std::exception& e) {
            std::cerr << "Error: " << e.what() << std::endl;
            return 2;
        }
    }
    try {
        Config c = parse_args(argc, argv);
        wf_scene* scene = nullptr;
        check(wf_scene_create(c.scene.c_str(), c.width, c.height, &scene));
        std::unique_ptr<wf_scene, void (*)(wf_scene*)> scene_guard(scene, wf_scene_destroy);
        wf_scene_info info{};
        check(wf_scene_info_get(scene, &info));
        wf_render_params p{};
        p.spp = c.spp;
        p.seed = c.seed;
        p.max_depth = c.max_depth;
        p.threads = c.threads;
        p.lds_nodes = -1;
        const int mode = mode_id(c.mode);
        const size_t px = (size_t)c.width * c.height;
        std::vector<float> rgb(3 * px), var(px);
        wf_render_stats st{};
        std::vector<double> runs;
        wf_renderer* r = nullptr;
        if (mode != WF_MODE_CPU) check(wf_renderer_create(scene, mode, &r));
        std::unique_ptr<wf_renderer, void (*)(wf_renderer*)> r_guard(r, wf_renderer_destroy);
        for (int k = 0; k < c.warmup + c.reps; ++k) {
            st = wf_render_stats{};
            if (mode == WF_MODE_CPU) check(wf_render_cpu(scene, &p, &st, rgb.data(), var.data()));
            else check(wf_renderer_render(r, &p, &st, rgb.data(), var.data()));
            if (k >= c.warmup) runs.push_back(st.render_ms);
        }
        if (!c.out_prefix.empty()) {
            write_png(c.out_prefix + ".png", c.width, c.height, tonemap_srgb8(rgb));
            write_pfm(c.out_prefix + ".pfm", c.width, c.height, rgb);
        }
        const uint64_t pfm_hash = fnv1a64(rgb.data(), rgb.size() * sizeof(float));
        long long nonfinite = 0;
        for (size_t i = 0; i < px; ++i)
            if (!std::isfinite(rgb[3 * i]) || !std::isfinite(rgb[3 * i + 1]) || !std::isfinite(rgb[3 * i + 2])) ++nonfinite;
        const double med = percentile(runs, 0.5);
        const double rays = (double)st.extension_rays + (double)st.shadow_rays;
        std::printf("%s %s %dx%d spp %d: median %.3f ms, %.1f Mrays/s (%llu ext + %llu shadow), hash %s\n",
                    c.scene.c_str(), c.mode.c_str(), c.width, c.height, c.spp, med, rays / (med * 1e3),
                    (unsigned long long)st.extension_rays, (unsigned long long)st.shadow_rays, hex64(pfm_hash).c_str());
        if (!c.report_path.empty()) {
            wf_device_facts dev{};
            std::string dname = "cpu", drv, rt, arch;
            uint64_t total = 0;
            if (mode != WF_MODE_CPU && wf_query_device(&dev) == WF_OK) {
                dname = dev.name;
                drv = std::to_string(dev.driver_version);
                rt = std::to_string(dev.runtime_version);
                arch = dev.arch;
                total = dev.total_memory_bytes;
            }
            std::string g = git_revision();
            std::ostringstream o;
            o << "{\n  \"schema_version\": 1,\n  \"git_revision\": " << (g.empty() ? "null" : jstr(g)) << ",\n";
            o << "  \"scene\": {\"name\": " << jstr(c.scene) << ", \"hash\": " << jstr(hex64(info.hash)) << "},\n";
            o << "  \"config\": {\"width\": " << c.width << ", \"height\": " << c.height << ", \"spp\": " << c.spp
              << ", \"seed\": " << c.seed << ", \"max_depth\": " << c.max_depth << ", \"mode\": " << jstr(c.mode)
              << "},\n";
            o << "  \"device\": {\"name\": " << jstr(dname) << ", \"driver_version\": " << jstr(drv)
              << ", \"cuda_runtime_version\": " << jstr(rt) << ", \"cuda_driver_version\": " << jstr(drv)
              << ", \"compute_capability\": " << jstr(arch) << ", \"total_memory_bytes\": " << total << "},\n";
            o << "  \"timings_ms\": {\"runs\": [";
            for (size_t i = 0; i < runs.size(); ++i) o << (i ? ", " : "") << jnum(runs[i]);
            o << "], \"min\": " << jnum(percentile(runs, 0.0)) << ", \"median\": " << jnum(med)
              << ", \"p10\": " << jnum(percentile(runs, 0.1)) << ", \"p90\": " << jnum(percentile(runs, 0.9))
              << ", \"max\": " << jnum(percentile(runs, 1.0)) << "},\n";
            const double sps = med > 0 ? (double)px * c.spp / (med / 1e3) : 0.0;
            o << "  \"throughput\": {\"samples_per_second\": " << jnum(sps) << ", \"rays_per_second\": "
              << (rays > 0 && med > 0 ? jnum(rays / (med / 1e3)) : "null") << "},\n";
            o << "  \"peak_vram_bytes\": " << st.peak_vram_bytes << ",\n";
            o << "  \"queue_high_water\": {\"lambert\": " << st.queue_high_water[0] << ", \"metal\": "
              << st.queue_high_water[1] << ", \"dielectric\": " << st.queue_high_water[2] << ", \"shadow\": "
              << st.queue_high_water[3] << ", \"active\": " << st.queue_high_water[4] << "},\n";
            std::string png = c.out_prefix.empty() ? "" : c.out_prefix + ".png";
            std::string pfm = c.out_prefix.empty() ? "" : c.out_prefix + ".pfm";
            o << "  \"outputs\": {\"png\": " << jstr(png) << ", \"pfm\": " << jstr(pfm) << ", \"pfm_hash\": "
              << jstr(hex64(pfm_hash)) << "},\n";
            o << "  \"validity\": {\"nonfinite_pixels\": " << nonfinite << ", \"queue_overflows\": "
              << st.queue_overflows << ", \"invalid_hits\": " << st.invalid_hits << "},\n";
            o << "  \"mi355x\": {\"extension_rays\": " << st.extension_rays << ", \"shadow_rays\": " << st.shadow_rays
              << ", \"mrays_per_second\": " << jnum(med > 0 ? rays / (med * 1e3) : 0.0)
              << ", \"extension_mrays_per_second\": " << jnum(med > 0 ? (double)st.extension_rays / (med * 1e3) : 0.0)
              << ", \"render_ms_p50\": " << jnum(med) << ", \"render_ms_p99\": " << jnum(percentile(runs, 0.99))
              << ", \"iterations\": " << st.iterations << ", \"arch\": " << jstr(arch) << "}\n}\n";
            std::ofstream f(c.report_path);
            if (!f) throw std::runtime_error("cannot open report file for writing: " + c.report_path);
            f << o.str();
            if (!f) throw std::runtime_error("failed to write report file");
        }
        return 0;
    } catch (const ConfigError& e) {
        std::cerr << "Error: " << e.what() << "\n" << usage();
        return 1;
    } catch (const std::exception& e) {
        std::cerr << "Error: " << e.what() << std::endl;
        return 1;
    }
}
