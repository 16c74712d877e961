// GpuRenderer (src/gpu/gpu_renderer.cu:100-392): scene upload, film, counters, the progressive
// API and the download with the shared finalize. One HIP stream per renderer (default priority);
// every HIP failure becomes std::runtime_error. Peak device memory is the hipMemGetInfo delta
// across construction (the reference adds the allocations to that delta a second time,
// src/gpu/gpu_renderer.cu:233-237).
#include <algorithm>
#include <cstring>
#include <exception>
#include <thread>

#include "wf_megakernel.h"
#include "wf_persistent.h"
#include "wf_renderer.h"
#include "wf_wavefront.h"

namespace wf {

namespace gpu {

void DeviceScene::upload(const Scene& s) {
    release();
    view = host_view(s);
    auto put = [&](const void* src, size_t bytes) -> void* {
        if (bytes == 0) return nullptr;
        void* d = nullptr;
        check(hipMalloc(&d, bytes), "hipMalloc scene");
        allocs.push_back(d);
        check(hipMemcpy(d, src, bytes, hipMemcpyHostToDevice), "hipMemcpy scene");
        this->bytes += bytes;
        return d;
    };
    view.nodes = (const Node*)put(s.dnodes.data(), s.dnodes.size() * sizeof(Node));
    view.prims = (const Prim*)put(s.dprims.data(), s.dprims.size() * sizeof(Prim));
    view.materials = (const Material*)put(s.materials.data(), s.materials.size() * sizeof(Material));
    view.lights = (const Light*)put(s.lights.data(), s.lights.size() * sizeof(Light));
}

void DeviceScene::release() {
    for (void* p : allocs) (void)hipFree(p);
    allocs.clear();
    bytes = 0;
}

}  // namespace gpu

struct GpuRenderer::Impl {
    GpuMode mode;
    int width = 0, height = 0, pixels = 0;
    gpu::DeviceScene scene;
    gpu::DeviceFilm film{};
    // Pinned host copy of the film sums (rgb 3P, lum P, lum2 P floats), allocated at the first
    // download: the copies then run at DMA speed instead of through pageable staging.
    float* pinned = nullptr;
    hipEvent_t dl_ev[4] = {};   // one per download range (download_to)
    gpu::DevCounters* d_ctr = nullptr;
    gpu::DevCounters* h_ctr = nullptr;   // pinned host mirror of the counters
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    std::unique_ptr<gpu::WavefrontEngine> wf;
    std::unique_ptr<gpu::PersistentEngine> pe;
    uint64_t peak_bytes = 0, own_bytes = 0;
    std::vector<BounceTrace> trace;

    ~Impl() {
        wf.reset();
        pe.reset();
        if (film.rgb) (void)hipFree(film.rgb);
        if (film.lum) (void)hipFree(film.lum);
        if (film.lum2) (void)hipFree(film.lum2);
        if (d_ctr) (void)hipFree(d_ctr);
        if (h_ctr) (void)hipHostFree(h_ctr);
        if (pinned) (void)hipHostFree(pinned);
        for (hipEvent_t e : dl_ev)
            if (e) (void)hipEventDestroy(e);
        if (ev0) (void)hipEventDestroy(ev0);
        if (ev1) (void)hipEventDestroy(ev1);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

GpuRenderer::GpuRenderer(const Scene& scene, GpuMode mode) : impl_(std::make_unique<Impl>()) {
    if (!scene.finalized()) throw std::runtime_error("Scene not finalized (no BVH nodes)");
    Impl& m = *impl_;
    m.mode = mode;
    size_t free0 = 0, total = 0;
    gpu::check(hipMemGetInfo(&free0, &total), "hipMemGetInfo");
    gpu::check(hipStreamCreate(&m.stream), "hipStreamCreate");
    gpu::check(hipEventCreate(&m.ev0), "hipEventCreate");
    gpu::check(hipEventCreate(&m.ev1), "hipEventCreate");
    m.scene.upload(scene);
    m.width = scene.width;
    m.height = scene.height;
    m.pixels = scene.width * scene.height;
    const size_t P = (size_t)m.pixels;
    gpu::check(hipMalloc(&m.film.rgb, 3 * P * sizeof(float)), "hipMalloc film.rgb");
    gpu::check(hipMalloc(&m.film.lum, P * sizeof(float)), "hipMalloc film.lum");
    gpu::check(hipMalloc(&m.film.lum2, P * sizeof(float)), "hipMalloc film.lum2");
    m.film.pixels = m.pixels;
    gpu::check(hipMalloc(&m.d_ctr, sizeof(gpu::DevCounters)), "hipMalloc counters");
    gpu::check(hipHostMalloc((void**)&m.h_ctr, sizeof(gpu::DevCounters), hipHostMallocDefault), "hipHostMalloc counters");
    m.own_bytes = m.scene.bytes + 5 * P * sizeof(float) + sizeof(gpu::DevCounters);
    if (mode == GpuMode::Wavefront) {
        m.wf = std::make_unique<gpu::WavefrontEngine>(m.scene.view, m.width, m.height);
        m.own_bytes += m.wf->bytes();
    } else if (mode == GpuMode::Persistent) {
        m.pe = std::make_unique<gpu::PersistentEngine>();
    }
    reset();
    size_t free1 = 0;
    gpu::check(hipMemGetInfo(&free1, &total), "hipMemGetInfo");
    m.peak_bytes = free0 > free1 ? (uint64_t)(free0 - free1) : m.own_bytes;
}

GpuRenderer::~GpuRenderer() = default;

void GpuRenderer::reset() {
    Impl& m = *impl_;
    const size_t P = (size_t)m.pixels;
    gpu::check(hipMemsetAsync(m.film.rgb, 0, 3 * P * sizeof(float), m.stream), "hipMemset rgb");
    gpu::check(hipMemsetAsync(m.film.lum, 0, P * sizeof(float), m.stream), "hipMemset lum");
    gpu::check(hipMemsetAsync(m.film.lum2, 0, P * sizeof(float), m.stream), "hipMemset lum2");
    gpu::check(hipStreamSynchronize(m.stream), "reset sync");
}

void GpuRenderer::run_passes(int first_sample, int count, const RenderParams& p, RenderStats* stats) {
    Impl& m = *impl_;
    if (count < 0 || first_sample < 0) throw std::runtime_error("run_passes: negative sample range");
    if (p.max_depth < 0 || p.max_depth > 0xffff) throw std::runtime_error("run_passes: max_depth out of range");
    gpu::check(hipMemsetAsync(m.d_ctr, 0, sizeof(gpu::DevCounters), m.stream), "hipMemset counters");
    m.trace.clear();
    StageTimes st;
    int iterations = 0;
    gpu::check(hipEventRecord(m.ev0, m.stream), "hipEventRecord");
    if (count > 0) {
        if (m.mode == GpuMode::Megakernel) {
            gpu::run_megakernel(m.scene.view, m.film, m.d_ctr, (uint32_t)first_sample, count, p.seed, p.max_depth,
                                p.lds_nodes, m.stream);
        } else if (m.mode == GpuMode::Wavefront) {
            gpu::PassOptions opt;
            opt.queue_capacity = p.queue_capacity;
            opt.lds_nodes = p.lds_nodes;
            opt.scan_tile = p.scan_tile;
            opt.stage_timing = p.stage_timing;
            for (int k = 0; k < count; ++k)
                m.wf->run_pass(m.film, m.d_ctr, (uint32_t)(first_sample + k), p.seed, p.max_depth, opt,
                               p.stage_timing ? &st : nullptr, p.trace_bounces ? &m.trace : nullptr, &iterations,
                               m.stream);
        } else {
            gpu::PersistentJob job;
            job.scene = &m.scene.view;
            job.film = m.film;
            job.first_sample = first_sample;
            job.spp = count;
            job.seed = p.seed;
            job.max_depth = p.max_depth;
            gpu::PersistentTuning tune;
            tune.block_size = p.block_size;
            tune.pool_paths = p.pool_paths;
            tune.lds_kib = p.lds_kib;
            tune.refill = p.refill;
            tune.leaf_batch = p.leaf_batch;
            tune.dir_bins = p.dir_bins;
            tune.shade_block = p.shade_block;
            tune.graph_iters = p.graph_iters;
            gpu::PersistentResult pr;
            m.pe->render(&job, 1, tune, m.d_ctr, m.stream, &pr, p.stage_timing ? &st : nullptr);
            iterations = pr.iterations;
        }
    }
    gpu::check(hipEventRecord(m.ev1, m.stream), "hipEventRecord");
    gpu::check(hipEventSynchronize(m.ev1), "render sync");
    gpu::check(hipGetLastError(), "render");
    if (stats) {
        float ms = 0.0f;
        gpu::check(hipEventElapsedTime(&ms, m.ev0, m.ev1), "hipEventElapsedTime");
        gpu::check(hipMemcpyAsync(m.h_ctr, m.d_ctr, sizeof(gpu::DevCounters), hipMemcpyDeviceToHost, m.stream),
                   "copy counters");
        gpu::check(hipStreamSynchronize(m.stream), "counters sync");
        const gpu::DevCounters c = *m.h_ctr;
        stats->render_ms = ms;
        stats->extension_rays += c.extension_rays;
        stats->shadow_rays += c.shadow_rays;
        stats->invalid_hits += c.invalid_hits;
        stats->queue_overflows += c.overflows;
        for (int j = 0; j < 5; ++j) stats->queue_high_water[j] = std::max(stats->queue_high_water[j], (long long)c.high_water[j]);
        stats->peak_vram_bytes = std::max(m.peak_bytes, m.pe ? m.own_bytes + m.pe->bytes() : m.own_bytes);
        stats->iterations += iterations;
        stats->stages.extend += st.extend;
        stats->stages.classify += st.classify;
        stats->stages.shade_lambert += st.shade_lambert;
        stats->stages.shade_metal += st.shade_metal;
        stats->stages.shade_dielectric += st.shade_dielectric;
        stats->stages.shadow += st.shadow;
        stats->stages.compact += st.compact;
        stats->stages.other += st.other;
        for (int k = 0; k < 8; ++k) {
            stats->stages.items[k] += st.items[k];
            stats->stages.launches[k] += st.launches[k];
            stats->stages.bytes[k] += st.bytes[k];
            stats->stages.threads[k] = std::max(stats->stages.threads[k], st.threads[k]);
        }
    }
}

namespace {

// Copies the requested film sums into the renderer's pinned buffer (rgb at 0, lum at 3P, lum2 at
// 4P floats) on its stream and waits for them.
const float* stage_film(GpuRenderer::Impl& m, bool rgb, bool lum, bool lum2) {
    const size_t P = (size_t)m.pixels;
    if (!m.pinned) gpu::check(hipHostMalloc((void**)&m.pinned, 5 * P * sizeof(float), hipHostMallocDefault), "hipHostMalloc film");
    if (rgb)
        gpu::check(hipMemcpyAsync(m.pinned, m.film.rgb, 3 * P * sizeof(float), hipMemcpyDeviceToHost, m.stream),
                   "download rgb");
    if (lum)
        gpu::check(hipMemcpyAsync(m.pinned + 3 * P, m.film.lum, P * sizeof(float), hipMemcpyDeviceToHost, m.stream),
                   "download lum");
    if (lum2)
        gpu::check(hipMemcpyAsync(m.pinned + 4 * P, m.film.lum2, P * sizeof(float), hipMemcpyDeviceToHost, m.stream),
                   "download lum2");
    gpu::check(hipStreamSynchronize(m.stream), "download sync");
    return m.pinned;
}

}  // namespace

void GpuRenderer::download_sums(std::vector<float>* rgb, std::vector<float>* lum, std::vector<float>* lum2) {
    Impl& m = *impl_;
    const size_t P = (size_t)m.pixels;
    const float* h = stage_film(m, rgb != nullptr, lum != nullptr, lum2 != nullptr);
    if (rgb) rgb->assign(h, h + 3 * P);
    if (lum) lum->assign(h + 3 * P, h + 4 * P);
    if (lum2) lum2->assign(h + 4 * P, h + 5 * P);
}

RenderResult GpuRenderer::download(int spp_done) {
    RenderResult r;
    r.width = impl_->width;
    r.height = impl_->height;
    r.spp = spp_done;
    r.rgb.resize(3 * (size_t)impl_->pixels);
    r.lum_var.resize((size_t)impl_->pixels);
    download_to(spp_done, r.rgb.data(), r.lum_var.data());
    return r;
}

// The film sums are copied into the pinned buffer in up to four pixel ranges, with an event after
// each; each range is finalized into the caller's buffers as soon as its copy has landed, so the
// host finalize of one range overlaps the copies of the next.
void GpuRenderer::download_to(int spp_done, float* rgb, float* lum_var) {
    if (spp_done < 1) throw std::runtime_error("download: spp_done must be >= 1");
    Impl& m = *impl_;
    const size_t P = (size_t)m.pixels;
    if (!m.pinned) gpu::check(hipHostMalloc((void**)&m.pinned, 5 * P * sizeof(float), hipHostMallocDefault), "hipHostMalloc film");
    const int K = P >= (size_t(1) << 16) ? 4 : 1;
    const size_t step = (P + K - 1) / K;
    for (int k = 0; k < K; ++k) {
        if (!m.dl_ev[k]) gpu::check(hipEventCreateWithFlags(&m.dl_ev[k], hipEventDisableTiming), "hipEventCreate");
        const size_t a = std::min(P, k * step), n = std::min(step, P - a);
        if (rgb && n)
            gpu::check(hipMemcpyAsync(m.pinned + 3 * a, m.film.rgb + 3 * a, 3 * n * sizeof(float),
                                      hipMemcpyDeviceToHost, m.stream), "download rgb");
        if (lum_var && n) {
            gpu::check(hipMemcpyAsync(m.pinned + 3 * P + a, m.film.lum + a, n * sizeof(float), hipMemcpyDeviceToHost,
                                      m.stream), "download lum");
            gpu::check(hipMemcpyAsync(m.pinned + 4 * P + a, m.film.lum2 + a, n * sizeof(float), hipMemcpyDeviceToHost,
                                      m.stream), "download lum2");
        }
        gpu::check(hipEventRecord(m.dl_ev[k], m.stream), "hipEventRecord");
    }
    // Each range is finalized on its own host thread as soon as its copy has landed (range 0 on
    // this one): the binary64 finalize of a 720p image takes longer than its copy.
    auto finish = [&](int k) {
        gpu::check(hipEventSynchronize(m.dl_ev[k]), "download sync");
        const size_t a = std::min(P, k * step), n = std::min(step, P - a);
        finalize_film(m.pinned + 3 * a, m.pinned + 3 * P + a, m.pinned + 4 * P + a, (int)n, spp_done,
                      rgb ? rgb + 3 * a : nullptr, lum_var ? lum_var + a : nullptr);
    };
    std::exception_ptr err[4];
    std::vector<std::thread> helpers;
    for (int k = 1; k < K; ++k)
        helpers.emplace_back([&, k] {
            try {
                finish(k);
            } catch (...) {
                err[k] = std::current_exception();
            }
        });
    try {
        finish(0);
    } catch (...) {
        err[0] = std::current_exception();
    }
    for (std::thread& t : helpers) t.join();
    for (int k = 0; k < K; ++k)
        if (err[k]) std::rethrow_exception(err[k]);
}

void GpuRenderer::download_mean(int spp_done, std::vector<float>* rgb) {
    if (spp_done < 1) throw std::runtime_error("download_mean: spp_done must be >= 1");
    const float* h = stage_film(*impl_, true, false, false);
    rgb->resize(3 * (size_t)impl_->pixels);
    finalize_film(h, nullptr, nullptr, impl_->pixels, spp_done, rgb->data(), nullptr);
}

RenderResult GpuRenderer::render(const RenderParams& p, RenderStats* stats) {
    reset();
    RenderStats local{};
    run_passes(0, p.spp, p, &local);
    if (stats) *stats = local;
    return download(p.spp);
}

uint64_t GpuRenderer::device_bytes() const {
    return impl_->own_bytes + (impl_->pe ? impl_->pe->bytes() : 0);
}

const std::string& GpuRenderer::device_name() const {
    static std::string name;
    if (name.empty()) {
        try {
            name = query_device().name;
        } catch (...) {
            name = "unknown";
        }
    }
    return name;
}

const std::vector<BounceTrace>& GpuRenderer::trace() const { return impl_->trace; }

DeviceFacts query_device() {
    int dev = 0;
    gpu::check(hipGetDevice(&dev), "hipGetDevice");
    hipDeviceProp_t prop;
    gpu::check(hipGetDeviceProperties(&prop, dev), "hipGetDeviceProperties");
    size_t fr = 0, tot = 0;
    gpu::check(hipMemGetInfo(&fr, &tot), "hipMemGetInfo");
    int drv = 0, rt = 0;
    (void)hipDriverGetVersion(&drv);
    (void)hipRuntimeGetVersion(&rt);
    DeviceFacts f;
    f.name = prop.name;
    f.arch = prop.gcnArchName;
    if (f.name.empty()) f.name = "AMD Instinct (" + f.arch.substr(0, f.arch.find(':')) + ")";
    f.driver_version = std::to_string(drv);
    f.runtime_version = std::to_string(rt);
    f.compute_units = prop.multiProcessorCount;
    f.total_memory_bytes = tot;
    f.free_memory_bytes = fr;
    return f;
}

}  // namespace wf
