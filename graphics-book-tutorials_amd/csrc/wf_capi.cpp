// C ABI over the C++ engine (include/wfamd.h). Every entry point catches exceptions and maps
// them to status codes with a per-thread message, so no C++ exception crosses the boundary.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include "../../include/wfamd.h"
#include "wf_cpu.h"
#include "wf_persistent.h"
#include "wf_queues.h"
#include "wf_renderer.h"
#include "wf_handles.h"
#include "wf_scene.h"

struct wf_renderer {
    std::unique_ptr<wf::GpuRenderer> r;
    std::string name;
};
struct wf_scan_ws {
    wf::gpu::ScanWorkspace ws;
};

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) { return wf::set_last_error(code, msg); }

template <class F>
int guard(F&& f) {
    try {
        return f();
    } catch (const std::invalid_argument& e) {
        return fail(WF_ERR_INVALID, e.what());
    } catch (const std::exception& e) {
        return fail(WF_ERR_RUNTIME, e.what());
    } catch (...) {
        return fail(WF_ERR_RUNTIME, "unknown error");
    }
}

wf::RenderParams to_params(const wf_render_params* p) {
    wf::RenderParams r;
    if (!p) return r;
    r.spp = p->spp;
    r.seed = p->seed;
    r.max_depth = p->max_depth;
    r.threads = p->threads;
    r.count_rays = p->count_rays != 0;
    r.stage_timing = p->stage_timing != 0;
    r.queue_capacity = p->queue_capacity;
    r.trace_bounces = p->trace_bounces != 0;
    r.block_size = p->block_size;
    r.pool_paths = p->pool_paths;
    r.scan_tile = p->scan_tile;
    r.lds_nodes = p->lds_nodes;
    r.lds_kib = p->lds_kib;
    r.refill = p->refill;
    r.leaf_batch = p->leaf_batch;
    r.dir_bins = p->dir_bins;
    r.shade_block = p->shade_block;
    r.graph_iters = p->graph_iters;
    return r;
}

void check_params(const wf::RenderParams& p) {
    if (p.spp < 1) throw std::invalid_argument("spp must be >= 1");
    if (p.max_depth < 0 || p.max_depth > 1024) throw std::invalid_argument("max_depth out of range");
}

void to_stats(const wf::RenderStats& s, wf_render_stats* o) {
    if (!o) return;
    std::memset(o, 0, sizeof(*o));
    o->render_ms = s.render_ms;
    o->total_ms = s.total_ms;
    o->extension_rays = s.extension_rays;
    o->shadow_rays = s.shadow_rays;
    for (int k = 0; k < 5; ++k) o->queue_high_water[k] = s.queue_high_water[k];
    o->queue_overflows = s.queue_overflows;
    o->invalid_hits = s.invalid_hits;
    o->peak_vram_bytes = s.peak_vram_bytes;
    o->stage_ms[0] = s.stages.extend;
    o->stage_ms[1] = s.stages.classify;
    o->stage_ms[2] = s.stages.shade_lambert;
    o->stage_ms[3] = s.stages.shade_metal;
    o->stage_ms[4] = s.stages.shade_dielectric;
    o->stage_ms[5] = s.stages.shadow;
    o->stage_ms[6] = s.stages.compact;
    o->stage_ms[7] = s.stages.other;
    o->iterations = s.iterations;
    for (int k = 0; k < 8; ++k) {
        o->stage_items[k] = s.stages.items[k];
        o->stage_launches[k] = s.stages.launches[k];
        o->stage_bytes[k] = s.stages.bytes[k];
        o->stage_threads[k] = s.stages.threads[k];
    }
}

float bits_f(int v) {
    float f;
    std::memcpy(&f, &v, 4);
    return f;
}
int bits_i(float f) {
    int v;
    std::memcpy(&v, &f, 4);
    return v;
}

// Uploads each distinct host scene once; the views stay valid while the cache lives.
struct SceneCache {
    std::map<const wf_scene*, std::unique_ptr<wf::gpu::DeviceScene>> map;
    const wf::SceneView* get(const wf_scene* s) {
        auto it = map.find(s);
        if (it == map.end()) {
            auto d = std::make_unique<wf::gpu::DeviceScene>();
            d->upload(s->s);
            it = map.emplace(s, std::move(d)).first;
        }
        return &it->second->view;
    }
    uint64_t bytes() const {
        uint64_t b = 0;
        for (auto& kv : map) b += kv.second->bytes;
        return b;
    }
};

// Fixed pinned host buffer for film downloads, allocated once: a copy into it runs at DMA speed,
// where a pageable destination goes through the runtime's staging. Films stream through its two
// halves in pixel chunks: the next chunk is copied into one half while the host finalizes the
// chunk in the other. (A buffer grown to each larger film cost the serving loop more than
// it saved: every growth page-locks tens of MB, and freeing the old one waits for the device.)
struct PinnedStage {
    static constexpr size_t kFloats = size_t(4) << 20;   // 16 MiB
    float* p = nullptr;
    hipEvent_t ev[2] = {};   // copies into half 0 / half 1 have landed
    PinnedStage() = default;
    PinnedStage(const PinnedStage&) = delete;
    PinnedStage& operator=(const PinnedStage&) = delete;
    ~PinnedStage() {
        if (p) (void)hipHostFree(p);
        for (hipEvent_t e : ev)
            if (e) (void)hipEventDestroy(e);
    }
    float* get() {
        if (!p) wf::gpu::check(hipHostMalloc((void**)&p, kFloats * sizeof(float), hipHostMallocDefault), "hipHostMalloc film stage");
        for (hipEvent_t& e : ev)
            if (!e) wf::gpu::check(hipEventCreateWithFlags(&e, hipEventDisableTiming), "hipEventCreate");
        return p;
    }
};

// Finalizes the n pixels of a staged chunk (rgb sums, then luminance sums, then squared sums) on up
// to four host threads, the calling one included: the binary64 variance of a chunk takes longer
// than its copy.
static void finalize_split(const float* h, size_t n, int spp, float* rgb, float* lum_var) {
    const size_t parts = n >= 65536 ? 4 : 1, step = (n + parts - 1) / parts;
    auto part = [&](size_t k) {
        const size_t b = std::min(n, k * step), m = std::min(step, n - b);
        if (m)
            wf::finalize_film(h + 3 * b, h + 3 * n + b, h + 4 * n + b, (int)m, spp, rgb ? rgb + 3 * b : nullptr,
                              lum_var ? lum_var + b : nullptr);
    };
    std::vector<std::thread> helpers;
    for (size_t k = 1; k < parts; ++k) helpers.emplace_back(part, k);
    part(0);
    for (std::thread& t : helpers) t.join();
}

struct DevFilmAlloc {
    wf::gpu::DeviceFilm film{};
    int pixels = 0;
    hipStream_t owner = nullptr;   // stream-ordered allocation (serving loop)
    // pool: the memory pool to allocate from (null: the device's default pool).
    void alloc(int px, hipStream_t stream, hipMemPool_t pool = nullptr) {
        pixels = px;
        owner = stream;
        const size_t bytes = (size_t)px * 5 * sizeof(float);
        if (pool) wf::gpu::check(hipMallocFromPoolAsync((void**)&film.rgb, bytes, pool, stream), "hipMallocFromPoolAsync film");
        else wf::gpu::check(hipMallocAsync((void**)&film.rgb, bytes, stream), "hipMallocAsync film");
        film.lum = film.rgb + 3 * (size_t)px;
        film.lum2 = film.lum + px;
        film.pixels = px;
        wf::gpu::check(hipMemsetAsync(film.rgb, 0, (size_t)px * 5 * sizeof(float), stream), "hipMemset film");
    }
    void download(int spp, float* rgb, float* lum_var, hipStream_t stream, PinnedStage& stage) {
        float* base = stage.get();
        const size_t P = (size_t)pixels, half = PinnedStage::kFloats / 2, chunk = half / 5;
        // Copies pixels [a, a + chunk) into half `s` of the stage and marks them with ev[s].
        auto issue = [&](size_t a, int s) {
            float* h = base + s * half;
            const size_t n = std::min(chunk, P - a);
            if (rgb)
                wf::gpu::check(hipMemcpyAsync(h, film.rgb + 3 * a, 3 * n * sizeof(float), hipMemcpyDeviceToHost, stream),
                               "download film");
            if (lum_var) {
                wf::gpu::check(hipMemcpyAsync(h + 3 * n, film.lum + a, n * sizeof(float), hipMemcpyDeviceToHost, stream),
                               "download film");
                wf::gpu::check(hipMemcpyAsync(h + 4 * n, film.lum2 + a, n * sizeof(float), hipMemcpyDeviceToHost, stream),
                               "download film");
            }
            wf::gpu::check(hipEventRecord(stage.ev[s], stream), "hipEventRecord");
        };
        if (P > 0) issue(0, 0);
        int s = 0;
        for (size_t a = 0; a < P; a += chunk, s ^= 1) {
            // The other half was finalized in the previous step, so the next chunk may land there.
            if (a + chunk < P) issue(a + chunk, s ^ 1);
            wf::gpu::check(hipEventSynchronize(stage.ev[s]), "download sync");
            const float* h = base + s * half;
            const size_t n = std::min(chunk, P - a);
            finalize_split(h, n, spp, rgb ? rgb + 3 * a : nullptr, lum_var ? lum_var + a : nullptr);
        }
    }
    void release() {
        if (film.rgb) (void)(owner ? hipFreeAsync(film.rgb, owner) : hipFree(film.rgb));
        film = wf::gpu::DeviceFilm{};
    }
};

}  // namespace

uint64_t wf::next_scene_uid() {
    static std::atomic<uint64_t> next{1};
    return next.fetch_add(1);
}

int wf::set_last_error(int code, const std::string& msg) {
    g_err = msg;
    return code;
}

extern "C" {

const char* wf_last_error(void) { return g_err.c_str(); }
const char* wf_version(void) { return "wfamd 0.1 (gfx950)"; }

// ---------------------------------------------------------------------------------------------
// scenes
// ---------------------------------------------------------------------------------------------
int wf_scene_create(const char* name, int width, int height, wf_scene** out) {
    return guard([&] {
        if (!name || !out) throw std::invalid_argument("null argument");
        if (width < 1 || height < 1 || width > 8192 || height > 8192) throw std::invalid_argument("image size outside 1..8192");
        auto s = std::make_unique<wf_scene>();
        s->s = wf::make_scene(name, width, height);
        *out = s.release();
        return (int)WF_OK;
    });
}

int wf_scene_from_arrays(const wf_scene_desc* d, wf_scene** out) {
    return guard([&] {
        if (!d || !out) throw std::invalid_argument("null argument");
        if (d->width < 1 || d->height < 1 || d->width > 8192 || d->height > 8192) throw std::invalid_argument("image size outside 1..8192");
        auto s = std::make_unique<wf_scene>();
        wf::Scene& sc = s->s;
        sc.name = "arrays";
        sc.width = d->width;
        sc.height = d->height;
        for (int i = 0; i < d->n_spheres; ++i) {
            const float* p = d->spheres + 8 * i;
            wf::Sphere sp{};
            sp.center = wf::Vec3{p[0], p[1], p[2]};
            sp.radius = p[3];
            sp.material = bits_i(p[4]);
            sc.spheres.push_back(sp);
        }
        for (int i = 0; i < d->n_vertices; ++i)
            sc.vertices.push_back(wf::Vec3{d->vertices[3 * i], d->vertices[3 * i + 1], d->vertices[3 * i + 2]});
        for (int i = 0; i < d->n_triangles; ++i) {
            const int* t = d->triangles + 4 * i;
            sc.triangles.push_back(wf::Triangle{t[0], t[1], t[2], t[3]});
        }
        for (int i = 0; i < d->n_materials; ++i) {
            const float* m = d->materials + 12 * i;
            wf::Material mt{};
            mt.type = bits_i(m[0]);
            mt.ior = m[1];
            mt.fuzz = m[2];
            mt.albedo = wf::Vec3{m[4], m[5], m[6]};
            mt.emission = wf::Vec3{m[8], m[9], m[10]};
            if (mt.type < 0 || mt.type > 3) throw std::invalid_argument("material type out of range");
            sc.materials.push_back(mt);
        }
        if (d->env) {
            const float* e = d->env;
            sc.env.env = wf::Vec3{e[0], e[1], e[2]};
            sc.env.sky_mode = bits_i(e[3]);
            sc.env.sky_lo = wf::Vec3{e[4], e[5], e[6]};
            sc.env.sky_light = bits_i(e[7]);
            sc.env.sky_hi = wf::Vec3{e[8], e[9], e[10]};
            sc.env.sun_light = bits_i(e[11]);
            sc.env.sun_dir = wf::Vec3{e[12], e[13], e[14]};
            sc.env.sun_E = wf::Vec3{e[16], e[17], e[18]};
        }
        if (d->camera) {
            const float* c = d->camera;
            sc.camera.position = wf::Vec3{c[0], c[1], c[2]};
            sc.camera.look_at = wf::Vec3{c[3], c[4], c[5]};
            sc.camera.up = wf::Vec3{c[6], c[7], c[8]};
            sc.camera.vfov_deg = c[9];
        }
        try {
            wf::finalize_scene(sc);
        } catch (const std::runtime_error& e) {
            throw std::invalid_argument(e.what());
        }
        *out = s.release();
        return (int)WF_OK;
    });
}

int wf_scene_orbit(wf_scene* scene, float yaw_deg) {
    return guard([&] {
        if (!scene) throw std::invalid_argument("null scene");
        wf::orbit_camera(scene->s, yaw_deg);
        scene->info_valid = false;
        return (int)WF_OK;
    });
}

int wf_scene_view(const wf_scene* scene, int width, int height, float yaw_deg, int orbit, wf_scene** out) {
    return guard([&] {
        if (!scene || !out) throw std::invalid_argument("null argument");
        if (width < 1 || height < 1 || width > 8192 || height > 8192) throw std::invalid_argument("image size outside 1..8192");
        auto s = std::make_unique<wf_scene>();
        s->s = scene->s;
        s->s.width = width;
        s->s.height = height;
        if (orbit) wf::orbit_camera(s->s, yaw_deg);
        *out = s.release();
        return (int)WF_OK;
    });
}

void wf_scene_destroy(wf_scene* scene) { delete scene; }

int wf_scene_info_get(const wf_scene* scene, wf_scene_info* out) {
    return guard([&] {
        if (!scene || !out) throw std::invalid_argument("null argument");
        if (scene->info_valid) {
            *out = scene->info;
            return (int)WF_OK;
        }
        const wf::Scene& s = scene->s;
        std::memset(out, 0, sizeof(*out));
        out->width = s.width;
        out->height = s.height;
        out->n_spheres = (int)s.spheres.size();
        out->n_triangles = (int)s.triangles.size();
        out->n_vertices = (int)s.vertices.size();
        out->n_materials = (int)s.materials.size();
        out->n_lights = (int)s.lights.size();
        out->n_nodes = (int)s.nodes.size();
        wf::BvhStats st = wf::bvh_stats(s);
        out->bvh_leaves = st.leaves;
        out->bvh_internal = st.internal;
        out->bvh_max_depth = st.max_depth;
        out->bvh_sah_cost = st.sah_cost;
        out->hash = wf::scene_hash(s);
        scene->info = *out;
        scene->info_valid = true;
        return (int)WF_OK;
    });
}

int wf_scene_export(const wf_scene* scene, float* spheres, float* vertices, int* triangles, float* materials, float* env,
                    float* camera) {
    return guard([&] {
        if (!scene) throw std::invalid_argument("null scene");
        const wf::Scene& s = scene->s;
        if (spheres)
            for (size_t i = 0; i < s.spheres.size(); ++i) {
                float* p = spheres + 8 * i;
                const wf::Sphere& sp = s.spheres[i];
                p[0] = sp.center.x; p[1] = sp.center.y; p[2] = sp.center.z; p[3] = sp.radius;
                p[4] = bits_f(sp.material); p[5] = p[6] = p[7] = 0.0f;
            }
        if (vertices)
            for (size_t i = 0; i < s.vertices.size(); ++i) {
                vertices[3 * i] = s.vertices[i].x;
                vertices[3 * i + 1] = s.vertices[i].y;
                vertices[3 * i + 2] = s.vertices[i].z;
            }
        if (triangles)
            for (size_t i = 0; i < s.triangles.size(); ++i) {
                const wf::Triangle& t = s.triangles[i];
                triangles[4 * i] = t.i0; triangles[4 * i + 1] = t.i1; triangles[4 * i + 2] = t.i2; triangles[4 * i + 3] = t.material;
            }
        if (materials)
            for (size_t i = 0; i < s.materials.size(); ++i) {
                const wf::Material& m = s.materials[i];
                float* p = materials + 12 * i;
                p[0] = bits_f(m.type); p[1] = m.ior; p[2] = m.fuzz; p[3] = 0.0f;
                p[4] = m.albedo.x; p[5] = m.albedo.y; p[6] = m.albedo.z; p[7] = 0.0f;
                p[8] = m.emission.x; p[9] = m.emission.y; p[10] = m.emission.z; p[11] = 0.0f;
            }
        if (env) {
            const wf::EnvParams& e = s.env;
            float v[20] = {e.env.x, e.env.y, e.env.z, bits_f(e.sky_mode), e.sky_lo.x, e.sky_lo.y, e.sky_lo.z,
                           bits_f(e.sky_light), e.sky_hi.x, e.sky_hi.y, e.sky_hi.z, bits_f(e.sun_light), e.sun_dir.x,
                           e.sun_dir.y, e.sun_dir.z, 0.0f, e.sun_E.x, e.sun_E.y, e.sun_E.z, 0.0f};
            std::memcpy(env, v, sizeof(v));
        }
        if (camera) {
            const wf::Camera& c = s.camera;
            float v[10] = {c.position.x, c.position.y, c.position.z, c.look_at.x, c.look_at.y, c.look_at.z,
                           c.up.x, c.up.y, c.up.z, c.vfov_deg};
            std::memcpy(camera, v, sizeof(v));
        }
        return (int)WF_OK;
    });
}

int wf_scene_bvh(const wf_scene* scene, int* nodes, int* prim_order) {
    return guard([&] {
        if (!scene) throw std::invalid_argument("null scene");
        const wf::Scene& s = scene->s;
        if (nodes)
            for (size_t i = 0; i < s.nodes.size(); ++i) {
                const wf::BvhNode& n = s.nodes[i];
                int* p = nodes + 12 * i;
                p[0] = bits_i(n.lo.x); p[1] = bits_i(n.lo.y); p[2] = bits_i(n.lo.z); p[3] = n.skip;
                p[4] = bits_i(n.hi.x); p[5] = bits_i(n.hi.y); p[6] = bits_i(n.hi.z);
                p[7] = n.prim_start; p[8] = n.prim_count; p[9] = p[10] = p[11] = 0;
            }
        if (prim_order) std::memcpy(prim_order, s.prim_order.data(), s.prim_order.size() * sizeof(int));
        return (int)WF_OK;
    });
}

int wf_scene_camera_rig(const wf_scene* scene, float* rig12) {
    return guard([&] {
        if (!scene || !rig12) throw std::invalid_argument("null argument");
        wf::CameraRig r = wf::make_camera_rig(scene->s.camera, scene->s.width, scene->s.height);
        float v[12] = {r.origin.x, r.origin.y, r.origin.z, r.lower_left.x, r.lower_left.y, r.lower_left.z,
                       r.horizontal.x, r.horizontal.y, r.horizontal.z, r.vertical.x, r.vertical.y, r.vertical.z};
        std::memcpy(rig12, v, sizeof(v));
        return (int)WF_OK;
    });
}

// ---------------------------------------------------------------------------------------------
// renderer
// ---------------------------------------------------------------------------------------------
int wf_renderer_create(const wf_scene* scene, int mode, wf_renderer** out) {
    return guard([&] {
        if (!scene || !out) throw std::invalid_argument("null argument");
        if (mode != WF_MODE_MEGAKERNEL && mode != WF_MODE_WAVEFRONT && mode != WF_MODE_PERSISTENT)
            throw std::invalid_argument("renderer mode must be megakernel, wavefront or persistent");
        auto r = std::make_unique<wf_renderer>();
        r->r = std::make_unique<wf::GpuRenderer>(scene->s, (wf::GpuMode)mode);
        *out = r.release();
        return (int)WF_OK;
    });
}

void wf_renderer_destroy(wf_renderer* r) { delete r; }

int wf_renderer_render(wf_renderer* r, const wf_render_params* p, wf_render_stats* stats, float* rgb, float* lum_var) {
    return guard([&] {
        if (!r || !p) throw std::invalid_argument("null argument");
        wf::RenderParams rp = to_params(p);
        check_params(rp);
        wf::RenderStats st{};
        auto t0 = std::chrono::steady_clock::now();
        // GpuRenderer::render's steps, with the image finalized straight into the caller's buffers.
        r->r->reset();
        r->r->run_passes(0, rp.spp, rp, &st);
        if (rgb || lum_var) r->r->download_to(rp.spp, rgb, lum_var);
        auto t1 = std::chrono::steady_clock::now();
        st.total_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
        to_stats(st, stats);
        return (int)WF_OK;
    });
}

int wf_renderer_reset(wf_renderer* r) {
    return guard([&] {
        if (!r) throw std::invalid_argument("null renderer");
        r->r->reset();
        return (int)WF_OK;
    });
}

int wf_renderer_run_passes(wf_renderer* r, int first_sample, int count, const wf_render_params* p,
                           wf_render_stats* stats) {
    return guard([&] {
        if (!r || !p) throw std::invalid_argument("null argument");
        if (first_sample < 0 || count < 0) throw std::invalid_argument("negative sample range");
        wf::RenderParams rp = to_params(p);
        if (rp.max_depth < 0 || rp.max_depth > 1024) throw std::invalid_argument("max_depth out of range");
        wf::RenderStats st{};
        r->r->run_passes(first_sample, count, rp, &st);
        to_stats(st, stats);
        return (int)WF_OK;
    });
}

int wf_renderer_download(wf_renderer* r, int spp_done, float* rgb, float* lum_var) {
    return guard([&] {
        if (!r) throw std::invalid_argument("null renderer");
        if (spp_done < 1) throw std::invalid_argument("spp_done must be >= 1");
        r->r->download_to(spp_done, rgb, lum_var);
        return (int)WF_OK;
    });
}

int wf_renderer_download_mean(wf_renderer* r, int spp_done, float* rgb) {
    return guard([&] {
        if (!r || !rgb) throw std::invalid_argument("null argument");
        if (spp_done < 1) throw std::invalid_argument("spp_done must be >= 1");
        std::vector<float> v;
        r->r->download_mean(spp_done, &v);
        std::memcpy(rgb, v.data(), v.size() * sizeof(float));
        return (int)WF_OK;
    });
}

int wf_renderer_download_sums(wf_renderer* r, float* rgb, float* lum, float* lum2) {
    return guard([&] {
        if (!r) throw std::invalid_argument("null renderer");
        std::vector<float> a, b, c;
        r->r->download_sums(rgb ? &a : nullptr, lum ? &b : nullptr, lum2 ? &c : nullptr);
        if (rgb) std::memcpy(rgb, a.data(), a.size() * sizeof(float));
        if (lum) std::memcpy(lum, b.data(), b.size() * sizeof(float));
        if (lum2) std::memcpy(lum2, c.data(), c.size() * sizeof(float));
        return (int)WF_OK;
    });
}

uint64_t wf_renderer_device_bytes(const wf_renderer* r) { return r ? r->r->device_bytes() : 0; }

const char* wf_renderer_device_name(const wf_renderer* r) {
    if (!r) return "";
    return r->r->device_name().c_str();
}

int wf_renderer_trace(const wf_renderer* r, wf_bounce_trace* out, int capacity) {
    if (!r) return -1;
    const auto& t = r->r->trace();
    const int n = (int)t.size();
    for (int i = 0; i < n && i < capacity && out; ++i) {
        out[i].active = t[i].active;
        out[i].lambert = t[i].lambert;
        out[i].metal = t[i].metal;
        out[i].dielectric = t[i].dielectric;
        out[i].shadow = t[i].shadow;
        out[i].next_active = t[i].next_active;
        out[i].list_hash = t[i].list_hash;
    }
    return n;
}

// ---------------------------------------------------------------------------------------------
// CPU arm
// ---------------------------------------------------------------------------------------------
int wf_render_cpu(const wf_scene* scene, const wf_render_params* p, wf_render_stats* stats, float* rgb, float* lum_var) {
    return guard([&] {
        if (!scene || !p) throw std::invalid_argument("null argument");
        wf::RenderParams rp = to_params(p);
        check_params(rp);
        wf::CpuCounts c;
        double ms = 0.0;
        auto t0 = std::chrono::steady_clock::now();
        wf::RenderResult res = wf::render_cpu(scene->s, rp.spp, rp.seed, rp.max_depth, rp.threads, &c, &ms);
        auto t1 = std::chrono::steady_clock::now();
        wf::RenderStats st{};
        st.render_ms = ms;
        st.total_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
        st.extension_rays = c.extension_rays;
        st.shadow_rays = c.shadow_rays;
        st.invalid_hits = (long long)c.invalid_hits;
        to_stats(st, stats);
        if (rgb) std::memcpy(rgb, res.rgb.data(), res.rgb.size() * sizeof(float));
        if (lum_var) std::memcpy(lum_var, res.lum_var.data(), res.lum_var.size() * sizeof(float));
        return (int)WF_OK;
    });
}

int wf_render_cpu_sums(const wf_scene* scene, int first_sample, int count, const wf_render_params* p,
                       wf_render_stats* stats, float* rgb, float* lum, float* lum2) {
    return guard([&] {
        if (!scene || !p || !rgb || !lum || !lum2) throw std::invalid_argument("null argument");
        if (first_sample < 0 || count < 0) throw std::invalid_argument("negative sample range");
        wf::RenderParams rp = to_params(p);
        const wf::Scene& s = scene->s;
        wf::FilmSums film;
        film.width = s.width;
        film.height = s.height;
        const size_t P = (size_t)s.width * s.height;
        film.rgb.assign(rgb, rgb + 3 * P);
        film.lum.assign(lum, lum + P);
        film.lum2.assign(lum2, lum2 + P);
        wf::CpuCounts c;
        auto t0 = std::chrono::steady_clock::now();
        wf::cpu_render_passes(s, first_sample, count, rp.seed, rp.max_depth, rp.threads, &film, &c);
        auto t1 = std::chrono::steady_clock::now();
        std::memcpy(rgb, film.rgb.data(), 3 * P * sizeof(float));
        std::memcpy(lum, film.lum.data(), P * sizeof(float));
        std::memcpy(lum2, film.lum2.data(), P * sizeof(float));
        wf::RenderStats st{};
        st.render_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
        st.extension_rays = c.extension_rays;
        st.shadow_rays = c.shadow_rays;
        st.invalid_hits = (long long)c.invalid_hits;
        to_stats(st, stats);
        return (int)WF_OK;
    });
}

int wf_finalize(const float* rgb_sum, const float* lum_sum, const float* lum2_sum, int pixels, int spp, float* rgb,
                float* lum_var) {
    return guard([&] {
        if (!rgb_sum || !lum_sum || !lum2_sum || pixels < 0 || spp < 1) throw std::invalid_argument("bad argument");
        wf::finalize_film(rgb_sum, lum_sum, lum2_sum, pixels, spp, rgb, lum_var);
        return (int)WF_OK;
    });
}

int wf_query_device(wf_device_facts* out) {
    return guard([&] {
        if (!out) throw std::invalid_argument("null argument");
        wf::DeviceFacts f = wf::query_device();
        std::memset(out, 0, sizeof(*out));
        std::strncpy(out->name, f.name.c_str(), sizeof(out->name) - 1);
        std::strncpy(out->arch, f.arch.c_str(), sizeof(out->arch) - 1);
        out->compute_units = f.compute_units;
        out->runtime_version = std::stoi(f.runtime_version);
        out->driver_version = std::stoi(f.driver_version);
        out->total_memory_bytes = f.total_memory_bytes;
        out->free_memory_bytes = f.free_memory_bytes;
        return (int)WF_OK;
    });
}

// ---------------------------------------------------------------------------------------------
// primitives
// ---------------------------------------------------------------------------------------------
int wf_scan_workspace_create(int capacity, wf_scan_ws** out) {
    return guard([&] {
        if (!out || capacity < 0) throw std::invalid_argument("bad argument");
        auto w = std::make_unique<wf_scan_ws>();
        wf::gpu::check(wf::gpu::scan_workspace_create(&w->ws, capacity), "scan_workspace_create");
        *out = w.release();
        return (int)WF_OK;
    });
}

void wf_scan_workspace_destroy(wf_scan_ws* ws) {
    if (!ws) return;
    wf::gpu::scan_workspace_destroy(&ws->ws);
    delete ws;
}

int wf_scan_exclusive_u32(const uint32_t* d_in, uint32_t* d_out, int n, uint32_t* d_total, wf_scan_ws* ws, void* stream,
                          int tile) {
    if (!ws) return fail(WF_ERR_INVALID, "null workspace");
    hipError_t e = wf::gpu::scan_exclusive_u32(d_in, d_out, n, d_total, ws->ws, (hipStream_t)stream,
                                               tile > 0 ? tile : wf::gpu::kDefaultScanTile);
    if (e == hipErrorInvalidValue) return fail(WF_ERR_INVALID, "scan: n out of range or bad tile");
    if (e != hipSuccess) return fail(WF_ERR_RUNTIME, hipGetErrorString(e));
    return WF_OK;
}

int wf_compact_u32(const uint32_t* d_values, const uint32_t* d_flags, uint32_t* d_out, uint32_t* d_count,
                   uint32_t* d_scratch, int n, wf_scan_ws* ws, void* stream, int tile) {
    (void)d_scratch;   // the reduce-then-scan compaction needs no per-element scratch
    if (!ws) return fail(WF_ERR_INVALID, "null workspace");
    hipError_t e = wf::gpu::compact_u32(d_values, d_flags, d_out, d_count, n, ws->ws, (hipStream_t)stream,
                                        tile > 0 ? tile : wf::gpu::kDefaultScanTile);
    if (e == hipErrorInvalidValue) return fail(WF_ERR_INVALID, "compact: n out of range or bad tile");
    if (e != hipSuccess) return fail(WF_ERR_RUNTIME, hipGetErrorString(e));
    return WF_OK;
}

int wf_compact_u32_noalloc(const uint32_t* d_values, const uint32_t* d_flags, uint32_t* d_out, uint32_t* d_count, int n,
                           void* stream, int tile) {
    if (n < 0 || n > wf::gpu::kNoAllocMaxN)
        return fail(WF_ERR_INVALID, "compact (no allocation): n outside 0..2^26, the size of the library's static partials");
    const int t = tile > 0 ? tile : wf::gpu::kDefaultScanTile;
    if (wf::gpu::scan_items_for_tile(t) == 0) return fail(WF_ERR_INVALID, "compact (no allocation): tile must be 512, 1024, 2048 or 4096");
    hipError_t e = wf::gpu::compact_u32_noalloc(d_values, d_flags, d_out, d_count, n, (hipStream_t)stream, t);
    if (e == hipErrorInvalidValue) return fail(WF_ERR_INVALID, "compact (no allocation): bad argument");
    if (e != hipSuccess) return fail(WF_ERR_RUNTIME, hipGetErrorString(e));
    return WF_OK;
}

int wf_test_queue_push(const uint32_t* d_flags, int n, uint32_t* d_items, uint32_t capacity, uint32_t* d_counter,
                       uint32_t* d_overflow, void* stream) {
    hipError_t e = wf::gpu::test_queue_push(d_flags, n, d_items, capacity, d_counter, d_overflow, (hipStream_t)stream);
    if (e == hipErrorInvalidValue) return fail(WF_ERR_INVALID, "queue push: n < 0");
    if (e != hipSuccess) return fail(WF_ERR_RUNTIME, hipGetErrorString(e));
    return WF_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------------------------
// batch
// ---------------------------------------------------------------------------------------------
namespace {

// Size and camera rig of a job's view (validated), filled into the engine job.
void job_view(const wf_job& in, wf::gpu::PersistentJob* j, int* w, int* h) {
    const wf::Scene& sc = in.scene->s;
    *w = in.width > 0 ? in.width : sc.width;
    *h = in.height > 0 ? in.height : sc.height;
    if (in.width < 0 || in.height < 0 || *w > 8192 || *h > 8192) throw std::invalid_argument("job image size outside 1..8192");
    j->has_view = true;
    j->width = *w;
    j->height = *h;
    j->cam = wf::view_rig(sc, *w, *h, in.orbit != 0, in.yaw_deg);
}

wf::gpu::PersistentTuning to_tuning(const wf_render_params* p) {
    wf::gpu::PersistentTuning t;
    if (p) {
        t.block_size = p->block_size;
        t.pool_paths = p->pool_paths;
        t.lds_kib = p->lds_kib;
        t.refill = p->refill;
        t.leaf_batch = p->leaf_batch;
        t.dir_bins = p->dir_bins;
        t.shade_block = p->shade_block;
        t.graph_iters = p->graph_iters;
    }
    return t;
}

// Pinned host mirror of a DevCounters block: counter reads are DMA copies into it, not pageable
// copies through the runtime's staging buffers.
struct PinnedCounters {
    wf::gpu::DevCounters* h = nullptr;
    PinnedCounters() = default;
    PinnedCounters(const PinnedCounters&) = delete;
    PinnedCounters& operator=(const PinnedCounters&) = delete;
    ~PinnedCounters() {
        if (h) (void)hipHostFree(h);
    }
    wf::gpu::DevCounters read(const wf::gpu::DevCounters* d, hipStream_t stream) {
        if (!h) wf::gpu::check(hipHostMalloc((void**)&h, sizeof(*h), hipHostMallocDefault), "hipHostMalloc counters");
        wf::gpu::check(hipMemcpyAsync(h, d, sizeof(*h), hipMemcpyDeviceToHost, stream), "copy counters");
        wf::gpu::check(hipStreamSynchronize(stream), "counters sync");
        return *h;
    }
};

struct BatchSetup {
    hipStream_t stream = nullptr;
    wf::gpu::DevCounters* ctr = nullptr;
    PinnedCounters hctr;
    SceneCache scenes;
    std::vector<DevFilmAlloc> films;
    std::vector<wf::gpu::PersistentJob> jobs;
    PinnedStage stage;
    ~BatchSetup() {
        for (DevFilmAlloc& f : films) f.release();
        if (ctr) (void)hipFree(ctr);
        if (stream) (void)hipStreamDestroy(stream);
    }
    void build(const wf_job* in, int n) {
        wf::gpu::check(hipStreamCreate(&stream), "hipStreamCreate");
        wf::gpu::check(hipMalloc(&ctr, sizeof(wf::gpu::DevCounters)), "hipMalloc counters");
        wf::gpu::check(hipMemsetAsync(ctr, 0, sizeof(wf::gpu::DevCounters), stream), "hipMemset counters");
        films.resize(n);
        for (int i = 0; i < n; ++i) {
            if (!in[i].scene) throw std::invalid_argument("job without a scene");
            if (in[i].spp < 1) throw std::invalid_argument("job spp must be >= 1");
            if (in[i].max_depth < 0 || in[i].max_depth > 1024) throw std::invalid_argument("job max_depth out of range");
            const wf::SceneView* v = scenes.get(in[i].scene);
            wf::gpu::PersistentJob j;
            int w = 0, h = 0;
            job_view(in[i], &j, &w, &h);
            films[i].alloc(w * h, stream);
            j.scene = v;
            j.film = films[i].film;
            j.first_sample = 0;
            j.spp = in[i].spp;
            j.seed = in[i].seed;
            j.max_depth = in[i].max_depth;
            jobs.push_back(j);
        }
    }
};

}  // namespace

extern "C" int wf_batch_render(const wf_job* jobs, int n_jobs, const wf_render_params* tuning, wf_batch_stats* stats) {
    return guard([&] {
        if (!jobs || n_jobs < 1) throw std::invalid_argument("no jobs");
        BatchSetup b;
        b.build(jobs, n_jobs);
        wf::gpu::PersistentEngine eng;
        wf::gpu::PersistentTuning t = to_tuning(tuning);
        hipEvent_t e0, e1;
        wf::gpu::check(hipEventCreate(&e0), "hipEventCreate");
        wf::gpu::check(hipEventCreate(&e1), "hipEventCreate");
        wf::gpu::PersistentResult pr;
        wf::StageTimes stages;
        const bool timed = tuning && tuning->stage_timing;
        wf::gpu::check(hipEventRecord(e0, b.stream), "hipEventRecord");
        eng.render(b.jobs.data(), n_jobs, t, b.ctr, b.stream, &pr, timed ? &stages : nullptr);
        wf::gpu::check(hipEventRecord(e1, b.stream), "hipEventRecord");
        wf::gpu::check(hipEventSynchronize(e1), "batch sync");
        float ms = 0.0f;
        wf::gpu::check(hipEventElapsedTime(&ms, e0, e1), "hipEventElapsedTime");
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        for (int i = 0; i < n_jobs; ++i) b.films[i].download(jobs[i].spp, jobs[i].rgb, jobs[i].lum_var, b.stream, b.stage);
        if (stats) {
            const wf::gpu::DevCounters c = b.hctr.read(b.ctr, b.stream);
            std::memset(stats, 0, sizeof(*stats));
            stats->render_ms = ms;
            stats->extension_rays = c.extension_rays;
            stats->shadow_rays = c.shadow_rays;
            stats->invalid_hits = c.invalid_hits;
            stats->queue_overflows = c.overflows;
            stats->iterations = pr.iterations;
            stats->graph_launches = pr.graph_launches;
            stats->block_size = t.block_size;
            stats->pool_paths = eng.pool_paths();
            stats->lds_kib = t.lds_kib;
            stats->refill = t.refill;
            stats->leaf_batch = t.leaf_batch;
            stats->dir_bins = t.dir_bins;
            stats->shade_block = t.shade_block;
            stats->device_bytes = eng.bytes() + b.scenes.bytes();
            const double ms8[8] = {stages.extend,           stages.classify, stages.shade_lambert, stages.shade_metal,
                                   stages.shade_dielectric, stages.shadow,   stages.compact,       stages.other};
            for (int k = 0; k < 8; ++k) {
                stats->stage_ms[k] = ms8[k];
                stats->stage_items[k] = stages.items[k];
                stats->stage_launches[k] = stages.launches[k];
                stats->stage_bytes[k] = stages.bytes[k];
                stats->stage_threads[k] = stages.threads[k];
            }
        }
        return (int)WF_OK;
    });
}

extern "C" uint64_t wf_cached_memory_bytes(void) { return wf::gpu::cached_block_bytes(); }

extern "C" int wf_release_cached_memory(uint64_t* freed_bytes) {
    return guard([&] {
        const uint64_t n = wf::gpu::release_cached_blocks();
        if (freed_bytes) *freed_bytes = n;
        return (int)WF_OK;
    });
}

extern "C" int wf_autotune(const wf_job* jobs, int n_jobs, int probe_spp, wf_render_params* tuning) {
    return wf_autotune_budget(jobs, n_jobs, probe_spp, 0, tuning);
}

extern "C" int wf_autotune_budget(const wf_job* jobs, int n_jobs, int probe_spp, uint64_t state_budget_bytes,
                                  wf_render_params* tuning) {
    return guard([&] {
        if (!jobs || n_jobs < 1 || !tuning || probe_spp < 1) throw std::invalid_argument("bad argument");
        BatchSetup b;
        b.build(jobs, n_jobs);
        wf::gpu::PersistentTuning best =
            wf::gpu::autotune_persistent(b.jobs.data(), n_jobs, b.stream, probe_spp, nullptr, state_budget_bytes);
        tuning->block_size = best.block_size;
        tuning->pool_paths = best.pool_paths;
        tuning->lds_kib = best.lds_kib;
        tuning->refill = best.refill;
        tuning->leaf_batch = best.leaf_batch;
        tuning->dir_bins = best.dir_bins;
        tuning->shade_block = best.shade_block;
        return (int)WF_OK;
    });
}

// ---------------------------------------------------------------------------------------------
// serving loop
// ---------------------------------------------------------------------------------------------
namespace {

// Device bytes of a scene's upload (DeviceScene::upload: nodes, primitives, materials, lights).
uint64_t scene_upload_bytes(const wf::Scene& s) {
    return (uint64_t)s.dnodes.size() * sizeof(wf::Node) + (uint64_t)s.dprims.size() * sizeof(wf::Prim) +
           (uint64_t)s.materials.size() * sizeof(wf::Material) + (uint64_t)s.lights.size() * sizeof(wf::Light);
}

// The serving loop's own stream-ordered memory pool: freed request memory stays in it for the next
// requests (release threshold at the maximum) without changing the device's default pool, which
// other renderers and batches of the process use. Destroyed after everything allocated from it.
struct ServerPool {
    hipMemPool_t pool = nullptr;
    ServerPool() = default;
    ServerPool(const ServerPool&) = delete;
    ServerPool& operator=(const ServerPool&) = delete;
    void create(int dev) {
        hipMemPoolProps props{};
        props.allocType = hipMemAllocationTypePinned;
        props.handleTypes = hipMemHandleTypeNone;
        props.location.type = hipMemLocationTypeDevice;
        props.location.id = dev;
        wf::gpu::check(hipMemPoolCreate(&pool, &props), "hipMemPoolCreate");
        uint64_t keep = ~0ull;
        wf::gpu::check(hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &keep), "hipMemPoolSetAttribute");
    }
    ~ServerPool() {
        if (pool) (void)hipMemPoolDestroy(pool);
    }
};

}  // namespace

struct wf_server {
    // A request renders in consecutive sample ranges (chunks), each an engine job with its own
    // sample buffer of pixels x chunk spp x 16 B; chunks are resolved into the request's film in
    // sample order, so the film holds the bits of one render of all its samples, and the memory a
    // request holds is bounded by its pixels, not its samples.
    struct Chunk {
        int64_t seq;      // engine job
        int spp;
        uint64_t bytes;
    };
    struct Req {
        int64_t id;
        wf_job job;
        uint64_t scene_uid = 0;
        int px = 0;                      // pixels
        int chunk_spp = 1;
        int next_sample = 0;             // samples admitted
        int resolved = 0;                // samples resolved into the film
        std::deque<Chunk> chunks;        // admitted, not resolved, in sample order
        wf::gpu::PersistentJob tmpl;     // the request's view (scene, size, camera, film, seed, depth)
        DevFilmAlloc film;
        uint64_t film_bytes = 0;
        bool film_held = false;          // film allocated and counted in in_use
        int state = 0;                   // 0 queued, 1 running, 2 done, 3 failed
    };
    // Uploaded scenes keyed by the handle's uid; `refs` counts the running requests on a scene.
    // An unused scene stays resident until the budget needs its memory (least recently used first).
    struct CachedScene {
        std::unique_ptr<wf::gpu::DeviceScene> dev;
        int refs = 0;
        uint64_t last_use = 0;
    };
    static constexpr int kChunksInFlight = 2;
    uint64_t budget = 0, in_use = 0;
    wf::gpu::PersistentTuning tune;
    ServerPool mem;                   // declared before everything allocated from it
    hipStream_t stream = nullptr;
    // Finished requests are resolved and downloaded here while the next launch runs on `stream`.
    hipStream_t side = nullptr;
    PinnedStage stage;                // finished films are downloaded through it on `side`
    int pending = -1;                 // snapshot slot of the last launch, not yet collected
    wf::gpu::DevCounters seen{};      // counters as of the last collected snapshot
    wf::gpu::DevCounters* ctr = nullptr;
    PinnedCounters hctr;
    std::map<uint64_t, CachedScene> scenes;
    uint64_t use_clock = 0;
    wf::gpu::PersistentEngine eng;
    std::deque<Req*> queue;
    std::map<int64_t, std::unique_ptr<Req>> reqs;
    std::vector<Req*> running;        // in admission order
    int64_t next_id = 1, submitted = 0, admitted = 0, completed = 0, rejected = 0, failed = 0;
    double busy_ms = 0.0;

    ~wf_server() {
        if (stream) (void)hipStreamSynchronize(stream);
        if (side) (void)hipStreamSynchronize(side);
        for (auto& kv : reqs) kv.second->film.release();
        if (stream) (void)hipStreamSynchronize(stream);
        if (side) (void)hipStreamSynchronize(side);
        if (ctr) (void)hipFree(ctr);
        if (side) (void)hipStreamDestroy(side);
        if (stream) (void)hipStreamDestroy(stream);
    }

    uint64_t scene_bytes() const {
        uint64_t b = 0;
        for (auto& kv : scenes) b += kv.second.dev->bytes;
        return b;
    }

    // Frees unused cached scenes, least recently used first, until `need` more bytes fit the
    // budget; false if they cannot. A scene leaves the engine's scene table before its memory is
    // freed, after the device has finished every launch that may stage it.
    bool make_room(uint64_t need) {
        if (in_use + need <= budget) return true;
        std::vector<std::pair<uint64_t, uint64_t>> idle;   // (last use, uid)
        uint64_t freeable = 0;
        for (auto& kv : scenes)
            if (kv.second.refs == 0) {
                idle.push_back({kv.second.last_use, kv.first});
                freeable += kv.second.dev->bytes;
            }
        if (in_use + need > budget + freeable) return false;
        std::sort(idle.begin(), idle.end());
        bool synced = false;
        for (auto& e : idle) {
            if (in_use + need <= budget) break;
            CachedScene& c = scenes[e.second];
            eng.forget_scene(&c.dev->view);
            if (!synced) {
                wf::gpu::check(hipStreamSynchronize(stream), "server sync");
                wf::gpu::check(hipStreamSynchronize(side), "server sync");
                synced = true;
            }
            in_use -= c.dev->bytes;
            scenes.erase(e.second);
        }
        return in_use + need <= budget;
    }

    Chunk next_chunk(const Req& r) const {
        Chunk c{};
        c.spp = std::min(r.chunk_spp, r.job.spp - r.next_sample);
        c.bytes = (uint64_t)r.px * (uint64_t)c.spp * sizeof(float4);
        return c;
    }

    void admit_chunk(Req& r, Chunk c) {
        wf::gpu::PersistentJob j = r.tmpl;
        j.first_sample = r.next_sample;
        j.spp = c.spp;
        c.seq = eng.admit(j);
        r.chunks.push_back(c);
        r.next_sample += c.spp;
        in_use += c.bytes;
    }

    // Releases what a request holds; it is reported failed (wf_server_poll < 0).
    void fail_request(Req* r) {
        if (r->film_held) {
            r->film.release();
            in_use -= r->film_bytes;
            r->film_held = false;
        }
        r->state = 3;
        failed++;
    }
};

extern "C" {

int wf_server_create(uint64_t budget_bytes, const wf_render_params* tuning, wf_server** out) {
    return guard([&] {
        if (!out) throw std::invalid_argument("null argument");
        auto s = std::make_unique<wf_server>();
        s->tune = to_tuning(tuning);
        // Shorter graph launches than a single render's: a step collects completions once per
        // launch, and 8 pool iterations per launch cut the serving latency (p50 57 -> 52 ms, p99
        // 136 -> 130 ms at 75% load) at the same capacity; 4 cut it further at 3% of capacity.
        if (s->tune.graph_iters <= 0) s->tune.graph_iters = 8;
        // Pool: the largest power of two up to 2^26 slots that takes at most half the budget (the
        // rest is for requests' films and sample buffers), at least 2^20.
        if (s->tune.pool_paths <= 0) {
            uint64_t p = 1ull << 26;
            while (p > (1ull << 20) && p * wf::gpu::PersistentEngine::kBytesPerSlot > budget_bytes / 2) p >>= 1;
            s->tune.pool_paths = (int)p;
        }
        int dev = 0;
        wf::gpu::check(hipGetDevice(&dev), "hipGetDevice");
        s->mem.create(dev);
        wf::gpu::check(hipStreamCreate(&s->stream), "hipStreamCreate");
        wf::gpu::check(hipStreamCreate(&s->side), "hipStreamCreate");
        s->stage.get();   // page-locked once, here, rather than inside the first step
        wf::gpu::check(hipMalloc(&s->ctr, sizeof(wf::gpu::DevCounters)), "hipMalloc counters");
        wf::gpu::check(hipMemsetAsync(s->ctr, 0, sizeof(wf::gpu::DevCounters), s->stream), "hipMemset counters");
        s->eng.open(s->tune, s->ctr, s->stream);
        s->eng.set_mem_pool(s->mem.pool);
        s->budget = budget_bytes;
        s->in_use = s->eng.pool_bytes();
        if (s->in_use > budget_bytes) throw std::invalid_argument("budget smaller than the path pool");
        *out = s.release();
        return (int)WF_OK;
    });
}

void wf_server_destroy(wf_server* s) { delete s; }

int wf_server_submit(wf_server* s, const wf_job* job, int64_t* id) {
    return guard([&] {
        if (!s || !job || !job->scene || !id) throw std::invalid_argument("null argument");
        if (job->spp < 1 || job->max_depth < 0 || job->max_depth > 1024) throw std::invalid_argument("bad job");
        if (job->width < 0 || job->height < 0 || job->width > 8192 || job->height > 8192)
            throw std::invalid_argument("job image size outside 1..8192");
        s->submitted++;
        const wf::Scene& sc = job->scene->s;
        const uint64_t px = (uint64_t)(job->width > 0 ? job->width : sc.width) *
                            (uint64_t)(job->height > 0 ? job->height : sc.height);
        const uint64_t film = px * 5 * sizeof(float);
        const uint64_t scene = s->scenes.count(job->scene->uid) ? 0 : scene_upload_bytes(sc);
        // Samples per chunk: about one pool of paths, fewer if the budget cannot hold one such
        // chunk beside the film (and the scene, if it is not resident); at least one.
        const uint64_t pool = s->eng.pool_bytes();
        const uint64_t free_for_chunk = s->budget > pool + film + scene ? s->budget - pool - film - scene : 0;
        const uint64_t by_budget = free_for_chunk / (px * sizeof(float4));
        const uint64_t by_pool = std::max<uint64_t>(1, (uint64_t)s->eng.pool_paths() / px);
        const uint64_t chunk = std::min<uint64_t>({by_pool, by_budget, (uint64_t)job->spp});
        if (chunk < 1) {
            s->rejected++;
            return fail(WF_ERR_BUDGET, "request can never fit the device-memory budget");
        }
        auto r = std::make_unique<wf_server::Req>();
        r->id = s->next_id++;
        r->job = *job;
        r->scene_uid = job->scene->uid;
        r->px = (int)px;
        r->chunk_spp = (int)chunk;
        r->film_bytes = film;
        *id = r->id;
        s->queue.push_back(r.get());
        s->reqs.emplace(r->id, std::move(r));
        return (int)WF_OK;
    });
}

int wf_server_step(wf_server* s, int launches) {
    return guard([&] {
        if (!s) throw std::invalid_argument("null server");
        auto t0 = std::chrono::steady_clock::now();
        // Next chunks of running requests first (oldest request first, up to kChunksInFlight each),
        // while they fit the budget and the job ring.
        bool full = false;
        for (wf_server::Req* r : s->running) {
            while (!full && r->next_sample < r->job.spp && (int)r->chunks.size() < wf_server::kChunksInFlight) {
                const wf_server::Chunk c = s->next_chunk(*r);
                if (s->in_use + c.bytes > s->budget || !s->eng.can_admit(r->tmpl.scene)) {
                    full = true;
                    break;
                }
                s->admit_chunk(*r, c);
            }
        }
        // Then queued requests in arrival order: the scene (uploaded unless resident), the film and
        // the first chunk must fit; an unused resident scene may be evicted to make room.
        while (!full && !s->queue.empty()) {
            wf_server::Req* r = s->queue.front();
            const auto it = s->scenes.find(r->scene_uid);
            const bool resident = it != s->scenes.end();
            const wf_server::Chunk c = s->next_chunk(*r);
            const uint64_t need = r->film_bytes + c.bytes + (resident ? 0 : scene_upload_bytes(r->job.scene->s));
            if (!s->make_room(need)) break;
            if (resident && !s->eng.can_admit(&it->second.dev->view)) break;
            if (!resident && !s->eng.can_admit(nullptr)) break;   // a new scene needs a free slot
            s->queue.pop_front();
            bool counted = false;   // the request holds a reference to its cached scene
            try {
                wf_server::CachedScene& cs = s->scenes[r->scene_uid];
                if (!cs.dev) {
                    auto d = std::make_unique<wf::gpu::DeviceScene>();
                    d->upload(r->job.scene->s);
                    s->in_use += d->bytes;
                    cs.dev = std::move(d);
                }
                cs.refs += 1;
                counted = true;
                cs.last_use = ++s->use_clock;
                int w = 0, h = 0;
                job_view(r->job, &r->tmpl, &w, &h);
                r->film.alloc(w * h, s->stream, s->mem.pool);
                r->film_held = true;
                s->in_use += r->film_bytes;
                r->tmpl.scene = &cs.dev->view;
                r->tmpl.film = r->film.film;
                r->tmpl.seed = r->job.seed;
                r->tmpl.max_depth = r->job.max_depth;
                s->admit_chunk(*r, c);
            } catch (const std::exception&) {
                // Not admitted after all: give back what it took, report it failed and pass the
                // error on; the queue and the running requests stay consistent.
                auto it2 = s->scenes.find(r->scene_uid);
                if (it2 != s->scenes.end()) {
                    if (counted) it2->second.refs -= 1;
                    if (!it2->second.dev) s->scenes.erase(it2);
                }
                s->fail_request(r);
                throw;
            }
            r->state = 1;
            s->running.push_back(r);
            s->admitted++;
        }
        // Queue this step's launches first, then finish what the previous step's launches
        // completed: the host work below overlaps the launches just queued.
        int launched = -1;
        if (!s->running.empty() && launches > 0) {
            s->eng.run(launches);
            launched = s->eng.snapshot(s->ctr);
        }
        if (s->pending >= 0) {
            const int slot = s->pending;
            s->eng.collect(slot, &s->seen);
            bool waited = false;
            std::vector<wf_server::Req*> still;
            for (wf_server::Req* r : s->running) {
                // Resolve complete chunks in sample order; a chunk completed before an older one
                // waits for it.
                while (!r->chunks.empty() && s->eng.job_complete(r->chunks.front().seq)) {
                    if (!waited) {   // the side stream starts after the snapshot's launches
                        wf::gpu::check(hipStreamWaitEvent(s->side, s->eng.snapshot_event(slot), 0), "hipStreamWaitEvent");
                        waited = true;
                    }
                    const wf_server::Chunk c = r->chunks.front();
                    s->eng.finish_job(c.seq, s->side);
                    r->chunks.pop_front();
                    r->resolved += c.spp;
                    s->in_use -= c.bytes;
                }
                if (r->resolved == r->job.spp) {
                    r->film.download(r->job.spp, r->job.rgb, r->job.lum_var, s->side, s->stage);
                    r->film.owner = s->side;
                    r->film.release();
                    r->film_held = false;
                    s->in_use -= r->film_bytes;
                    r->state = 2;
                    s->completed++;
                    s->scenes[r->scene_uid].refs -= 1;
                } else {
                    still.push_back(r);
                }
            }
            s->running.swap(still);
        }
        s->pending = launched;
        s->busy_ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        return (int)WF_OK;
    });
}

int wf_server_poll(wf_server* s, int64_t id) {
    if (!s) return -1;
    auto it = s->reqs.find(id);
    if (it == s->reqs.end()) return -1;
    if (it->second->state == 3) return -2;
    return it->second->state == 2 ? 1 : 0;
}

int wf_server_stats_get(wf_server* s, wf_server_stats* out) {
    return guard([&] {
        if (!s || !out) throw std::invalid_argument("null argument");
        std::memset(out, 0, sizeof(*out));
        out->submitted = s->submitted;
        out->admitted = s->admitted;
        out->completed = s->completed;
        out->rejected = s->rejected;
        out->in_flight = (int64_t)s->running.size();
        // While requests run, the counters of the last collected snapshot (reading the device
        // counters would wait for the queued launch); when idle, the device counters.
        wf::gpu::DevCounters c = s->seen;
        if (s->running.empty()) c = s->hctr.read(s->ctr, s->stream);
        out->extension_rays = c.extension_rays;
        out->shadow_rays = c.shadow_rays;
        out->bytes_in_use = s->in_use;
        out->budget_bytes = s->budget;
        out->busy_ms = s->busy_ms;
        return (int)WF_OK;
    });
}

}  // extern "C"
