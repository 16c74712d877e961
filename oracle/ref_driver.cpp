// C entry points over the reference's own host code, compiled from its sources where they lie
// (see oracle/Makefile, target `ref`). Used only where those sources exist, to pin the oracle and
// to write golden fixtures; the reference code itself never enters the repository.
#include <cstring>
#include <string>
#include <vector>

#include "wf/bvh.hpp"
#include "wf/film.hpp"
#include "wf/path_core.hpp"
#include "wf/renderer.hpp"
#include "wf/scene.hpp"

using namespace wf;

namespace {
int ib(float f) { int v; std::memcpy(&v, &f, 4); return v; }
float fb(int v) { float f; std::memcpy(&f, &v, 4); return f; }
}

extern "C" {

void* ref_make_scene(const char* name, int w, int h) {
    try {
        return new Scene(make_scene(name, w, h));
    } catch (...) {
        return nullptr;
    }
}

// Builds a reference Scene from arrays in the product's C-ABI layout and finalizes it.
void* ref_scene_from_arrays(const float* spheres, int ns, const float* vertices, int nv, const int* tris, int nt,
                            const float* mats, int nm, const float* env, const float* cam, int w, int h) {
    Scene* s = new Scene();
    s->width = w;
    s->height = h;
    for (int i = 0; i < ns; ++i) {
        Sphere sp{};
        sp.center = Vec3{spheres[8 * i], spheres[8 * i + 1], spheres[8 * i + 2]};
        sp.radius = spheres[8 * i + 3];
        sp.material = ib(spheres[8 * i + 4]);
        s->spheres.push_back(sp);
    }
    for (int i = 0; i < nv; ++i) s->vertices.push_back(Vec3{vertices[3 * i], vertices[3 * i + 1], vertices[3 * i + 2]});
    for (int i = 0; i < nt; ++i) s->triangles.push_back(Triangle{tris[4 * i], tris[4 * i + 1], tris[4 * i + 2], tris[4 * i + 3]});
    for (int i = 0; i < nm; ++i) {
        Material m{};
        m.type = ib(mats[12 * i]);
        m.ior = mats[12 * i + 1];
        m.albedo = Vec3{mats[12 * i + 4], mats[12 * i + 5], mats[12 * i + 6]};
        m.emission = Vec3{mats[12 * i + 8], mats[12 * i + 9], mats[12 * i + 10]};
        s->materials.push_back(m);
    }
    s->env = Vec3{env[0], env[1], env[2]};
    s->camera.position = Vec3{cam[0], cam[1], cam[2]};
    s->camera.look_at = Vec3{cam[3], cam[4], cam[5]};
    s->camera.up = Vec3{cam[6], cam[7], cam[8]};
    s->camera.vfov_deg = cam[9];
    try {
        finalize_scene(*s);
    } catch (...) {
        delete s;
        return nullptr;
    }
    return s;
}

void ref_scene_free(void* s) { delete (Scene*)s; }

void ref_scene_counts(void* h, int* out) {
    Scene* s = (Scene*)h;
    out[0] = (int)s->spheres.size();
    out[1] = (int)s->vertices.size();
    out[2] = (int)s->triangles.size();
    out[3] = (int)s->materials.size();
    out[4] = (int)s->nodes.size();
    out[5] = (int)s->lights.size();
}

void ref_scene_export(void* h, float* spheres, float* vertices, int* tris, float* mats, float* env, float* cam) {
    Scene* s = (Scene*)h;
    for (size_t i = 0; i < s->spheres.size(); ++i) {
        float* p = spheres + 8 * i;
        p[0] = s->spheres[i].center.x; p[1] = s->spheres[i].center.y; p[2] = s->spheres[i].center.z;
        p[3] = s->spheres[i].radius; p[4] = fb(s->spheres[i].material); p[5] = p[6] = p[7] = 0.0f;
    }
    for (size_t i = 0; i < s->vertices.size(); ++i) {
        vertices[3 * i] = s->vertices[i].x; vertices[3 * i + 1] = s->vertices[i].y; vertices[3 * i + 2] = s->vertices[i].z;
    }
    for (size_t i = 0; i < s->triangles.size(); ++i) {
        tris[4 * i] = s->triangles[i].i0; tris[4 * i + 1] = s->triangles[i].i1;
        tris[4 * i + 2] = s->triangles[i].i2; tris[4 * i + 3] = s->triangles[i].material;
    }
    for (size_t i = 0; i < s->materials.size(); ++i) {
        float* p = mats + 12 * i;
        const Material& m = s->materials[i];
        p[0] = fb(m.type); p[1] = m.ior; p[2] = 0.0f; p[3] = 0.0f;
        p[4] = m.albedo.x; p[5] = m.albedo.y; p[6] = m.albedo.z; p[7] = 0.0f;
        p[8] = m.emission.x; p[9] = m.emission.y; p[10] = m.emission.z; p[11] = 0.0f;
    }
    std::memset(env, 0, 20 * sizeof(float));
    env[0] = s->env.x; env[1] = s->env.y; env[2] = s->env.z;
    const Camera& c = s->camera;
    float v[10] = {c.position.x, c.position.y, c.position.z, c.look_at.x, c.look_at.y, c.look_at.z,
                   c.up.x, c.up.y, c.up.z, c.vfov_deg};
    std::memcpy(cam, v, sizeof(v));
}

void ref_scene_bvh(void* h, int* nodes, int* order) {
    Scene* s = (Scene*)h;
    for (size_t i = 0; i < s->nodes.size(); ++i) {
        const BvhNode& n = s->nodes[i];
        int* p = nodes + 12 * i;
        p[0] = ib(n.lo.x); p[1] = ib(n.lo.y); p[2] = ib(n.lo.z); p[3] = n.skip;
        p[4] = ib(n.hi.x); p[5] = ib(n.hi.y); p[6] = ib(n.hi.z);
        p[7] = n.prim_start; p[8] = n.prim_count; p[9] = p[10] = p[11] = 0;
    }
    std::memcpy(order, s->prim_order.data(), s->prim_order.size() * sizeof(int));
}

unsigned long long ref_scene_hash(void* h) { return scene_hash(*(Scene*)h); }

void ref_camera_rig(void* h, float* r) {
    Scene* s = (Scene*)h;
    CameraRig g = make_camera_rig(s->camera, s->width, s->height);
    float v[12] = {g.origin.x, g.origin.y, g.origin.z, g.lower_left.x, g.lower_left.y, g.lower_left.z,
                   g.horizontal.x, g.horizontal.y, g.horizontal.z, g.vertical.x, g.vertical.y, g.vertical.z};
    std::memcpy(r, v, sizeof(v));
}

unsigned int ref_hash_u32(unsigned int x) { return hash_u32(x); }
void ref_path_key(unsigned int seed, unsigned int pixel, unsigned int sample, unsigned int* k) {
    PathKey p = make_path_key(seed, pixel, sample);
    k[0] = p.k0;
    k[1] = p.k1;
}
float ref_rng(unsigned int k0, unsigned int k1, unsigned int depth, unsigned int dim) {
    return rng(PathKey{k0, k1}, depth, dim);
}

void ref_intersect(void* h, const float* rays, int n, int anyhit, float* t, int* prim) {
    Scene* s = (Scene*)h;
    SceneView v = s->view();
    for (int i = 0; i < n; ++i) {
        const float* q = rays + 8 * i;
        Ray r;
        r.o = Vec3{q[0], q[1], q[2]};
        r.tmin = q[3];
        r.d = Vec3{q[4], q[5], q[6]};
        r.tmax = q[7];
        HitRecord hr;
        intersect(v, r, &hr, anyhit != 0);
        t[i] = hr.t;
        prim[i] = hr.prim;
    }
}

void ref_trace_path(void* h, int px, int py, unsigned int sample, unsigned int seed, int max_depth, float* L) {
    Scene* s = (Scene*)h;
    SceneView v = s->view();
    Vec3 r = trace_path(v, px, py, sample, seed, max_depth);
    L[0] = r.x;
    L[1] = r.y;
    L[2] = r.z;
}

// render_cpu (mean and the reference CPU's float unbiased variance).
void ref_render_cpu(void* h, int spp, unsigned int seed, int max_depth, int threads, float* rgb, float* lum_var) {
    Scene* s = (Scene*)h;
    RenderParams p;
    p.spp = spp;
    p.seed = seed;
    p.max_depth = max_depth;
    p.threads = threads;
    RenderStats st;
    RenderResult r = render_cpu(*s, p, &st);
    std::memcpy(rgb, r.rgb.data(), r.rgb.size() * sizeof(float));
    std::memcpy(lum_var, r.lum_var.data(), r.lum_var.size() * sizeof(float));
}

}  // extern "C"
