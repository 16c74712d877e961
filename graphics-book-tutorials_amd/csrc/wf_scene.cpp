// Host scene construction, validation, lights, BVH and device layout. See wf_scene.h.
#include "wf_scene.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <stdexcept>

namespace wf {

CameraRig make_camera_rig(const Camera& cam, int width, int height) {
    // Pinhole rig (include/wf/scene.hpp:129-141), evaluated on the host for every arm.
    float aspect = (float)width / (float)height;
    float h = tanf(cam.vfov_deg * 0.5f * kPi / 180.0f);
    Vec3 w = normalize(cam.position - cam.look_at);
    Vec3 u = normalize(cross(cam.up, w));
    Vec3 v = cross(w, u);
    CameraRig r;
    r.origin = cam.position;
    r.horizontal = u * (2.0f * h * aspect);
    r.vertical = v * (2.0f * h);
    r.lower_left = r.origin - r.horizontal * 0.5f - r.vertical * 0.5f - w;
    return r;
}

// ---------------------------------------------------------------------------------------------
// validation and finalize (src/scene.cpp:10-104)
// ---------------------------------------------------------------------------------------------
std::string validate_scene(const Scene& s) {
    if (s.spheres.empty() && s.triangles.empty()) return "no primitives";
    if (s.materials.empty()) return "no materials";
    const int nm = (int)s.materials.size(), nv = (int)s.vertices.size();
    for (const Sphere& sp : s.spheres) {
        if (sp.material < 0 || sp.material >= nm) return "sphere material index out of range";
        if (sp.radius <= 0.0f || !is_finite1(sp.radius)) return "sphere has invalid radius";
        if (!is_finite3(sp.center)) return "sphere centre is non-finite";
    }
    for (const Triangle& t : s.triangles) {
        if (t.material < 0 || t.material >= nm) return "triangle material index out of range";
        if (t.i0 < 0 || t.i0 >= nv || t.i1 < 0 || t.i1 >= nv || t.i2 < 0 || t.i2 >= nv)
            return "triangle vertex index out of range";
    }
    for (const Vec3& v : s.vertices)
        if (!is_finite3(v)) return "vertex is non-finite";
    if (s.width < 1 || s.height < 1) return "image size must be positive";
    return "";
}

static bool emits(const Material& m) {
    return m.emission.x > 0.0f || m.emission.y > 0.0f || m.emission.z > 0.0f;
}

void finalize_scene(Scene& s) {
    s.nodes.clear();
    s.lights.clear();
    s.prim_order.clear();
    s.dnodes.clear();
    s.dprims.clear();
    std::string err = validate_scene(s);
    if (!err.empty()) throw std::runtime_error(err);
    // Lights: emissive spheres first, then emissive triangles, in id order.
    const int ns = (int)s.spheres.size();
    for (int i = 0; i < ns; ++i) {
        const Sphere& sp = s.spheres[i];
        if (emits(s.materials[sp.material])) {
            float area = 4.0f * 3.14159265358979323846f * sp.radius * sp.radius;
            s.lights.push_back(Light{-1, area, i, 0});
        }
    }
    for (int i = 0; i < (int)s.triangles.size(); ++i) {
        const Triangle& t = s.triangles[i];
        if (emits(s.materials[t.material])) {
            Vec3 e1 = s.vertices[t.i1] - s.vertices[t.i0];
            Vec3 e2 = s.vertices[t.i2] - s.vertices[t.i0];
            float area = 0.5f * length(cross(e1, e2));
            s.lights.push_back(Light{-1, area, ns + i, 0});
        }
    }
    try {
        build_bvh(s);
        build_device_layout(s);
    } catch (...) {
        s.nodes.clear();
        s.lights.clear();
        s.prim_order.clear();
        throw;
    }
}

// ---------------------------------------------------------------------------------------------
// BVH (src/bvh_build.cpp:60-251). The reference recurses on vectors of primitive ids; here one id
// array is partitioned in place with a stable partition, which visits primitives in the same
// order, and nodes are emitted directly in preorder (creation order equals preorder because the
// tree is built depth first, left first), so the skip of a left child is its right sibling and a
// right child inherits its parent's skip (src/bvh_build.cpp:191-226).
// ---------------------------------------------------------------------------------------------
namespace {

struct Box {
    Vec3 lo{kInf, kInf, kInf}, hi{-kInf, -kInf, -kInf};
    void grow(Vec3 p) {
        lo.x = min2(lo.x, p.x); lo.y = min2(lo.y, p.y); lo.z = min2(lo.z, p.z);
        hi.x = max2(hi.x, p.x); hi.y = max2(hi.y, p.y); hi.z = max2(hi.z, p.z);
    }
    void grow(const Box& b) {
        lo.x = min2(lo.x, b.lo.x); lo.y = min2(lo.y, b.lo.y); lo.z = min2(lo.z, b.lo.z);
        hi.x = max2(hi.x, b.hi.x); hi.y = max2(hi.y, b.hi.y); hi.z = max2(hi.z, b.hi.z);
    }
    float area() const {
        Vec3 d = hi - lo;
        return 2.0f * (d.x * d.y + d.y * d.z + d.z * d.x);
    }
};

struct Builder {
    std::vector<Box> pbox;
    std::vector<Vec3> pcen;
    std::vector<int> ids, tmp;
    std::vector<BvhNode>* out;
    std::vector<int>* order;

    static float axis_of(Vec3 v, int a) { return a == 0 ? v.x : (a == 1 ? v.y : v.z); }

    static int bin_of(float c, float amin, float range) {
        int b = (int)((c - amin) / range * (8 - 0.0001f));
        return b < 0 ? 0 : (b >= 8 ? 7 : b);
    }

    // Builds the subtree over ids[first, first+count) with the given skip; returns its root index.
    int build(int first, int count, int skip) {
        const int idx = (int)out->size();
        out->push_back(BvhNode{});
        Box bounds, cb;
        for (int k = first; k < first + count; ++k) {
            bounds.grow(pbox[ids[k]]);
            cb.grow(pcen[ids[k]]);
        }
        {
            BvhNode& n = (*out)[idx];
            n.lo = bounds.lo;
            n.hi = bounds.hi;
            n.skip = skip;
            n.pad[0] = n.pad[1] = n.pad[2] = 0;
        }
        if (count <= 4) {
            BvhNode& n = (*out)[idx];
            n.prim_start = (int)order->size();
            n.prim_count = count;
            for (int k = first; k < first + count; ++k) order->push_back(ids[k]);
            return idx;
        }
        (*out)[idx].prim_start = -1;
        (*out)[idx].prim_count = 0;

        Vec3 ext = cb.hi - cb.lo;
        int axis = 0;
        if (ext.y > ext.x) axis = 1;
        if (ext.z > ext.x && ext.z > ext.y) axis = 2;

        int n_left;
        if (ext.x < 1e-6f && ext.y < 1e-6f && ext.z < 1e-6f) {
            n_left = count / 2;   // identical centroids: split by position
        } else {
            const float amin = axis_of(cb.lo, axis), amax = axis_of(cb.hi, axis);
            const float range = amax - amin;
            Box bins[8];
            int cnt[8] = {0, 0, 0, 0, 0, 0, 0, 0};
            for (int k = first; k < first + count; ++k) {
                int b = bin_of(axis_of(pcen[ids[k]], axis), amin, range);
                bins[b].grow(pbox[ids[k]]);
                cnt[b]++;
            }
            float best_cost = 1e9f;
            int best = 1;
            const float parent_area = bounds.area();
            for (int split = 1; split < 8; ++split) {
                Box lb, rb;
                int lc = 0, rc = 0;
                for (int i = 0; i < split; ++i) { lb.grow(bins[i]); lc += cnt[i]; }
                for (int i = split; i < 8; ++i) { rb.grow(bins[i]); rc += cnt[i]; }
                float cost = (lc * lb.area() + rc * rb.area()) / parent_area;
                if (cost < best_cost) { best_cost = cost; best = split; }
            }
            // Stable partition: left keeps input order, then right keeps input order.
            int nl = 0, nr = 0;
            for (int k = first; k < first + count; ++k) {
                int id = ids[k];
                if (bin_of(axis_of(pcen[id], axis), amin, range) < best) ids[first + nl++] = id;
                else tmp[nr++] = id;
            }
            std::memcpy(&ids[first + nl], tmp.data(), sizeof(int) * (size_t)nr);
            n_left = nl;
            if (nl == 0 || nr == 0) n_left = count / 2;   // empty side: split by position
        }
        // Right child index = idx + 1 + size(left subtree); it is only known after the left
        // subtree is built, so the left child's skip is patched afterwards.
        const int left = build(first, n_left, -2);
        const int right_idx = (int)out->size();
        patch_skip(left, right_idx);
        build(first + n_left, count - n_left, skip);
        return idx;
    }

    // The left child was built with placeholder skip -2; its real skip is its right sibling.
    // Only the left child itself carries the placeholder (its own right descendants inherit
    // the placeholder too), so walk the right spine of the left subtree.
    void patch_skip(int node, int value) {
        for (;;) {
            BvhNode& n = (*out)[node];
            n.skip = value;
            if (n.prim_count > 0) return;
            node = (*out)[node + 1].skip;   // right child of this internal node
        }
    }
};

}  // namespace

void build_bvh(Scene& s) {
    if (s.spheres.empty() && s.triangles.empty()) throw std::runtime_error("Cannot build BVH: no primitives");
    const int ns = (int)s.spheres.size(), np = s.n_prims();
    Builder b;
    b.pbox.resize(np);
    b.pcen.resize(np);
    for (int p = 0; p < np; ++p) {
        if (p < ns) {
            const Sphere& sp = s.spheres[p];
            Vec3 r{sp.radius, sp.radius, sp.radius};
            b.pbox[p].lo = sp.center - r;
            b.pbox[p].hi = sp.center + r;
            b.pcen[p] = sp.center;
        } else {
            const Triangle& t = s.triangles[p - ns];
            Box bb;
            bb.lo = bb.hi = s.vertices[t.i0];
            bb.grow(s.vertices[t.i1]);
            bb.grow(s.vertices[t.i2]);
            b.pbox[p] = bb;
            Vec3 c = s.vertices[t.i0] + s.vertices[t.i1] + s.vertices[t.i2];
            b.pcen[p] = c * (1.0f / 3.0f);
        }
    }
    b.ids.resize(np);
    for (int p = 0; p < np; ++p) b.ids[p] = p;
    b.tmp.resize(np);
    s.nodes.clear();
    s.prim_order.clear();
    s.nodes.reserve((size_t)np);
    s.prim_order.reserve((size_t)np);
    b.out = &s.nodes;
    b.order = &s.prim_order;
    b.build(0, np, -1);
}

std::string validate_bvh(const Scene& s) {
    const int np = s.n_prims(), nn = (int)s.nodes.size();
    std::vector<int> seen(np, 0);
    for (int p : s.prim_order) {
        if (p < 0 || p >= np) return "primitive index out of range";
        seen[p]++;
    }
    if ((int)s.prim_order.size() != np) return "not all primitives in BVH";
    for (int c : seen)
        if (c != 1) return "primitive appears multiple times or not at all";
    auto contains = [](Vec3 plo, Vec3 phi, Vec3 clo, Vec3 chi) {
        const float e = 1e-5f;
        return plo.x - e <= clo.x && chi.x <= phi.x + e && plo.y - e <= clo.y && chi.y <= phi.y + e &&
               plo.z - e <= clo.z && chi.z <= phi.z + e;
    };
    const int ns = (int)s.spheres.size();
    for (int i = 0; i < nn; ++i) {
        const BvhNode& n = s.nodes[i];
        if (n.skip < -1 || n.skip >= nn) return "skip link out of bounds";
        if (n.skip != -1 && n.skip <= i) return "skip link does not point forward";
        if (n.prim_count > 0) {
            if (n.prim_count > 4) return "leaf has more than 4 primitives";
            if (n.prim_start < 0 || n.prim_start + n.prim_count > (int)s.prim_order.size())
                return "leaf prim_start/count out of bounds";
            for (int k = 0; k < n.prim_count; ++k) {
                int p = s.prim_order[n.prim_start + k];
                Vec3 lo, hi;
                if (p < ns) {
                    const Sphere& sp = s.spheres[p];
                    Vec3 r{sp.radius, sp.radius, sp.radius};
                    lo = sp.center - r;
                    hi = sp.center + r;
                } else {
                    const Triangle& t = s.triangles[p - ns];
                    Box bb;
                    bb.lo = bb.hi = s.vertices[t.i0];
                    bb.grow(s.vertices[t.i1]);
                    bb.grow(s.vertices[t.i2]);
                    lo = bb.lo;
                    hi = bb.hi;
                }
                if (!contains(n.lo, n.hi, lo, hi)) return "leaf box does not contain its primitive";
            }
        } else {
            if (n.prim_start != -1) return "internal node has prim_start != -1";
            if (i + 1 >= nn) return "left child (N+1) out of bounds";
            int r = s.nodes[i + 1].skip;
            if (r <= i + 1 || r >= nn) return "right child index invalid";
            if (s.nodes[r].skip != n.skip) return "right child does not share parent's skip";
            if (!contains(n.lo, n.hi, s.nodes[i + 1].lo, s.nodes[i + 1].hi)) return "parent box does not contain left child";
            if (!contains(n.lo, n.hi, s.nodes[r].lo, s.nodes[r].hi)) return "parent box does not contain right child";
        }
    }
    if (nn > 0 && s.nodes.back().skip != -1) return "last node skip is not -1";
    return "";
}

void build_device_layout(Scene& s) {
    const int nn = (int)s.nodes.size(), np = s.n_prims(), ns = (int)s.spheres.size();
    // Breadth-first order of the logical tree for the first kTopFirstNodes nodes.
    std::vector<int> newidx(nn, -1);
    int next = 0;
    {
        std::vector<int> queue;
        queue.reserve(std::min(nn, kTopFirstNodes) + 2);
        queue.push_back(0);
        size_t head = 0;
        while (head < queue.size() && next < kTopFirstNodes) {
            int n = queue[head++];
            newidx[n] = next++;
            if (s.nodes[n].prim_count == 0) {
                queue.push_back(n + 1);
                queue.push_back(s.nodes[n + 1].skip);
            }
        }
    }
    for (int n = 0; n < nn; ++n)
        if (newidx[n] < 0) newidx[n] = next++;
    s.dnodes.assign(nn, Node{});
    for (int n = 0; n < nn; ++n) {
        const BvhNode& src = s.nodes[n];
        Node d;
        d.lo = src.lo;
        d.hi = src.hi;
        d.skip = src.skip >= 0 ? newidx[src.skip] : -1;
        if (src.prim_count > 0) {
            if (src.prim_start >= (1 << 28)) throw std::runtime_error("too many primitives for the node encoding");
            int code = (src.prim_start << 2) | (src.prim_count - 1);
            d.child = -(code + 1);
        } else {
            d.child = newidx[n + 1];
        }
        s.dnodes[newidx[n]] = d;
    }
    // Device traversal follows these links without range checks: every skip is -1 or a node, every
    // child a node or the code of a leaf whose primitives exist, and every box has lo <= hi.
    for (const Node& d : s.dnodes) {
        bool ok = d.skip >= -1 && d.skip < nn && d.lo.x <= d.hi.x && d.lo.y <= d.hi.y && d.lo.z <= d.hi.z;
        if (d.child >= 0) {
            ok = ok && d.child < nn;
        } else {
            const int code = -(d.child + 1);
            ok = ok && (code >> 2) + (code & 3) + 1 <= np;
        }
        if (!ok) throw std::runtime_error("device BVH layout has an invalid link or box");
    }
    s.dprims.assign(np, Prim{});
    std::vector<int> slot_of(np, -1);
    for (int k = 0; k < np; ++k) {
        const int id = s.prim_order[k];
        slot_of[id] = k;
        Prim p{};
        p.id = id;
        if (id < ns) {
            const Sphere& sp = s.spheres[id];
            p.p0 = sp.center;
            p.radius = sp.radius;
            p.material = sp.material;
        } else {
            const Triangle& t = s.triangles[id - ns];
            p.p0 = s.vertices[t.i0];
            p.p1 = s.vertices[t.i1];
            p.p2 = s.vertices[t.i2];
            p.material = t.material;
        }
        s.dprims[k] = p;
    }
    for (Light& l : s.lights) l.slot = slot_of[l.id];
}

SceneView host_view(const Scene& s) {
    SceneView v{};
    v.nodes = s.dnodes.data();
    v.prims = s.dprims.data();
    v.materials = s.materials.data();
    v.lights = s.lights.empty() ? nullptr : s.lights.data();
    v.n_nodes = (int)s.dnodes.size();
    v.n_prims = (int)s.dprims.size();
    v.n_spheres = (int)s.spheres.size();
    v.n_materials = (int)s.materials.size();
    v.n_lights = (int)s.lights.size();
    v.n_sel = v.n_lights + (s.env.sky_light ? 1 : 0) + (s.env.sun_light ? 1 : 0);
    v.sky_pdf = v.n_sel > 0 ? kInv4Pi / (float)v.n_sel : 0.0f;
    v.env = s.env;
    v.camera = make_camera_rig(s.camera, s.width, s.height);
    v.width = s.width;
    v.height = s.height;
    return v;
}

BvhStats bvh_stats(const Scene& s) {
    BvhStats st;
    if (s.nodes.empty()) return st;
    auto area = [](Vec3 lo, Vec3 hi) {
        float dx = hi.x - lo.x, dy = hi.y - lo.y, dz = hi.z - lo.z;
        return 2.0f * (dx * dy + dy * dz + dz * dx);
    };
    const double root = area(s.nodes[0].lo, s.nodes[0].hi);
    std::vector<std::pair<int, int>> stack{{0, 0}};
    while (!stack.empty()) {
        auto [i, d] = stack.back();
        stack.pop_back();
        const BvhNode& n = s.nodes[i];
        double a = area(n.lo, n.hi) / root;
        st.nodes++;
        st.max_depth = std::max(st.max_depth, d);
        if (n.prim_count > 0) {
            st.leaves++;
            st.sah_cost += a * n.prim_count;
        } else {
            st.internal++;
            st.sah_cost += a;
            stack.push_back({i + 1, d + 1});
            stack.push_back({s.nodes[i + 1].skip, d + 1});
        }
    }
    return st;
}

// ---------------------------------------------------------------------------------------------
// scene hash (src/scene.cpp:107-160)
// ---------------------------------------------------------------------------------------------
namespace {
struct Fnv {
    uint64_t h = 14695981039346656037ULL;
    void put(const void* p, size_t n) {
        const uint8_t* b = (const uint8_t*)p;
        for (size_t i = 0; i < n; ++i) {
            h ^= b[i];
            h *= 1099511628211ULL;
        }
    }
    template <class T> void put(const T& v) { put(&v, sizeof(T)); }
};
}  // namespace

uint64_t scene_hash(const Scene& s) {
    Fnv f;
    for (const Sphere& sp : s.spheres) { f.put(sp.center); f.put(sp.radius); f.put(sp.material); }
    for (const Vec3& v : s.vertices) f.put(v);
    for (const Triangle& t : s.triangles) { f.put(t.i0); f.put(t.i1); f.put(t.i2); f.put(t.material); }
    for (const Material& m : s.materials) { f.put(m.type); f.put(m.ior); f.put(m.albedo); f.put(m.emission); }
    f.put(s.env.env);
    f.put(s.camera.position);
    f.put(s.camera.look_at);
    f.put(s.camera.up);
    f.put(s.camera.vfov_deg);
    f.put(s.width);
    f.put(s.height);
    bool ext = s.env.sky_mode != 0 || s.env.sky_light || s.env.sun_light;
    for (const Material& m : s.materials) ext = ext || m.fuzz != 0.0f;
    if (ext) {
        for (const Material& m : s.materials) f.put(m.fuzz);
        f.put(s.env.sky_mode); f.put(s.env.sky_lo); f.put(s.env.sky_hi); f.put(s.env.sky_light);
        f.put(s.env.sun_light); f.put(s.env.sun_dir); f.put(s.env.sun_E);
    }
    return f.h;
}

// ---------------------------------------------------------------------------------------------
// procedural scenes
// ---------------------------------------------------------------------------------------------
namespace {

Material mat(int type, Vec3 albedo, Vec3 emission = Vec3{0, 0, 0}, float ior = 0.0f, float fuzz = 0.0f) {
    Material m{};
    m.type = type;
    m.ior = ior;
    m.fuzz = fuzz;
    m.albedo = albedo;
    m.emission = emission;
    return m;
}

Sphere sphere(Vec3 c, float r, int m) {
    Sphere s{};
    s.center = c;
    s.radius = r;
    s.material = m;
    return s;
}

void set_camera(Scene& s, Vec3 pos, Vec3 at, float vfov) {
    s.camera.position = pos;
    s.camera.look_at = at;
    s.camera.up = Vec3{0.0f, 1.0f, 0.0f};
    s.camera.vfov_deg = vfov;
}

void constant_env(Scene& s, Vec3 env) {
    s.env = EnvParams{};
    s.env.env = env;
}

// The reference's "cornell" (src/scenes.cpp:30-127): two coloured spheres in a triangle box with a
// ceiling light of the Lambert type and emission 5.
void scene_ref_cornell(Scene& s) {
    const Vec3 g7{0.7f, 0.7f, 0.7f};
    s.materials = {mat(kLambert, g7), mat(kLambert, {0.7f, 0.2f, 0.2f}), mat(kLambert, {0.2f, 0.7f, 0.2f}),
                   mat(kLambert, g7, {5.0f, 5.0f, 5.0f})};
    s.spheres = {sphere({-1.5f, 0.0f, -2.0f}, 0.3f, 1), sphere({1.5f, 0.0f, -2.0f}, 0.3f, 2)};
    s.vertices = {{-3.0f, -2.0f, -4.0f}, {3.0f, -2.0f, -4.0f}, {3.0f, 2.0f, -4.0f}, {-3.0f, 2.0f, -4.0f},
                  {-3.0f, -2.0f, 0.0f},  {3.0f, -2.0f, 0.0f},  {-3.0f, 2.0f, 0.0f}, {3.0f, 2.0f, 0.0f},
                  {-1.0f, 1.99f, -2.5f}, {1.0f, 1.99f, -2.5f}, {1.0f, 1.99f, -1.5f}, {-1.0f, 1.99f, -1.5f}};
    s.triangles = {{0, 1, 2, 0}, {0, 2, 3, 0}, {0, 4, 5, 0}, {0, 5, 1, 0},
                   {3, 2, 7, 0}, {3, 7, 6, 0}, {8, 9, 10, 3}, {8, 10, 11, 3}};
    set_camera(s, {0.0f, 0.0f, 2.5f}, {0.0f, 0.0f, -2.0f}, 40.0f);
    constant_env(s, {0.0f, 0.0f, 0.0f});
}

// The reference's "divergence" (src/scenes.cpp:129-229): 30x30 checkerboard of small spheres,
// Lambert / metal / dielectric, a small emissive sphere and a dim constant environment.
void scene_ref_divergence(Scene& s) {
    s.materials = {mat(kLambert, {0.5f, 0.5f, 0.5f}), mat(kLambert, {0.7f, 0.7f, 0.7f}),
                   mat(kLambert, {0.7f, 0.3f, 0.3f}), mat(kLambert, {0.3f, 0.3f, 0.7f}),
                   mat(kMetal, {0.9f, 0.9f, 0.9f}),   mat(kDielectric, {1.0f, 1.0f, 1.0f}, {0, 0, 0}, 1.5f),
                   mat(kEmissive, {0.0f, 0.0f, 0.0f}, {1.0f, 1.0f, 1.0f})};
    s.vertices = {{-15.0f, -1.0f, -15.0f}, {15.0f, -1.0f, -15.0f}, {15.0f, -1.0f, 15.0f}, {-15.0f, -1.0f, 15.0f}};
    s.triangles = {{0, 1, 2, 0}, {0, 2, 3, 0}};
    const float spacing = 0.33f, radius = 0.16f;
    for (int i = 0; i < 30; ++i)
        for (int j = 0; j < 30; ++j) {
            int m = ((i + j) & 1) == 0 ? (i + j) % 4 : ((i & 1) ? 5 : 4);
            s.spheres.push_back(sphere({-4.95f + i * spacing, 0.2f, -4.95f + j * spacing}, radius, m));
        }
    s.spheres.push_back(sphere({0.0f, 5.0f, 0.0f}, 0.2f, 6));
    set_camera(s, {0.0f, 2.0f, 6.0f}, {0.0f, 0.0f, 0.0f}, 45.0f);
    constant_env(s, {0.05f, 0.05f, 0.05f});
}

// The reference's "trianglehv" (src/scenes.cpp:231-313): 241x241-vertex sin*cos height field,
// two metal spheres and an emissive sphere.
void scene_ref_trianglehv(Scene& s) {
    s.materials = {mat(kLambert, {0.7f, 0.7f, 0.7f}), mat(kMetal, {0.9f, 0.9f, 0.9f}),
                   mat(kEmissive, {0.0f, 0.0f, 0.0f}, {3.0f, 3.0f, 3.0f})};
    const int g = 240;
    const float scale = 10.0f, hscale = 1.0f;
    s.vertices.reserve((size_t)(g + 1) * (g + 1));
    for (int i = 0; i <= g; ++i)
        for (int j = 0; j <= g; ++j) {
            float x = (i - g / 2) * scale / (g / 2);
            float z = (j - g / 2) * scale / (g / 2);
            float h = hscale * (sinf(x) * cosf(z) + 0.5f * (x * x + z * z) / 100.0f);
            s.vertices.push_back(Vec3{x, h, z});
        }
    s.triangles.reserve((size_t)2 * g * g);
    for (int i = 0; i < g; ++i)
        for (int j = 0; j < g; ++j) {
            int v0 = i * (g + 1) + j, v1 = v0 + 1, v2 = v0 + (g + 1), v3 = v2 + 1;
            s.triangles.push_back({v0, v1, v2, 0});
            s.triangles.push_back({v1, v3, v2, 0});
        }
    s.spheres = {sphere({0.0f, 2.0f, 0.0f}, 0.5f, 1), sphere({3.0f, 1.5f, -2.0f}, 0.4f, 1),
                 sphere({0.0f, 5.0f, 0.0f}, 0.3f, 2)};
    set_camera(s, {5.0f, 3.0f, 5.0f}, {0.0f, 1.0f, 0.0f}, 50.0f);
    constant_env(s, {0.05f, 0.05f, 0.1f});
}

// The reference's white furnace (src/scenes.cpp:315-344).
void scene_ref_furnace(Scene& s) {
    s.materials = {mat(kLambert, {0.5f, 0.5f, 0.5f}, {1.0f, 1.0f, 1.0f})};
    s.spheres = {sphere({0.0f, 0.0f, 0.0f}, 5.0f, 0)};
    set_camera(s, {0.0f, 0.0f, 0.0f}, {0.0f, 0.0f, -1.0f}, 90.0f);
    constant_env(s, {0.0f, 0.0f, 0.0f});
}

// Indexed quad a-b-c-d as triangles (a,b,c), (a,c,d).
void quad(Scene& s, Vec3 a, Vec3 b, Vec3 c, Vec3 d, int m) {
    int base = (int)s.vertices.size();
    s.vertices.insert(s.vertices.end(), {a, b, c, d});
    s.triangles.push_back({base, base + 1, base + 2, m});
    s.triangles.push_back({base, base + 2, base + 3, m});
}

// Axis-aligned box [0,size] rotated about +y by deg, then translated: 8 vertices, 12 triangles.
void rotated_box(Scene& s, Vec3 size, double deg, Vec3 offset, int m) {
    const double th = deg * 3.14159265358979323846 / 180.0, c = std::cos(th), sn = std::sin(th);
    int base = (int)s.vertices.size();
    for (int k = 0; k < 8; ++k) {
        double x = (k & 1) ? size.x : 0.0, y = (k & 2) ? size.y : 0.0, z = (k & 4) ? size.z : 0.0;
        double rx = c * x + sn * z, rz = -sn * x + c * z;
        s.vertices.push_back(Vec3{(float)(rx + offset.x), (float)(y + offset.y), (float)(rz + offset.z)});
    }
    // Faces by corner bit index (x=1, y=2, z=4), each as a quad.
    const int f[6][4] = {{0, 2, 3, 1}, {4, 5, 7, 6}, {0, 1, 5, 4}, {2, 6, 7, 3}, {0, 4, 6, 2}, {1, 3, 7, 5}};
    for (auto& q : f) {
        s.triangles.push_back({base + q[0], base + q[1], base + q[2], m});
        s.triangles.push_back({base + q[0], base + q[2], base + q[3], m});
    }
}

// Classic 555-unit Cornell box: five walls, a 130x105 ceiling light with Le 15 and two rotated
// boxes; white 0.73, red (0.65,0.05,0.05) at x=555, green (0.12,0.45,0.15) at x=0.
void scene_cornell555(Scene& s) {
    s.materials = {mat(kLambert, {0.73f, 0.73f, 0.73f}), mat(kLambert, {0.65f, 0.05f, 0.05f}),
                   mat(kLambert, {0.12f, 0.45f, 0.15f}), mat(kEmissive, {0, 0, 0}, {15.0f, 15.0f, 15.0f})};
    const float L = 555.0f;
    quad(s, {0, 0, 0}, {0, L, 0}, {0, L, L}, {0, 0, L}, 2);     // green, x = 0
    quad(s, {L, 0, 0}, {L, 0, L}, {L, L, L}, {L, L, 0}, 1);     // red, x = 555
    quad(s, {0, 0, 0}, {0, 0, L}, {L, 0, L}, {L, 0, 0}, 0);     // floor
    quad(s, {0, L, 0}, {L, L, 0}, {L, L, L}, {0, L, L}, 0);     // ceiling
    quad(s, {0, 0, L}, {0, L, L}, {L, L, L}, {L, 0, L}, 0);     // back
    quad(s, {213, 554, 227}, {343, 554, 227}, {343, 554, 332}, {213, 554, 332}, 3);   // light
    rotated_box(s, {165, 330, 165}, 15.0, {265, 0, 295}, 0);
    rotated_box(s, {165, 165, 165}, -18.0, {130, 0, 65}, 0);
    set_camera(s, {278.0f, 278.0f, -800.0f}, {278.0f, 278.0f, 0.0f}, 40.0f);
    constant_env(s, {0.0f, 0.0f, 0.0f});
}

// Scene-generation uniforms from the project's counter hash.
float scene_u(uint32_t seed, uint32_t index, uint32_t dim) {
    return rng(make_path_key(seed, index, 0xFFFFFFFFu), 0, dim);
}

// 1024 spheres on a jittered 32x32 grid (spacing 1, jitter U[-0.1,0.1], radius U[0.2,0.4], resting
// on y=0) over a Lambert ground quad; material uniform over {Lambert albedo U[0,1]^3, metal albedo
// U[0.5,1]^3 fuzz U[0,0.3], dielectric IOR 1.5}; sky gradient white -> (0.5,0.7,1.0) as the light.
void scene_sphere_grid(Scene& s) {
    const uint32_t seed = 2024u;
    s.materials.push_back(mat(kLambert, {0.5f, 0.5f, 0.5f}));
    const float G = 100.0f;
    quad(s, {-G, 0, -G}, {G, 0, -G}, {G, 0, G}, {-G, 0, G}, 0);
    for (int gi = 0; gi < 32; ++gi)
        for (int gj = 0; gj < 32; ++gj) {
            const uint32_t k = (uint32_t)(gi * 32 + gj);
            double cx = (gi - 15.5) + (0.2 * scene_u(seed, k, 0) - 0.1);
            double cz = (gj - 15.5) + (0.2 * scene_u(seed, k, 1) - 0.1);
            float r = (float)(0.2 + 0.2 * scene_u(seed, k, 2));
            int type = (int)(scene_u(seed, k, 3) * 3.0f);
            if (type > 2) type = 2;
            Material m;
            if (type == 0) {
                m = mat(kLambert, {scene_u(seed, k, 4), scene_u(seed, k, 5), scene_u(seed, k, 6)});
            } else if (type == 1) {
                m = mat(kMetal, {(float)(0.5 + 0.5 * scene_u(seed, k, 4)), (float)(0.5 + 0.5 * scene_u(seed, k, 5)),
                                 (float)(0.5 + 0.5 * scene_u(seed, k, 6))},
                        {0, 0, 0}, 0.0f, (float)(0.3 * scene_u(seed, k, 7)));
            } else {
                m = mat(kDielectric, {1.0f, 1.0f, 1.0f}, {0, 0, 0}, 1.5f);
            }
            s.materials.push_back(m);
            s.spheres.push_back(sphere({(float)cx, r, (float)cz}, r, (int)s.materials.size() - 1));
        }
    set_camera(s, {18.0f, 6.0f, 18.0f}, {0.0f, 0.0f, 0.0f}, 40.0f);
    s.env = EnvParams{};
    s.env.sky_mode = 1;
    s.env.sky_lo = Vec3{1.0f, 1.0f, 1.0f};
    s.env.sky_hi = Vec3{0.5f, 0.7f, 1.0f};
    s.env.sky_light = 1;
}

// 2D gradient noise with quintic fade and hashed unit gradients, in binary64.
double grad_noise(double x, double z, uint32_t seed) {
    const double fx0 = std::floor(x), fz0 = std::floor(z);
    const int ix = (int)fx0, iz = (int)fz0;
    const double fx = x - fx0, fz = z - fz0;
    double g[4];
    for (int c = 0; c < 4; ++c) {
        const int dx = c & 1, dz = c >> 1;
        uint32_t h = hash_u32(hash_u32((uint32_t)(ix + dx) * 0x8da6b343u ^ (uint32_t)(iz + dz) * 0xd8163841u) ^ seed);
        double a = (double)(h >> 8) * (1.0 / 16777216.0) * 6.283185307179586;
        g[c] = std::cos(a) * (fx - dx) + std::sin(a) * (fz - dz);
    }
    auto fade = [](double t) { return t * t * t * (t * (t * 6.0 - 15.0) + 10.0); };
    const double u = fade(fx), v = fade(fz);
    const double a = g[0] + (g[1] - g[0]) * u, b = g[2] + (g[3] - g[2]) * u;
    return (a + (b - a) * v) * 1.4142135623730951;
}

// 1024x1024-vertex height field over [-5,5]^2 (2,093,058 triangles). Height: 6-octave fBm of
// gradient noise, amplitude 0.25, lacunarity 2, gain 0.5, base frequency 0.4, seed 7. Lambert
// albedo from a 16-band colour ramp by triangle height; a directional sun plus a sky gradient.
void scene_heightfield(Scene& s) {
    const int N = 1024;
    const double ext = 5.0, f0 = 0.4;
    const uint32_t seed = 7u;
    std::vector<double> hgt((size_t)N * N);
    double hmin = 1e30, hmax = -1e30;
    for (int i = 0; i < N; ++i)
        for (int j = 0; j < N; ++j) {
            double x = -ext + 2.0 * ext * i / (N - 1), z = -ext + 2.0 * ext * j / (N - 1);
            double h = 0.0, amp = 0.25, f = f0;
            for (int o = 0; o < 6; ++o) {
                h += amp * grad_noise(x * f, z * f, seed + (uint32_t)o * 0x9e3779b9u);
                amp *= 0.5;
                f *= 2.0;
            }
            hgt[(size_t)i * N + j] = h;
            hmin = std::min(hmin, h);
            hmax = std::max(hmax, h);
        }
    s.vertices.resize((size_t)N * N);
    for (int i = 0; i < N; ++i)
        for (int j = 0; j < N; ++j) {
            double x = -ext + 2.0 * ext * i / (N - 1), z = -ext + 2.0 * ext * j / (N - 1);
            s.vertices[(size_t)i * N + j] = Vec3{(float)x, (float)hgt[(size_t)i * N + j], (float)z};
        }
    // Colour ramp: water, sand, grass, rock, snow.
    const int bands = 16;
    const double st[5] = {0.0, 0.3, 0.45, 0.75, 1.0};
    const double col[5][3] = {{0.10, 0.22, 0.45}, {0.76, 0.70, 0.50}, {0.25, 0.50, 0.18}, {0.45, 0.40, 0.35}, {0.92, 0.92, 0.95}};
    for (int b = 0; b < bands; ++b) {
        double t = (b + 0.5) / bands;
        int k = 0;
        while (k < 3 && t > st[k + 1]) ++k;
        double w = (t - st[k]) / (st[k + 1] - st[k]);
        Vec3 a{(float)(col[k][0] + (col[k + 1][0] - col[k][0]) * w), (float)(col[k][1] + (col[k + 1][1] - col[k][1]) * w),
               (float)(col[k][2] + (col[k + 1][2] - col[k][2]) * w)};
        s.materials.push_back(mat(kLambert, a));
    }
    s.triangles.reserve((size_t)2 * (N - 1) * (N - 1));
    auto band = [&](int a, int b, int c) {
        double h = (hgt[a] + hgt[b] + hgt[c]) / 3.0;
        int k = (int)((h - hmin) / (hmax - hmin) * bands);
        return k < 0 ? 0 : (k >= bands ? bands - 1 : k);
    };
    for (int i = 0; i < N - 1; ++i)
        for (int j = 0; j < N - 1; ++j) {
            int v0 = i * N + j, v1 = v0 + 1, v2 = v0 + N, v3 = v2 + 1;
            s.triangles.push_back({v0, v1, v2, band(v0, v1, v2)});
            s.triangles.push_back({v1, v3, v2, band(v1, v3, v2)});
        }
    set_camera(s, {6.0f, 2.5f, 6.0f}, {0.0f, 0.0f, 0.0f}, 45.0f);
    s.env = EnvParams{};
    s.env.sky_mode = 1;
    s.env.sky_lo = Vec3{0.5f, 0.5f, 0.5f};
    s.env.sky_hi = Vec3{0.25f, 0.35f, 0.5f};
    s.env.sky_light = 1;
    s.env.sun_light = 1;
    s.env.sun_dir = normalize(Vec3{0.4f, 0.75f, 0.35f});
    s.env.sun_E = Vec3{2.6f, 2.5f, 2.3f};
}

}  // namespace

std::vector<std::string> scene_names() {
    return {"cornell", "divergence", "trianglehv", "furnace", "cornell555", "sphere_grid", "heightfield"};
}

Scene make_scene(const std::string& name, int width, int height) {
    Scene s;
    s.name = name;
    s.width = width;
    s.height = height;
    if (name == "cornell") scene_ref_cornell(s);
    else if (name == "divergence") scene_ref_divergence(s);
    else if (name == "trianglehv") scene_ref_trianglehv(s);
    else if (name == "furnace") scene_ref_furnace(s);
    else if (name == "cornell555") scene_cornell555(s);
    else if (name == "sphere_grid") scene_sphere_grid(s);
    else if (name == "heightfield") scene_heightfield(s);
    else throw std::invalid_argument("Unknown scene name: " + name);
    finalize_scene(s);
    return s;
}

Camera orbited(const Camera& c, float yaw_deg) {
    float dx = c.position.x - c.look_at.x, dy = c.position.y - c.look_at.y, dz = c.position.z - c.look_at.z;
    float dist = std::sqrt(dx * dx + dy * dy + dz * dz);
    float pitch = std::asin(std::max(-1.0f, std::min(1.0f, dy / std::max(dist, 1e-6f))));
    float yaw = yaw_deg * kPi / 180.0f;
    float cp = std::cos(pitch);
    Camera o = c;
    o.position = Vec3{c.look_at.x + dist * cp * std::sin(yaw), c.look_at.y + dist * std::sin(pitch),
                      c.look_at.z + dist * cp * std::cos(yaw)};
    return o;
}

void orbit_camera(Scene& s, float yaw_deg) { s.camera = orbited(s.camera, yaw_deg); }

CameraRig view_rig(const Scene& s, int width, int height, bool orbit, float yaw_deg) {
    return make_camera_rig(orbit ? orbited(s.camera, yaw_deg) : s.camera, width, height);
}

}  // namespace wf
