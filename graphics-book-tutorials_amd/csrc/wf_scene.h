// Host scene model: owning arrays in the reference layout, the procedural scenes, validation,
// the light list, the binned-SAH BVH and the conversion to the device layout of wf_core.h.
//   Scene types          include/wf/scene.hpp:12-100
//   finalize / validate  src/scene.cpp:10-104
//   scene hash           src/scene.cpp:107-160
//   BVH build            src/bvh_build.cpp:60-251
//   procedural scenes    src/scenes.cpp:29-344
#pragma once
#include <cstdint>
#include <string>
#include <vector>

#include "wf_core.h"

namespace wf {

struct Sphere {
    Vec3 center;
    float radius;
    int material;
    int pad[3];
};

struct Triangle {
    int i0, i1, i2;
    int material;
};

struct Camera {
    Vec3 position;
    Vec3 look_at;
    Vec3 up;
    float vfov_deg;
};

// Reference-layout BVH node (48 bytes, preorder: the left child of node N is N+1).
struct BvhNode {
    Vec3 lo;
    int skip;
    Vec3 hi;
    int prim_start;   // leaf: first entry in prim_order; internal: -1
    int prim_count;   // leaf: 1..4; internal: 0
    int pad[3];
};

struct Scene {
    std::string name;
    std::vector<Sphere> spheres;
    std::vector<Vec3> vertices;
    std::vector<Triangle> triangles;
    std::vector<Material> materials;
    EnvParams env{};
    Camera camera{};
    int width = 1280, height = 720;

    // Filled by finalize_scene.
    std::vector<BvhNode> nodes;       // reference layout
    std::vector<int> prim_order;      // leaf entries -> primitive ids
    std::vector<Light> lights;        // emissive primitives (slot filled from prim_order)
    // Device layout (wf_core.h), also used by the CPU arm.
    std::vector<Node> dnodes;
    std::vector<Prim> dprims;

    int n_prims() const { return (int)spheres.size() + (int)triangles.size(); }
    bool finalized() const { return !nodes.empty(); }
};

CameraRig make_camera_rig(const Camera& cam, int width, int height);

// Empty string if valid, else the reason (bad material or vertex index, non-finite vertex or
// centre, bad radius, no primitives, no materials).
std::string validate_scene(const Scene& s);

// Validates, builds the light list and the BVH, and derives the device layout.
// Throws std::runtime_error on invalid input and leaves nodes and lights empty.
void finalize_scene(Scene& s);

// Binned SAH, 8 bins on the longest centroid axis, leaf <= 4, median fallbacks. Output is the
// reference's node array and prim_order, bit for bit.
void build_bvh(Scene& s);

// Empty string if the reference-layout BVH satisfies the structural invariants.
std::string validate_bvh(const Scene& s);

// Number of leading breadth-first nodes placed at the start of the device node array.
constexpr int kTopFirstNodes = 8192;

// Rebuilds dnodes/dprims/lights[].slot from nodes/prim_order.
void build_device_layout(Scene& s);

// FNV-1a 64 over the reference's canonical scene bytes; extension fields (fuzz, sky, sun) are
// appended only when set, so the reference scenes hash as in the reference.
uint64_t scene_hash(const Scene& s);

// Procedural scenes. Reference scenes: "cornell", "divergence", "trianglehv", "furnace".
// Benchmark scenes: "cornell555" (classic 555-unit Cornell box), "sphere_grid" (1024 spheres of
// three material types under a sky-gradient emitter), "heightfield" (1024x1024-vertex fBm terrain,
// sun plus sky). Throws std::invalid_argument for an unknown name. The result is finalized.
Scene make_scene(const std::string& name, int width, int height);
std::vector<std::string> scene_names();

// Orbit the camera about its look-at point to an absolute yaw (degrees), keeping the distance
// and pitch of the current position (the viewer's orbit, src/viewer/viewer.cpp:80-90).
void orbit_camera(Scene& s, float yaw_deg);
Camera orbited(const Camera& c, float yaw_deg);

// Camera rig of a view of the scene: its own image size and, with orbit, the camera orbited to
// yaw_deg. Jobs of the batched engines render such views of one shared scene.
CameraRig view_rig(const Scene& s, int width, int height, bool orbit, float yaw_deg);

// View over the device-layout arrays held by the scene (host pointers).
SceneView host_view(const Scene& s);

// Leaf/internal/depth statistics of the reference-layout tree.
struct BvhStats {
    int nodes = 0, leaves = 0, internal = 0, max_depth = 0;
    double sah_cost = 0.0;
};
BvhStats bvh_stats(const Scene& s);

}  // namespace wf
