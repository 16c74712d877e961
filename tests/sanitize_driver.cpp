// Host code under AddressSanitizer and UndefinedBehaviorSanitizer (the reference's sanitizer
// build, CMakeLists.txt:13,21-25, restricted to what runs without a GPU): every scene's build and
// BVH (graphics-book-tutorials_amd/csrc/wf_scene.cpp), the CPU arm (wf_cpu.cpp), the `wf` CLI's
// host helpers (flag parsing, JSON reader, report validator, PNG/PFM writers, PFM reader, image
// difference) and the oracle (oracle/wf_oracle.c: BVH, traversal, renders, the wavefront
// restatement). Built by `make -C graphics-book-tutorials_amd sanitize` and run by
// tests/test_host_sanitizers.py; a sanitizer report makes it exit non-zero.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "wf_cli.h"
#include "wf_cpu.h"
#include "wf_gif.h"
#include "wf_json.h"
#include "wf_oracle.h"
#include "wf_scene.h"

namespace {

int failures = 0;

void expect(bool ok, const std::string& what) {
    if (!ok) {
        std::printf("FAIL %s\n", what.c_str());
        ++failures;
    }
}

float bits_f(int v) {
    float f;
    std::memcpy(&f, &v, 4);
    return f;
}

// The oracle's scene from a product scene (the C ABI's array layout, wf_scene_export).
or_scene* oracle_scene(const wf::Scene& s) {
    std::vector<float> sph(8 * s.spheres.size()), ver(3 * s.vertices.size()), mat(12 * s.materials.size());
    std::vector<int> tri(4 * s.triangles.size());
    for (size_t i = 0; i < s.spheres.size(); ++i) {
        const wf::Sphere& p = s.spheres[i];
        float* o = &sph[8 * i];
        o[0] = p.center.x, o[1] = p.center.y, o[2] = p.center.z, o[3] = p.radius, o[4] = bits_f(p.material);
    }
    for (size_t i = 0; i < s.vertices.size(); ++i)
        ver[3 * i] = s.vertices[i].x, ver[3 * i + 1] = s.vertices[i].y, ver[3 * i + 2] = s.vertices[i].z;
    for (size_t i = 0; i < s.triangles.size(); ++i) {
        const wf::Triangle& t = s.triangles[i];
        tri[4 * i] = t.i0, tri[4 * i + 1] = t.i1, tri[4 * i + 2] = t.i2, tri[4 * i + 3] = t.material;
    }
    for (size_t i = 0; i < s.materials.size(); ++i) {
        const wf::Material& m = s.materials[i];
        float* o = &mat[12 * i];
        o[0] = bits_f(m.type), o[1] = m.ior, o[2] = m.fuzz;
        o[4] = m.albedo.x, o[5] = m.albedo.y, o[6] = m.albedo.z;
        o[8] = m.emission.x, o[9] = m.emission.y, o[10] = m.emission.z;
    }
    const wf::EnvParams& e = s.env;
    const float env[20] = {e.env.x, e.env.y, e.env.z, bits_f(e.sky_mode), e.sky_lo.x, e.sky_lo.y, e.sky_lo.z,
                           bits_f(e.sky_light), e.sky_hi.x, e.sky_hi.y, e.sky_hi.z, bits_f(e.sun_light), e.sun_dir.x,
                           e.sun_dir.y, e.sun_dir.z, 0.0f, e.sun_E.x, e.sun_E.y, e.sun_E.z, 0.0f};
    const wf::Camera& c = s.camera;
    const float cam[10] = {c.position.x, c.position.y, c.position.z, c.look_at.x, c.look_at.y, c.look_at.z,
                           c.up.x, c.up.y, c.up.z, c.vfov_deg};
    return or_scene_new(sph.data(), (int)s.spheres.size(), ver.data(), (int)s.vertices.size(), tri.data(),
                        (int)s.triangles.size(), mat.data(), (int)s.materials.size(), env, cam, s.width, s.height);
}

void scenes_cpu_arm_and_oracle() {
    for (const std::string& name : wf::scene_names()) {
        const int w = 16, h = 9, spp = 2;
        wf::Scene s = wf::make_scene(name, w, h);
        expect(wf::validate_scene(s).empty(), name + " validate_scene");
        expect(wf::validate_bvh(s).empty(), name + " validate_bvh");
        const wf::BvhStats st = wf::bvh_stats(s);
        expect(st.nodes == (int)s.nodes.size() && st.leaves > 0, name + " bvh_stats");
        (void)wf::scene_hash(s);
        (void)wf::host_view(s);
        wf::CpuCounts cc;
        double ms = 0.0;
        const wf::RenderResult r = wf::render_cpu(s, spp, 1234u, 8, 3, &cc, &ms);
        expect(r.rgb.size() == (size_t)3 * w * h && r.lum_var.size() == (size_t)w * h, name + " render size");
        or_scene* o = oracle_scene(s);
        expect(o != nullptr, name + " oracle scene");
        if (o) {
            std::vector<float> rgb(3 * w * h), lum(w * h), lum2(w * h);
            uint64_t cnt[3] = {0, 0, 0};
            or_render(o, 0, spp, 1234u, 8, 2, rgb.data(), lum.data(), lum2.data(), cnt);
            expect(cnt[0] == cc.extension_rays && cnt[1] == cc.shadow_rays, name + " oracle rays = CPU arm");
            uint32_t rec[6 * 16];
            uint64_t hashes[16];
            uint64_t cnt2[3] = {0, 0, 0};
            const int nb = or_wavefront_pass(o, 0u, 1234u, 8, 50, rec, hashes, 16, rgb.data(), lum.data(), lum2.data(),
                                             cnt2);
            expect(nb >= 1, name + " wavefront pass");
            const float rays[16] = {0, 1, 0, 1e-4f, 0, -1, 0, 1e30f, 1, 2, 3, 1e-4f, -1, -0.5f, 0.25f, 1e30f};
            float t[2];
            int prim[2];
            or_intersect(o, rays, 2, 0, t, prim);
            or_intersect(o, rays, 2, 1, t, prim);
            or_scene_free(o);
        }
        wf::orbit_camera(s, 37.0f);
        (void)wf::view_rig(s, 8, 8, true, 90.0f);
    }
    // An invalid scene is refused without touching out-of-range indices.
    wf::Scene bad = wf::make_scene("cornell", 8, 8);
    bad.triangles[0].i1 = 1 << 20;
    expect(!wf::validate_scene(bad).empty(), "bad vertex index refused");
    bad = wf::make_scene("cornell", 8, 8);
    bad.triangles[0].material = -3;
    expect(!wf::validate_scene(bad).empty(), "bad material index refused");
}

void cli_helpers(const std::string& dir) {
    using namespace wfcli;
    const char* good[] = {"wf", "render", "--scene", "sphere_grid", "--width", "31", "--height", "7", "--spp", "3"};
    const Config c = parse_args(10, good);
    expect(c.width == 31 && c.height == 7 && c.spp == 3, "parse_args");
    const char* bad[][4] = {{"wf", "render", "--width", "0"}, {"wf", "render", "--spp", "x"},
                            {"wf", "render", "--bogus", "1"}, {"wf", "render", "--mode", "nope"}};
    for (auto& b : bad) {
        bool threw = false;
        try {
            (void)parse_args(4, b);
        } catch (const ConfigError&) {
            threw = true;
        }
        expect(threw, std::string("parse_args rejects ") + b[2]);
    }
    const char* texts[] = {"{\"a\": [1, 2.5e3, -0.0, true, null, \"\\u00e9\\ud83d\\ude00\"]}", "[1,]", "01", "\"\\x\"",
                           "{\"a\" 1}", "[[[[[[[[[[]]]]]]]]]]", "\"unterminated", "1e", "{}", "  7  "};
    for (const char* t : texts) {
        try {
            (void)json_parse(t);
        } catch (const JsonError&) {
        }
    }
    const std::string deep(300, '[');
    bool deep_refused = false;
    try {
        (void)json_parse(deep + std::string(300, ']'));
    } catch (const JsonError&) {
        deep_refused = true;
    }
    expect(deep_refused, "nesting limit");
    expect(!validate_report_json("{}").empty() && !validate_report_json("[").empty(), "validate_report_json");
    const int w = 5, h = 3;
    std::vector<float> img(3 * w * h);
    for (size_t i = 0; i < img.size(); ++i) img[i] = 0.1f * (float)i;
    write_pfm(dir + "/s.pfm", w, h, img);
    write_png(dir + "/s.png", w, h, tonemap_srgb8(img));
    int rw = 0, rh = 0;
    const std::vector<float> back = read_pfm(dir + "/s.pfm", &rw, &rh);
    expect(rw == w && rh == h && back == img, "PFM round trip");
    const ImageDiff d = image_diff(img, back);
    expect(d.max_abs_rel == 0.0 && !d.nonfinite_mismatch, "image_diff");
    std::FILE* f = std::fopen((dir + "/t.pfm").c_str(), "wb");
    std::fputs("PF\n4 4\n-1.0\n", f);
    std::fwrite(img.data(), 4, 7, f);   // truncated
    std::fclose(f);
    bool truncated = false;
    try {
        (void)read_pfm(dir + "/t.pfm", &rw, &rh);
    } catch (const std::runtime_error&) {
        truncated = true;
    }
    expect(truncated, "truncated PFM refused");
    std::vector<double> v = {3, 1, 2};
    expect(percentile(v, 0.5) == 2.0, "percentile");
    // GIF: quantise, caption (clipped at every edge), LZW over a table that fills and clears, file.
    const int gw = 161, gh = 121;
    std::vector<uint8_t> g8((size_t)3 * gw * gh);
    uint32_t lcg = 1u;
    for (size_t i = 0; i < g8.size(); ++i) {
        lcg = lcg * 1664525u + 1013904223u;
        g8[i] = (uint8_t)(lcg >> 24);
    }
    gif_caption(g8, gw, gh, -5, -3, "EDGE 0123456789 ABCDEFGHIJKLMNOPQRSTUVWXYZ .:-/%=+()", 2);
    gif_caption(g8, gw, gh, gw - 4, gh - 2, "X", 3);
    const std::vector<uint8_t> gi = gif_quantize(g8, gw, gh);
    GifLzwEnd end;
    (void)gif_lzw(gi.data(), gi.size(), &end);
    expect(end.clears > 0, "gif lzw clears its table");
    (void)gif_lzw(gi.data(), 1);
    (void)gif_lzw(gi.data(), 0);
    GifWriter gw_file(dir + "/s.gif", gw, gh);
    gw_file.frame(gi, 7);
    gw_file.frame(gif_quantize(std::vector<uint8_t>((size_t)3 * gw * gh, 200), gw, gh), 9);
    gw_file.close();
    expect(gif_auto_exposure(img) > 0.0f, "gif exposure");
    const char* gargs[] = {"wf", "gif", "o.gif", "--steps", "1,3,9", "--scenes", "cornell,furnace", "--exposure", "0.5"};
    const GifConfig gc = parse_gif_args(9, gargs);
    expect(gc.steps.size() == 3 && gc.scenes.size() == 2 && gc.exposure == 0.5f, "parse_gif_args");
}

}  // namespace

int main(int argc, char** argv) {
    const std::string dir = argc > 1 ? argv[1] : ".";
    scenes_cpu_arm_and_oracle();
    cli_helpers(dir);
    std::printf("%s (%d failures)\n", failures ? "FAILED" : "ok", failures);
    return failures ? 1 : 0;
}
