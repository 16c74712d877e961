// Command-line driver over the `wf` CLI's host helpers (graphics-book-tutorials_amd/cli/wf_cli.h)
// for tests/test_cli.py, which builds it with g++. Each command prints one line of JSON:
//   parse ARGS...            the parsed Config, or {"error": "..."} for a ConfigError
//   percentile P V...        percentile of V at P, or {"error": ...} for an empty set
//   crc32 TEXT               CRC-32 of TEXT's bytes
//   png PATH W H             writes a W x H PNG of the test pattern (x*7 + y*13 + c*29) mod 256
//   pfm PATH W H             writes a W x H PFM of the test pattern (x + 1000 y + 0.25 c)
//   tonemap V...             the 8-bit tone-mapped value of each V
//   readpfm PATH             width, height and the top-row-first floats' bits of a PFM, or {"exception": ...}
//   json TEXT                {"ok": true, "kind": K} when TEXT parses as one JSON value, else {"error": ...}
//   gifidx PATH W H KIND SEED   one-frame GIF of generated palette indices (KIND random | runs | few |
//                            const), the indices also raw in PATH.idx; prints the LZW end state
//   gifquant W H R G B       error-diffuses a flat W x H image of colour (R, G, B); prints the mean
//                            palette colour and the number of distinct indices
//   gifparse ARGS...         the parsed `wf gif` options (after "gif"), or {"error": "..."}
//   gifanim PATH W H         three frames (test pattern with a caption, flat grey, the pattern) with
//                            delays 10, 20, 30; the first frame's indices raw in PATH.idx
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "wf_cli.h"
#include "wf_gif.h"
#include "wf_json.h"

using namespace wfcli;

static unsigned char pattern8(int x, int y, int c) { return (unsigned char)((x * 7 + y * 13 + c * 29) % 256); }

int main(int argc, char** argv) {
    if (argc < 2) return 2;
    const std::string cmd = argv[1];
    try {
        if (cmd == "parse") {
            std::vector<const char*> a = {"wf"};
            for (int i = 2; i < argc; ++i) a.push_back(argv[i]);
            try {
                Config c = parse_args((int)a.size(), a.data());
                std::printf("{\"scene\": %s, \"width\": %d, \"height\": %d, \"spp\": %d, \"seed\": %u, \"max_depth\": %d, "
                            "\"mode\": %s, \"report\": %s, \"out\": %s, \"threads\": %d, \"reps\": %d, \"warmup\": %d}\n",
                            jstr(c.scene).c_str(), c.width, c.height, c.spp, c.seed, c.max_depth, jstr(c.mode).c_str(),
                            jstr(c.report_path).c_str(), jstr(c.out_prefix).c_str(), c.threads, c.reps, c.warmup);
            } catch (const ConfigError& e) {
                std::printf("{\"error\": %s}\n", jstr(e.what()).c_str());
            }
        } else if (cmd == "percentile") {
            const double p = std::atof(argv[2]);
            std::vector<double> v;
            for (int i = 3; i < argc; ++i) v.push_back(std::atof(argv[i]));
            try {
                std::printf("{\"value\": %s}\n", jnum(percentile(v, p)).c_str());
            } catch (const std::invalid_argument& e) {
                std::printf("{\"error\": %s}\n", jstr(e.what()).c_str());
            }
        } else if (cmd == "crc32") {
            const std::string t = argv[2];
            std::printf("{\"crc32\": %u}\n", crc32((const uint8_t*)t.data(), t.size()) ^ 0xFFFFFFFFu);
        } else if (cmd == "png" || cmd == "pfm") {
            const int w = std::atoi(argv[3]), h = std::atoi(argv[4]);
            if (cmd == "png") {
                std::vector<uint8_t> px((size_t)3 * w * h);
                for (int y = 0; y < h; ++y)
                    for (int x = 0; x < w; ++x)
                        for (int c = 0; c < 3; ++c) px[((size_t)y * w + x) * 3 + c] = pattern8(x, y, c);
                write_png(argv[2], w, h, px);
            } else {
                std::vector<float> px((size_t)3 * w * h);
                for (int y = 0; y < h; ++y)
                    for (int x = 0; x < w; ++x)
                        for (int c = 0; c < 3; ++c) px[((size_t)y * w + x) * 3 + c] = (float)x + 1000.0f * (float)y + 0.25f * (float)c;
                write_pfm(argv[2], w, h, px);
            }
            std::printf("{\"ok\": true}\n");
        } else if (cmd == "tonemap") {
            std::vector<float> v;
            for (int i = 2; i < argc; ++i) v.push_back(std::strtof(argv[i], nullptr));
            const std::vector<uint8_t> t = tonemap_srgb8(v);
            std::printf("{\"bytes\": [");
            for (size_t i = 0; i < t.size(); ++i) std::printf("%s%u", i ? ", " : "", (unsigned)t[i]);
            std::printf("]}\n");
        } else if (cmd == "readpfm") {
            int w = 0, h = 0;
            const std::vector<float> px = read_pfm(argv[2], &w, &h);
            std::printf("{\"width\": %d, \"height\": %d, \"bits\": [", w, h);
            for (size_t i = 0; i < px.size(); ++i) {
                uint32_t u;
                std::memcpy(&u, &px[i], 4);
                std::printf("%s%u", i ? ", " : "", u);
            }
            std::printf("]}\n");
        } else if (cmd == "json") {
            try {
                const JValue v = json_parse(argv[2]);
                std::printf("{\"ok\": true, \"kind\": %d}\n", (int)v.kind);
            } catch (const JsonError& e) {
                std::printf("{\"error\": %s}\n", jstr(e.what()).c_str());
            }
        } else if (cmd == "gifidx") {
            const std::string path = argv[2];
            const int w = std::atoi(argv[3]), h = std::atoi(argv[4]);
            const std::string kind = argv[5];
            uint32_t st = (uint32_t)std::strtoul(argv[6], nullptr, 10) * 2654435761u + 12345u;
            auto rnd = [&]() {
                st = st * 1664525u + 1013904223u;
                return st >> 8;
            };
            std::vector<uint8_t> idx((size_t)w * h);
            const uint32_t alphabet = kind == "few" ? 2u + rnd() % 3u : 256u;
            for (size_t i = 0; i < idx.size();) {
                if (kind == "runs") {
                    const size_t run = 1 + rnd() % 64;
                    const uint8_t v = (uint8_t)(rnd() % 7u);
                    for (size_t k = 0; k < run && i < idx.size(); ++k) idx[i++] = v;
                } else if (kind == "const") {
                    idx[i++] = 77;
                } else {
                    idx[i++] = (uint8_t)(rnd() % alphabet);
                }
            }
            GifLzwEnd end;
            const std::vector<uint8_t> data = gif_lzw(idx.data(), idx.size(), &end);
            GifWriter g(path, w, h);
            g.frame(idx, 5);
            g.close();
            std::FILE* f = std::fopen((path + ".idx").c_str(), "wb");
            std::fwrite(idx.data(), 1, idx.size(), f);
            std::fclose(f);
            std::printf("{\"next\": %d, \"size\": %d, \"wide_end\": %s, \"clears\": %d, \"bytes\": %zu}\n", end.next,
                        end.size, end.wide_end ? "true" : "false", end.clears, data.size());
        } else if (cmd == "gifparse") {
            std::vector<const char*> a = {"wf", "gif"};
            for (int i = 2; i < argc; ++i) a.push_back(argv[i]);
            try {
                const GifConfig c = parse_gif_args((int)a.size(), a.data());
                std::string scenes, steps;
                for (const std::string& n : c.scenes) scenes += (scenes.empty() ? "" : ", ") + jstr(n);
                for (int v : c.steps) steps += (steps.empty() ? "" : ", ") + std::to_string(v);
                std::printf("{\"out\": %s, \"scenes\": [%s], \"width\": %d, \"height\": %d, \"mode\": %s, "
                            "\"steps\": [%s], \"orbit\": %d, \"spp\": %d, \"seed\": %u, \"max_depth\": %d, "
                            "\"exposure\": %.9g, \"delay\": %d, \"hold\": %d, \"png_prefix\": %s}\n",
                            jstr(c.out).c_str(), scenes.c_str(), c.width, c.height, jstr(c.mode).c_str(), steps.c_str(),
                            c.orbit, c.spp, c.seed, c.max_depth, c.exposure, c.delay, c.hold, jstr(c.png_prefix).c_str());
            } catch (const ConfigError& e) {
                std::printf("{\"error\": %s}\n", jstr(e.what()).c_str());
            }
        } else if (cmd == "gifquant") {
            const int w = std::atoi(argv[2]), h = std::atoi(argv[3]);
            std::vector<uint8_t> rgb((size_t)3 * w * h);
            for (size_t i = 0; i < rgb.size(); ++i) rgb[i] = (uint8_t)std::atoi(argv[4 + i % 3]);
            const std::vector<uint8_t> idx = gif_quantize(rgb, w, h);
            uint8_t table[768];
            gif_palette(table);
            double m[3] = {0, 0, 0};
            std::vector<int> seen(256, 0);
            for (uint8_t v : idx) {
                seen[v] = 1;
                for (int c = 0; c < 3; ++c) m[c] += table[3 * v + c];
            }
            int distinct = 0;
            for (int v : seen) distinct += v;
            std::printf("{\"mean\": [%.6f, %.6f, %.6f], \"distinct\": %d}\n", m[0] / idx.size(), m[1] / idx.size(),
                        m[2] / idx.size(), distinct);
        } else if (cmd == "gifanim") {
            const std::string path = argv[2];
            const int w = std::atoi(argv[3]), h = std::atoi(argv[4]);
            std::vector<uint8_t> pat((size_t)3 * w * h), grey((size_t)3 * w * h, 128);
            for (int y = 0; y < h; ++y)
                for (int x = 0; x < w; ++x)
                    for (int c = 0; c < 3; ++c) pat[3 * ((size_t)y * w + x) + c] = pattern8(x, y, c);
            std::vector<uint8_t> cap = pat;
            gif_caption(cap, w, h, 2, 2, "ABC 123", 1);
            const std::vector<uint8_t> first = gif_quantize(cap, w, h);
            GifWriter g(path, w, h);
            g.frame(first, 10);
            g.frame(gif_quantize(grey, w, h), 20);
            g.frame(gif_quantize(pat, w, h), 30);
            g.close();
            std::FILE* f = std::fopen((path + ".idx").c_str(), "wb");
            std::fwrite(first.data(), 1, first.size(), f);
            std::fclose(f);
            std::printf("{\"frames\": %d}\n", g.frames());
        } else {
            return 2;
        }
    } catch (const std::exception& e) {
        std::printf("{\"exception\": %s}\n", jstr(e.what()).c_str());
    }
    return 0;
}
