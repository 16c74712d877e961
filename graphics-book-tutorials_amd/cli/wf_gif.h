// Animated GIF output for `wf gif`, the counterpart of the reference's animation tool
// (tools/wfgif.cpp: progressive-refinement frames of real renders written as one looping GIF) and of
// its viewer's orbit camera (src/viewer/viewer.cpp:56-90), headless: frames are rendered through the
// C ABI, tone-mapped, captioned with a small bitmap font over a darkened box (the reference's
// overlay, include/wf/visual.hpp:10-55), reduced to a fixed palette with error diffusion and
// compressed with the GIF variant of LZW.
//
// Format: GIF89a (W3C/CompuServe specification, 1990): a global 256-entry colour table, the NETSCAPE2.0
// application extension for looping, and per frame a graphic control extension (delay), an image
// descriptor and LZW data in sub-blocks of at most 255 bytes.
#pragma once
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <fstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "wf_cli.h"

namespace wfcli {

// ---- palette: a 6 x 6 x 6 colour cube (entries 0..215, index r * 36 + g * 6 + b, levels 0, 51,
// ..., 255) and a ramp of 40 greys (entries 216..255, 255 (k + 1) / 41 rounded), so that the
// neutral skies and walls of the scenes dither between greys rather than between colours ----
constexpr int kGifCube = 6, kGifGreys = 40;

inline int gif_cube_value(int q) { return 51 * q; }
inline int gif_grey_value(int k) { return (int)std::lround(255.0 * (k + 1) / (kGifGreys + 1)); }

inline void gif_palette(uint8_t* table768) {
    for (int i = 0; i < 216; ++i) {
        table768[3 * i + 0] = (uint8_t)gif_cube_value(i / 36);
        table768[3 * i + 1] = (uint8_t)gif_cube_value(i / 6 % 6);
        table768[3 * i + 2] = (uint8_t)gif_cube_value(i % 6);
    }
    for (int k = 0; k < kGifGreys; ++k)
        for (int c = 0; c < 3; ++c) table768[3 * (216 + k) + c] = (uint8_t)gif_grey_value(k);
}

// Nearest palette entry to v (squared distance): the cube's nearest is per channel; the ramp's is
// the grey nearest to the channels' mean. The cube wins ties.
inline int gif_nearest(const float* v, float* chosen) {
    float cube[3], dc = 0.0f;
    int q[3];
    for (int c = 0; c < 3; ++c) {
        q[c] = std::clamp((int)std::lround(v[c] / 51.0f), 0, kGifCube - 1);
        cube[c] = (float)gif_cube_value(q[c]);
        dc += (v[c] - cube[c]) * (v[c] - cube[c]);
    }
    const float mean = (v[0] + v[1] + v[2]) / 3.0f;
    const int k = std::clamp((int)std::lround(mean * (kGifGreys + 1) / 255.0f) - 1, 0, kGifGreys - 1);
    const float g = (float)gif_grey_value(k);
    float dg = 0.0f;
    for (int c = 0; c < 3; ++c) dg += (v[c] - g) * (v[c] - g);
    if (dg < dc) {
        chosen[0] = chosen[1] = chosen[2] = g;
        return 216 + k;
    }
    for (int c = 0; c < 3; ++c) chosen[c] = cube[c];
    return q[0] * 36 + q[1] * 6 + q[2];
}

// Floyd-Steinberg error diffusion (Floyd and Steinberg, 1976) onto the palette, left to right and
// top to bottom: each pixel's quantisation error goes 7/16 to its right neighbour and 3/16, 5/16,
// 1/16 to the lower-left, lower and lower-right ones. Two rows of carried error are kept.
inline std::vector<uint8_t> gif_quantize(const std::vector<uint8_t>& rgb8, int w, int h) {
    if (w < 1 || h < 1 || rgb8.size() != (size_t)3 * w * h) throw std::invalid_argument("gif_quantize: bad image size");
    std::vector<uint8_t> idx((size_t)w * h);
    std::vector<float> cur((size_t)3 * (w + 2), 0.0f), nxt((size_t)3 * (w + 2), 0.0f);   // index x + 1
    for (int y = 0; y < h; ++y) {
        std::fill(nxt.begin(), nxt.end(), 0.0f);
        for (int x = 0; x < w; ++x) {
            float v[3], got[3];
            for (int c = 0; c < 3; ++c)
                v[c] = std::clamp((float)rgb8[3 * ((size_t)y * w + x) + c] + cur[3 * (x + 1) + c], 0.0f, 255.0f);
            idx[(size_t)y * w + x] = (uint8_t)gif_nearest(v, got);
            for (int c = 0; c < 3; ++c) {
                const float e = v[c] - got[c];
                cur[3 * (x + 2) + c] += e * (7.0f / 16.0f);
                nxt[3 * (x + 0) + c] += e * (3.0f / 16.0f);
                nxt[3 * (x + 1) + c] += e * (5.0f / 16.0f);
                nxt[3 * (x + 2) + c] += e * (1.0f / 16.0f);
            }
        }
        std::swap(cur, nxt);
    }
    return idx;
}

// GIF LZW with a minimum code size of 8: clear code 256, end code 257, first free code 258, codes of
// 9 to 12 bits packed least significant bit first. The string table is a trie (child[code * 256 +
// byte], 0 = absent: no code below 258 is ever a child). A decoder adds its table entry one code
// later than the encoder, so the encoder widens its codes once its next free code passes 2^size,
// and sends the end code one bit wider when the decoder's last entry reaches 2^size. The table is
// cleared (clear code, 12 bits) when it is full.
struct GifLzwEnd {
    int next = 0, size = 0;   // next free code and code width when the last pixel code was sent
    bool wide_end = false;    // the end code went out one bit wider than that code
    int clears = 0;           // table clears after the first clear code
};

inline std::vector<uint8_t> gif_lzw(const uint8_t* px, size_t n, GifLzwEnd* end = nullptr) {
    std::vector<uint16_t> child((size_t)4096 * 256, 0);
    std::vector<uint8_t> out;
    uint32_t acc = 0;
    int bits = 0, size = 9, next = 258;
    bool fresh = true;   // no entry added since the last clear
    auto emit = [&](int code) {
        acc |= (uint32_t)code << bits;
        bits += size;
        while (bits >= 8) {
            out.push_back((uint8_t)(acc & 0xffu));
            acc >>= 8;
            bits -= 8;
        }
    };
    emit(256);
    if (n > 0) {
        int cur = px[0];
        for (size_t i = 1; i < n; ++i) {
            const int b = px[i];
            const int c = child[(size_t)cur * 256 + b];
            if (c) {
                cur = c;
                continue;
            }
            emit(cur);
            child[(size_t)cur * 256 + b] = (uint16_t)next++;
            fresh = false;
            if (next == 4096) {
                emit(256);
                if (end) end->clears += 1;
                std::fill(child.begin(), child.end(), (uint16_t)0);
                size = 9;
                next = 258;
                fresh = true;
            } else if (next > (1 << size)) {
                ++size;
            }
            cur = b;
        }
        emit(cur);
        if (end) {
            end->next = next;
            end->size = size;
        }
        if (!fresh && next == (1 << size) && size < 12) {
            ++size;
            if (end) end->wide_end = true;
        }
    }
    emit(257);
    if (bits > 0) out.push_back((uint8_t)(acc & 0xffu));
    return out;
}

// ---- captions: a 3 x 5 bitmap font (rows top to bottom, '#' = ink) ----
struct GifGlyph {
    char c;
    const char* rows[5];
};

inline const GifGlyph* gif_font() {
    static const GifGlyph font[] = {
        {'0', {"###", "#.#", "#.#", "#.#", "###"}}, {'1', {".#.", "##.", ".#.", ".#.", "###"}},
        {'2', {"###", "..#", "###", "#..", "###"}}, {'3', {"###", "..#", ".##", "..#", "###"}},
        {'4', {"#.#", "#.#", "###", "..#", "..#"}}, {'5', {"###", "#..", "###", "..#", "###"}},
        {'6', {"###", "#..", "###", "#.#", "###"}}, {'7', {"###", "..#", ".#.", ".#.", ".#."}},
        {'8', {"###", "#.#", "###", "#.#", "###"}}, {'9', {"###", "#.#", "###", "..#", "###"}},
        {'A', {".#.", "#.#", "###", "#.#", "#.#"}}, {'B', {"##.", "#.#", "##.", "#.#", "##."}},
        {'C', {".##", "#..", "#..", "#..", ".##"}}, {'D', {"##.", "#.#", "#.#", "#.#", "##."}},
        {'E', {"###", "#..", "##.", "#..", "###"}}, {'F', {"###", "#..", "##.", "#..", "#.."}},
        {'G', {".##", "#..", "#.#", "#.#", ".##"}}, {'H', {"#.#", "#.#", "###", "#.#", "#.#"}},
        {'I', {"###", ".#.", ".#.", ".#.", "###"}}, {'J', {"..#", "..#", "..#", "#.#", ".#."}},
        {'K', {"#.#", "##.", "#..", "##.", "#.#"}}, {'L', {"#..", "#..", "#..", "#..", "###"}},
        {'M', {"#.#", "###", "###", "#.#", "#.#"}}, {'N', {"###", "#.#", "#.#", "#.#", "#.#"}},
        {'O', {".#.", "#.#", "#.#", "#.#", ".#."}}, {'P', {"##.", "#.#", "##.", "#..", "#.."}},
        {'Q', {".#.", "#.#", "#.#", "##.", ".##"}}, {'R', {"##.", "#.#", "##.", "##.", "#.#"}},
        {'S', {".##", "#..", ".#.", "..#", "##."}}, {'T', {"###", ".#.", ".#.", ".#.", ".#."}},
        {'U', {"#.#", "#.#", "#.#", "#.#", "###"}}, {'V', {"#.#", "#.#", "#.#", "#.#", ".#."}},
        {'W', {"#.#", "#.#", "###", "###", "#.#"}}, {'X', {"#.#", "#.#", ".#.", "#.#", "#.#"}},
        {'Y', {"#.#", "#.#", ".#.", ".#.", ".#."}}, {'Z', {"###", "..#", ".#.", "#..", "###"}},
        {'.', {"...", "...", "...", "...", ".#."}}, {'-', {"...", "...", "###", "...", "..."}},
        {':', {"...", ".#.", "...", ".#.", "..."}}, {'/', {"..#", "..#", ".#.", "#..", "#.."}},
        {'%', {"#.#", "..#", ".#.", "#..", "#.#"}}, {'=', {"...", "###", "...", "###", "..."}},
        {'+', {"...", ".#.", "###", ".#.", "..."}}, {'(', {".#.", "#..", "#..", "#..", ".#."}},
        {')', {".#.", "..#", "..#", "..#", ".#."}}, {0, {nullptr, nullptr, nullptr, nullptr, nullptr}},
    };
    return font;
}

inline const GifGlyph* gif_glyph(char c) {
    if (c >= 'a' && c <= 'z') c = (char)(c - 'a' + 'A');
    for (const GifGlyph* g = gif_font(); g->c; ++g)
        if (g->c == c) return g;
    return nullptr;   // space and anything unknown: blank
}

// Draws `text` at (x, y) (top-left of the first glyph) with `scale` pixels per font dot: the box
// under it (one dot of margin) is darkened to a quarter, the glyphs are drawn in white. Clipped to
// the image.
inline void gif_caption(std::vector<uint8_t>& rgb8, int w, int h, int x, int y, const std::string& text, int scale) {
    const int adv = 4 * scale;
    const int x0 = x - scale, y0 = y - scale, x1 = x + (int)text.size() * adv, y1 = y + 6 * scale;
    for (int py = std::max(0, y0); py < std::min(h, y1); ++py)
        for (int px = std::max(0, x0); px < std::min(w, x1); ++px)
            for (int c = 0; c < 3; ++c) {
                uint8_t& v = rgb8[3 * ((size_t)py * w + px) + c];
                v = (uint8_t)(v >> 2);
            }
    for (size_t k = 0; k < text.size(); ++k) {
        const GifGlyph* g = gif_glyph(text[k]);
        if (!g) continue;
        for (int row = 0; row < 5; ++row)
            for (int col = 0; col < 3; ++col) {
                if (g->rows[row][col] != '#') continue;
                for (int dy = 0; dy < scale; ++dy)
                    for (int dx = 0; dx < scale; ++dx) {
                        const int px = x + (int)k * adv + col * scale + dx, py = y + row * scale + dy;
                        if (px < 0 || py < 0 || px >= w || py >= h) continue;
                        for (int c = 0; c < 3; ++c) rgb8[3 * ((size_t)py * w + px) + c] = 255;
                    }
            }
    }
}

// ---- file ----
class GifWriter {
public:
    GifWriter(const std::string& path, int w, int h) : f_(path, std::ios::binary), w_(w), h_(h) {
        if (w < 1 || h < 1 || w > 65535 || h > 65535) throw std::invalid_argument("gif: image size outside 1..65535");
        if (!f_) throw std::runtime_error("cannot open " + path + " for writing");
        f_.write("GIF89a", 6);
        put16(w);
        put16(h);
        put8(0xF7);   // global colour table, 8-bit colour resolution, 2^(7+1) entries
        put8(0);      // background colour index
        put8(0);      // square pixels
        uint8_t table[768];
        gif_palette(table);
        f_.write(reinterpret_cast<const char*>(table), 768);
        static const uint8_t loop[19] = {0x21, 0xFF, 0x0B, 'N', 'E', 'T', 'S', 'C', 'A', 'P', 'E', '2', '.', '0',
                                         0x03, 0x01, 0x00, 0x00, 0x00};   // loop forever
        f_.write(reinterpret_cast<const char*>(loop), sizeof(loop));
    }

    // One frame of palette indices (w * h), shown for delay_cs hundredths of a second.
    void frame(const std::vector<uint8_t>& idx, int delay_cs) {
        if (idx.size() != (size_t)w_ * h_) throw std::invalid_argument("gif: frame size does not match the image");
        const uint8_t gce[4] = {0x21, 0xF9, 0x04, 0x04};   // graphic control: leave the frame in place
        f_.write(reinterpret_cast<const char*>(gce), 4);
        put16(std::clamp(delay_cs, 0, 65535));
        put8(0);   // no transparent colour
        put8(0);
        put8(0x2C);   // image descriptor: the whole screen, no local table, not interlaced
        put16(0);
        put16(0);
        put16(w_);
        put16(h_);
        put8(0);
        put8(8);   // LZW minimum code size
        const std::vector<uint8_t> data = gif_lzw(idx.data(), idx.size());
        for (size_t pos = 0; pos < data.size(); pos += 255) {
            const size_t n = std::min<size_t>(255, data.size() - pos);
            put8((int)n);
            f_.write(reinterpret_cast<const char*>(data.data() + pos), (std::streamsize)n);
        }
        put8(0);
        ++frames_;
    }

    int frames() const { return frames_; }

    void close() {
        put8(0x3B);
        f_.close();
        if (!f_) throw std::runtime_error("gif: write failed");
    }

private:
    void put8(int v) { f_.put((char)(v & 0xff)); }
    void put16(int v) {
        put8(v);
        put8(v >> 8);
    }
    std::ofstream f_;
    int w_, h_, frames_ = 0;
};

// Exposure from the image's key (Reinhard et al., "Photographic tone reproduction", 2002): the
// geometric mean of the luminance of the lit pixels maps to 0.18, so every scene of an animation
// is visible whatever its absolute radiance. Pixels with no light (the black outside a box) are
// left out of the mean.
inline float gif_auto_exposure(const std::vector<float>& rgb) {
    double sum = 0.0;
    size_t n = 0;
    for (size_t i = 0; i + 2 < rgb.size(); i += 3) {
        const double l = 0.2126 * rgb[i] + 0.7152 * rgb[i + 1] + 0.0722 * rgb[i + 2];
        if (std::isfinite(l) && l > 1e-6) {
            sum += std::log(l);
            ++n;
        }
    }
    if (n == 0) return 1.0f;
    return (float)std::clamp(0.18 / std::exp(sum / (double)n), 0.01, 1000.0);
}

// ---- `wf gif` options ----
// wf gif OUT.gif [--scenes A,B,...] [--width W] [--height H] [--mode megakernel|wavefront|persistent]
//                [--steps 1,2,4,...] [--orbit FRAMES] [--spp N] [--seed S] [--max-depth D]
//                [--exposure X] [--delay CS] [--hold CS] [--final-png-prefix P]
// Progressive (the default): per scene one renderer, frames after each cumulative sample count of
// --steps (strictly increasing), the last frame held for --hold. Orbit (--orbit N, N >= 1): per
// scene N frames with the camera orbited by 360/N degrees each (the viewer's orbit), each a render
// of --spp samples. Exposure: given, or per scene from its last progressive / first orbit frame.
struct GifConfig {
    std::string out;
    std::vector<std::string> scenes = {"cornell", "divergence", "trianglehv"};
    int width = 448, height = 252;
    std::string mode = "megakernel";
    std::vector<int> steps = {1, 2, 4, 8, 16, 32, 64, 128, 256};
    int orbit = 0, spp = 16;
    unsigned seed = 1234;
    int max_depth = 8;
    float exposure = 0.0f;   // 0: automatic
    int delay = 25, hold = 150;
    std::string png_prefix;
};

inline std::string gif_usage() {
    return "Usage: wf gif OUT.gif [--scenes A,B,...] [--width W] [--height H] [--mode megakernel|wavefront|persistent]\n"
           "       [--steps 1,2,4,...] [--orbit FRAMES] [--spp N] [--seed S] [--max-depth D] [--exposure X]\n"
           "       [--delay CS] [--hold CS] [--final-png-prefix P]\n";
}

inline std::vector<std::string> split_commas(const std::string& s) {
    std::vector<std::string> out;
    size_t a = 0;
    for (;;) {
        const size_t b = s.find(',', a);
        out.push_back(s.substr(a, b == std::string::npos ? std::string::npos : b - a));
        if (b == std::string::npos) break;
        a = b + 1;
    }
    return out;
}

inline GifConfig parse_gif_args(int argc, const char* const* argv) {
    GifConfig c;
    if (argc < 3 || std::string(argv[1]) != "gif") throw ConfigError("usage: wf gif OUT.gif [options]");
    c.out = argv[2];
    if (c.out.empty() || c.out[0] == '-') throw ConfigError("the first argument must be the output file");
    for (int i = 3; i < argc; ++i) {
        const std::string f = argv[i];
        auto val = [&]() -> const char* {
            if (i + 1 >= argc) throw ConfigError(f + ": missing value");
            return argv[++i];
        };
        auto num = [&](int lo, int hi) {
            int v;
            if (!parse_int(val(), v)) throw ConfigError(f + ": non-numeric value");
            if (v < lo || v > hi)
                throw ConfigError(f + ": value outside " + std::to_string(lo) + ".." + std::to_string(hi));
            return v;
        };
        if (f == "--scenes") {
            c.scenes = split_commas(val());
            for (const std::string& n : c.scenes)
                if (n.empty()) throw ConfigError("--scenes: empty scene name");
        } else if (f == "--width") c.width = num(1, 4096);
        else if (f == "--height") c.height = num(1, 4096);
        else if (f == "--mode") {
            c.mode = val();
            if (c.mode != "megakernel" && c.mode != "wavefront" && c.mode != "persistent")
                throw ConfigError("--mode: unknown value \"" + c.mode + "\" (a GPU arm: megakernel, wavefront, persistent)");
        } else if (f == "--steps") {
            c.steps.clear();
            for (const std::string& t : split_commas(val())) {
                int v;
                if (!parse_int(t.c_str(), v)) throw ConfigError("--steps: non-numeric value");
                if (v < 1 || v > 65536) throw ConfigError("--steps: value outside 1..65536");
                if (!c.steps.empty() && v <= c.steps.back()) throw ConfigError("--steps: values must increase");
                c.steps.push_back(v);
            }
        } else if (f == "--orbit") c.orbit = num(0, 720);
        else if (f == "--spp") c.spp = num(1, 65536);
        else if (f == "--seed") {
            if (!parse_uint(val(), c.seed)) throw ConfigError("--seed: non-numeric value");
        } else if (f == "--max-depth") c.max_depth = num(1, 64);
        else if (f == "--exposure") {
            const char* v = val();
            char* end = nullptr;
            const double x = std::strtod(v, &end);
            if (!end || *end != '\0' || !(x > 0.0) || !std::isfinite(x)) throw ConfigError("--exposure: not a positive number");
            c.exposure = (float)x;
        } else if (f == "--delay") c.delay = num(1, 65535);
        else if (f == "--hold") c.hold = num(1, 65535);
        else if (f == "--final-png-prefix") c.png_prefix = val();
        else throw ConfigError("unknown flag: " + f);
    }
    return c;
}

}  // namespace wfcli
