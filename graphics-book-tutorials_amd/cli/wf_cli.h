// Host-side support for the `wf` CLI: strict flag parsing (src/config.cpp:62-175), percentiles
// (src/report.cpp:12-35), report.json schema v1 (include/wf/report.hpp:41-49, src/report.cpp:37-346)
// with an `mi355x` extension object, and PNG (stored deflate) / PFM output (src/image_io.cpp).
#pragma once
#include <algorithm>
#include <charconv>
#include <climits>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

namespace wfcli {

struct ConfigError : std::runtime_error {
    using std::runtime_error::runtime_error;
};

struct Config {
    std::string scene = "cornell";
    int width = 1280, height = 720, spp = 64;
    unsigned seed = 1234;
    int max_depth = 8;
    std::string mode = "wavefront";
    std::string report_path, out_prefix;
    int threads = 0, reps = 1, warmup = 0;
};

inline std::string usage() {
    return "Usage: wf render --scene NAME --width W --height H --spp N --seed S --max-depth D\n"
           "       --mode {cpu,megakernel,wavefront,persistent} --report PATH --out PREFIX --threads N --reps N "
           "--warmup N\n"
           "Scenes: cornell divergence trianglehv furnace cornell555 sphere_grid heightfield\n"
           "Also:  wf validate-report REPORT.json | wf imgdiff A.pfm B.pfm [--tol T] | wf gif OUT.gif [options]\n";
}

inline bool parse_int(const char* s, int& out) {
    if (!s || !*s) return false;
    const char* e = s + std::strlen(s);
    auto r = std::from_chars(s, e, out);
    return r.ec == std::errc() && r.ptr == e;
}
inline bool parse_uint(const char* s, unsigned& out) {
    if (!s || !*s) return false;
    const char* e = s + std::strlen(s);
    auto r = std::from_chars(s, e, out);
    return r.ec == std::errc() && r.ptr == e;
}

// Strict parser: unknown flag, missing or non-numeric value, out-of-range bounds -> ConfigError
// naming the flag (bounds of src/config.cpp:62-175).
inline Config parse_args(int argc, const char* const* argv) {
    Config c;
    if (argc < 2 || std::string(argv[1]) != "render") throw ConfigError("argv[1] must be \"render\"");
    for (int i = 2; i < argc; ++i) {
        std::string f = argv[i];
        auto val = [&]() -> const char* {
            if (i + 1 >= argc) throw ConfigError(f + ": missing value");
            return argv[++i];
        };
        auto num = [&](int lo, int hi, const char* what) {
            int v;
            if (!parse_int(val(), v)) throw ConfigError(f + ": non-numeric value");
            if (v < lo || v > hi) throw ConfigError(f + ": " + what);
            return v;
        };
        if (f == "--scene") c.scene = val();
        else if (f == "--width") c.width = num(1, 8192, "value outside 1..8192");
        else if (f == "--height") c.height = num(1, 8192, "value outside 1..8192");
        else if (f == "--spp") c.spp = num(1, INT_MAX, "value < 1");
        else if (f == "--seed") {
            if (!parse_uint(val(), c.seed)) throw ConfigError("--seed: non-numeric value");
        } else if (f == "--max-depth") c.max_depth = num(1, 64, "value outside 1..64");
        else if (f == "--mode") {
            c.mode = val();
            if (c.mode != "cpu" && c.mode != "megakernel" && c.mode != "wavefront" && c.mode != "persistent")
                throw ConfigError("--mode: unknown value \"" + c.mode + "\"");
        } else if (f == "--report") c.report_path = val();
        else if (f == "--out") c.out_prefix = val();
        else if (f == "--threads") {
            int v;
            if (!parse_int(val(), v)) throw ConfigError("--threads: non-numeric value");
            c.threads = v;
        } else if (f == "--reps") c.reps = num(1, INT_MAX, "value < 1");
        else if (f == "--warmup") c.warmup = num(0, INT_MAX, "value < 0");
        else throw ConfigError("unknown flag: " + f);
    }
    return c;
}

// Linear interpolation between order statistics (src/report.cpp:12-35).
inline double percentile(std::vector<double> v, double p) {
    if (v.empty()) throw std::invalid_argument("percentile of empty input");
    std::sort(v.begin(), v.end());
    if (p <= 0) return v.front();
    if (p >= 1) return v.back();
    double x = p * (double)(v.size() - 1);
    size_t i = (size_t)x;
    double f = x - (double)i;
    if (i + 1 >= v.size()) return v.back();
    return v[i] + (v[i + 1] - v[i]) * f;
}

inline uint64_t fnv1a64(const void* p, size_t n) {
    uint64_t h = 14695981039346656037ULL;
    const unsigned char* b = (const unsigned char*)p;
    for (size_t i = 0; i < n; ++i) {
        h ^= b[i];
        h *= 1099511628211ULL;
    }
    return h;
}

inline std::string hex64(uint64_t v) {
    char buf[17];
    std::snprintf(buf, sizeof buf, "%016llx", (unsigned long long)v);
    return buf;
}

inline std::string jnum(double v) {
    if (!std::isfinite(v)) return "null";
    std::ostringstream o;
    o.precision(17);
    o << v;
    return o.str();
}

inline std::string jstr(const std::string& s) {
    std::string o = "\"";
    for (char ch : s) {
        if (ch == '"' || ch == '\\') o += '\\', o += ch;
        else if ((unsigned char)ch < 0x20) {
            char b[8];
            std::snprintf(b, sizeof b, "\\u%04x", ch);
            o += b;
        } else o += ch;
    }
    return o + "\"";
}

// ---- images ----
inline uint32_t crc32(const uint8_t* d, size_t n, uint32_t c = 0xFFFFFFFFu) {
    for (size_t i = 0; i < n; ++i) {
        c ^= d[i];
        for (int k = 0; k < 8; ++k) c = (c >> 1) ^ (0xEDB88320u & (0u - (c & 1u)));
    }
    return c;
}

// Reinhard x/(1+x) then the sRGB transfer, 8 bits (src/film.cpp:8-32).
inline std::vector<uint8_t> tonemap_srgb8(const std::vector<float>& rgb, float exposure = 1.0f) {
    std::vector<uint8_t> out(rgb.size());
    for (size_t i = 0; i < rgb.size(); ++i) {
        float x = rgb[i] * exposure;
        float t = x / (1.0f + x);
        float s = t <= 0.0031308f ? 12.92f * t : 1.055f * std::pow(t, 1.0f / 2.4f) - 0.055f;
        s = std::max(0.0f, std::min(1.0f, s));
        out[i] = (uint8_t)(s * 255.0f + 0.5f);
    }
    return out;
}

// PNG with zlib stored (uncompressed) deflate blocks.
inline void write_png(const std::string& path, int w, int h, const std::vector<uint8_t>& rgb8) {
    std::vector<uint8_t> raw;
    raw.reserve((size_t)(3 * w + 1) * h);
    for (int y = 0; y < h; ++y) {
        raw.push_back(0);
        raw.insert(raw.end(), rgb8.begin() + (size_t)3 * w * y, rgb8.begin() + (size_t)3 * w * (y + 1));
    }
    std::vector<uint8_t> z = {0x78, 0x01};
    uint32_t a = 1, b = 0;
    for (uint8_t v : raw) {
        a = (a + v) % 65521;
        b = (b + a) % 65521;
    }
    for (size_t off = 0; off < raw.size() || off == 0; off += 65535) {
        size_t n = std::min<size_t>(65535, raw.size() - off);
        bool last = off + n >= raw.size();
        z.push_back(last ? 1 : 0);
        z.push_back(n & 0xff);
        z.push_back(n >> 8);
        z.push_back(~n & 0xff);
        z.push_back((~n >> 8) & 0xff);
        z.insert(z.end(), raw.begin() + off, raw.begin() + off + n);
        if (last) break;
    }
    uint32_t ad = (b << 16) | a;
    for (int s = 24; s >= 0; s -= 8) z.push_back((ad >> s) & 0xff);
    std::ofstream f(path, std::ios::binary);
    if (!f) throw std::runtime_error("cannot write " + path);
    const uint8_t sig[8] = {0x89, 'P', 'N', 'G', '\r', '\n', 0x1a, '\n'};
    f.write((const char*)sig, 8);
    auto chunk = [&](const char* type, const std::vector<uint8_t>& data) {
        uint8_t len[4] = {(uint8_t)(data.size() >> 24), (uint8_t)(data.size() >> 16), (uint8_t)(data.size() >> 8),
                          (uint8_t)data.size()};
        f.write((const char*)len, 4);
        std::vector<uint8_t> td(type, type + 4);
        td.insert(td.end(), data.begin(), data.end());
        f.write((const char*)td.data(), (std::streamsize)td.size());
        uint32_t c = crc32(td.data(), td.size()) ^ 0xFFFFFFFFu;
        uint8_t cb[4] = {(uint8_t)(c >> 24), (uint8_t)(c >> 16), (uint8_t)(c >> 8), (uint8_t)c};
        f.write((const char*)cb, 4);
    };
    std::vector<uint8_t> ihdr = {(uint8_t)(w >> 24), (uint8_t)(w >> 16), (uint8_t)(w >> 8), (uint8_t)w,
                                 (uint8_t)(h >> 24), (uint8_t)(h >> 16), (uint8_t)(h >> 8), (uint8_t)h,
                                 8, 2, 0, 0, 0};
    chunk("IHDR", ihdr);
    chunk("IDAT", z);
    chunk("IEND", {});
}

// PFM (little endian, bottom row first).
inline void write_pfm(const std::string& path, int w, int h, const std::vector<float>& rgb) {
    std::ofstream f(path, std::ios::binary);
    if (!f) throw std::runtime_error("cannot write " + path);
    f << "PF\n" << w << " " << h << "\n-1.0\n";
    for (int y = h - 1; y >= 0; --y) f.write((const char*)(rgb.data() + (size_t)3 * w * y), (std::streamsize)(12 * w));
}

// Reads a PFM image into top-row-first RGB floats (src/image_io.cpp:188-250). "PF" holds three
// channels, "Pf" one (returned as grey RGB); the scale's sign gives the byte order (negative:
// little endian). The header is magic, width, height and scale separated by whitespace, then
// exactly one whitespace byte before the rows, bottom row first. Errors name the problem.
inline std::vector<float> read_pfm(const std::string& path, int* width, int* height) {
    std::ifstream f(path, std::ios::binary);
    if (!f) throw std::runtime_error("read_pfm: cannot open " + path);
    std::string magic;
    f >> magic;
    if (magic != "PF" && magic != "Pf") throw std::runtime_error("read_pfm: not a PFM file");
    const int channels = magic == "PF" ? 3 : 1;
    long long w = 0, h = 0;
    double scale = 0.0;
    if (!(f >> w >> h) || w <= 0 || h <= 0 || w > (1 << 16) || h > (1 << 16))
        throw std::runtime_error("read_pfm: invalid dimensions");
    if (!(f >> scale) || scale == 0.0 || !std::isfinite(scale)) throw std::runtime_error("read_pfm: bad scale");
    const int sep = f.get();
    if (sep != '\n' && sep != '\r' && sep != ' ' && sep != '\t') throw std::runtime_error("read_pfm: bad header");
    const bool big = scale > 0.0;
    const size_t row_bytes = (size_t)w * channels * 4;
    std::vector<unsigned char> row(row_bytes);
    std::vector<float> rgb((size_t)3 * w * h);
    for (long long y = h - 1; y >= 0; --y) {
        f.read((char*)row.data(), (std::streamsize)row_bytes);
        if (f.gcount() != (std::streamsize)row_bytes) throw std::runtime_error("read_pfm: truncated file");
        for (long long x = 0; x < w; ++x)
            for (int c = 0; c < 3; ++c) {
                const unsigned char* b = row.data() + ((size_t)x * channels + (channels == 3 ? c : 0)) * 4;
                const uint32_t u = big ? ((uint32_t)b[0] << 24 | (uint32_t)b[1] << 16 | (uint32_t)b[2] << 8 | b[3])
                                       : ((uint32_t)b[3] << 24 | (uint32_t)b[2] << 16 | (uint32_t)b[1] << 8 | b[0]);
                float v;
                std::memcpy(&v, &u, 4);
                rgb[((size_t)y * w + x) * 3 + c] = v;
            }
    }
    *width = (int)w;
    *height = (int)h;
    return rgb;
}

// Difference of two images of equal size (tools/imgdiff.cpp): the largest relative difference
// |a-b| / max(|a|,|b|) over channel pairs where that maximum exceeds 1e-20, and the relative
// difference of the channel sums |sum a - sum b| / sum a (0 when sum a <= 1e-12). A non-finite
// value must compare equal to its counterpart (the same infinity); a NaN never does.
struct ImageDiff {
    double max_abs_rel = 0.0;
    double mean_rel = 0.0;
    bool nonfinite_mismatch = false;
};

inline ImageDiff image_diff(const std::vector<float>& a, const std::vector<float>& b) {
    if (a.size() != b.size()) throw std::invalid_argument("image_diff: sizes differ");
    ImageDiff d;
    double sa = 0.0, sb = 0.0;
    for (size_t i = 0; i < a.size(); ++i) {
        const float x = a[i], y = b[i];
        if (!std::isfinite(x) || !std::isfinite(y)) {
            if (!(x == y)) d.nonfinite_mismatch = true;
            continue;
        }
        sa += x;
        sb += y;
        const double m = std::max(std::fabs((double)x), std::fabs((double)y));
        if (m > 1e-20) d.max_abs_rel = std::max(d.max_abs_rel, std::fabs((double)x - (double)y) / m);
    }
    if (sa > 1e-12) d.mean_rel = std::fabs(sa - sb) / sa;
    return d;
}

}  // namespace wfcli
