// A small JSON reader for the `wf` CLI: report.json validation (src/report.cpp:348-475) and the
// `wf validate-report` command. Values are a tagged tree; objects keep their members in document
// order. The parser is strict RFC 8259: one value per document, no trailing commas, no comments,
// \uXXXX escapes (surrogate pairs included) decoded to UTF-8, nesting limited to kMaxDepth.
#pragma once
#include <cctype>
#include <cstdlib>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

namespace wfcli {

struct JValue {
    enum Kind { Null, Bool, Number, String, Array, Object };
    Kind kind = Null;
    bool b = false;
    double num = 0.0;
    std::string str;
    std::vector<JValue> items;                               // Array
    std::vector<std::pair<std::string, JValue>> members;     // Object, document order

    // Member `key` of an object, or nullptr (also when this is not an object).
    const JValue* get(const std::string& key) const {
        if (kind != Object) return nullptr;
        for (const auto& m : members)
            if (m.first == key) return &m.second;
        return nullptr;
    }
};

struct JsonError : std::runtime_error {
    using std::runtime_error::runtime_error;
};

class JsonReader {
public:
    static constexpr int kMaxDepth = 256;

    explicit JsonReader(const std::string& text) : s_(text) {}

    JValue document() {
        skip_ws();
        JValue v = value(0);
        skip_ws();
        if (i_ != s_.size()) fail("unexpected text after the value");
        return v;
    }

private:
    const std::string& s_;
    size_t i_ = 0;

    [[noreturn]] void fail(const std::string& what) const {
        throw JsonError(what + " at offset " + std::to_string(i_));
    }
    void skip_ws() {
        while (i_ < s_.size() && (s_[i_] == ' ' || s_[i_] == '\t' || s_[i_] == '\n' || s_[i_] == '\r')) ++i_;
    }
    bool take(char c) {
        if (i_ < s_.size() && s_[i_] == c) {
            ++i_;
            return true;
        }
        return false;
    }
    void expect_word(const char* w) {
        for (const char* p = w; *p; ++p)
            if (!take(*p)) fail(std::string("expected '") + w + "'");
    }

    JValue value(int depth) {
        if (depth > kMaxDepth) fail("nesting too deep");
        if (i_ >= s_.size()) fail("unexpected end of input");
        JValue v;
        const char c = s_[i_];
        if (c == '{') {
            ++i_;
            v.kind = JValue::Object;
            skip_ws();
            if (take('}')) return v;
            for (;;) {
                skip_ws();
                if (i_ >= s_.size() || s_[i_] != '"') fail("expected a member name");
                std::string key = string_body();
                skip_ws();
                if (!take(':')) fail("expected ':'");
                skip_ws();
                v.members.emplace_back(std::move(key), value(depth + 1));
                skip_ws();
                if (take('}')) return v;
                if (!take(',')) fail("expected ',' or '}'");
            }
        }
        if (c == '[') {
            ++i_;
            v.kind = JValue::Array;
            skip_ws();
            if (take(']')) return v;
            for (;;) {
                skip_ws();
                v.items.push_back(value(depth + 1));
                skip_ws();
                if (take(']')) return v;
                if (!take(',')) fail("expected ',' or ']'");
            }
        }
        if (c == '"') {
            v.kind = JValue::String;
            v.str = string_body();
            return v;
        }
        if (c == 't') {
            expect_word("true");
            v.kind = JValue::Bool;
            v.b = true;
            return v;
        }
        if (c == 'f') {
            expect_word("false");
            v.kind = JValue::Bool;
            return v;
        }
        if (c == 'n') {
            expect_word("null");
            return v;
        }
        v.kind = JValue::Number;
        v.num = number();
        return v;
    }

    // -? (0 | [1-9][0-9]*) (. [0-9]+)? ([eE] [+-]? [0-9]+)?
    double number() {
        const size_t start = i_;
        auto digits = [&]() {
            size_t n = 0;
            while (i_ < s_.size() && std::isdigit((unsigned char)s_[i_])) ++i_, ++n;
            return n;
        };
        take('-');
        if (take('0')) {
            if (i_ < s_.size() && std::isdigit((unsigned char)s_[i_])) fail("leading zero");
        } else if (digits() == 0) {
            fail("invalid value");
        }
        if (take('.') && digits() == 0) fail("digit expected after '.'");
        if (i_ < s_.size() && (s_[i_] == 'e' || s_[i_] == 'E')) {
            ++i_;
            if (!take('+')) take('-');
            if (digits() == 0) fail("digit expected in exponent");
        }
        const std::string lit = s_.substr(start, i_ - start);
        return std::strtod(lit.c_str(), nullptr);
    }

    unsigned hex4() {
        if (i_ + 4 > s_.size()) fail("truncated \\u escape");
        unsigned v = 0;
        for (int k = 0; k < 4; ++k) {
            const char h = s_[i_++];
            v <<= 4;
            if (h >= '0' && h <= '9') v |= (unsigned)(h - '0');
            else if (h >= 'a' && h <= 'f') v |= (unsigned)(h - 'a' + 10);
            else if (h >= 'A' && h <= 'F') v |= (unsigned)(h - 'A' + 10);
            else fail("bad hex digit in \\u escape");
        }
        return v;
    }

    static void utf8(std::string& o, unsigned cp) {
        if (cp < 0x80) {
            o += (char)cp;
        } else if (cp < 0x800) {
            o += (char)(0xC0 | (cp >> 6));
            o += (char)(0x80 | (cp & 0x3F));
        } else if (cp < 0x10000) {
            o += (char)(0xE0 | (cp >> 12));
            o += (char)(0x80 | ((cp >> 6) & 0x3F));
            o += (char)(0x80 | (cp & 0x3F));
        } else {
            o += (char)(0xF0 | (cp >> 18));
            o += (char)(0x80 | ((cp >> 12) & 0x3F));
            o += (char)(0x80 | ((cp >> 6) & 0x3F));
            o += (char)(0x80 | (cp & 0x3F));
        }
    }

    // A string starting at the opening quote; returns its decoded text.
    std::string string_body() {
        ++i_;
        std::string o;
        for (;;) {
            if (i_ >= s_.size()) fail("unterminated string");
            const char c = s_[i_++];
            if (c == '"') return o;
            if ((unsigned char)c < 0x20) fail("control character in string");
            if (c != '\\') {
                o += c;
                continue;
            }
            if (i_ >= s_.size()) fail("unterminated escape");
            const char e = s_[i_++];
            switch (e) {
                case '"': o += '"'; break;
                case '\\': o += '\\'; break;
                case '/': o += '/'; break;
                case 'b': o += '\b'; break;
                case 'f': o += '\f'; break;
                case 'n': o += '\n'; break;
                case 'r': o += '\r'; break;
                case 't': o += '\t'; break;
                case 'u': {
                    unsigned cp = hex4();
                    if (cp >= 0xD800 && cp <= 0xDBFF) {
                        if (!(take('\\') && take('u'))) fail("unpaired high surrogate");
                        const unsigned lo = hex4();
                        if (lo < 0xDC00 || lo > 0xDFFF) fail("bad low surrogate");
                        cp = 0x10000 + ((cp - 0xD800) << 10) + (lo - 0xDC00);
                    } else if (cp >= 0xDC00 && cp <= 0xDFFF) {
                        fail("unpaired low surrogate");
                    }
                    utf8(o, cp);
                    break;
                }
                default: fail("bad escape");
            }
        }
    }
};

inline JValue json_parse(const std::string& text) { return JsonReader(text).document(); }

// Checks a report.json document against schema version 1 (the rules of src/report.cpp:348-475):
// every required key present with its type, scene.hash and outputs.pfm_hash 16 hex digits, a
// non-empty timings_ms.runs of numbers, min <= median <= max. Returns "" when the document is
// valid, else a description of the first problem found. Keys beyond the schema (such as the
// `mi355x` extension object) are allowed.
inline std::string validate_report_json(const std::string& text) {
    JValue root;
    try {
        root = json_parse(text);
    } catch (const JsonError& e) {
        return std::string("not valid JSON: ") + e.what();
    }
    if (root.kind != JValue::Object) return "the document is not an object";
    auto is = [](const JValue* v, JValue::Kind k) { return v && v->kind == k; };
    auto hex16 = [](const std::string& s) {
        if (s.size() != 16) return false;
        for (char c : s)
            if (!std::isxdigit((unsigned char)c)) return false;
        return true;
    };
    auto object = [&](const JValue& parent, const std::string& key, const JValue** out) -> std::string {
        const JValue* v = parent.get(key);
        if (!v) return "missing " + key;
        if (v->kind != JValue::Object) return key + " is not an object";
        *out = v;
        return "";
    };
    auto field = [&](const JValue& obj, const std::string& where, const std::string& key, JValue::Kind k,
                     const char* kind_name) -> std::string {
        const JValue* v = obj.get(key);
        if (!v) return where + " has no " + key;
        if (v->kind != k) return where + "." + key + " is not a " + kind_name;
        return "";
    };
    std::string p;

    const JValue* sv = root.get("schema_version");
    if (!sv) return "missing schema_version";
    if (sv->kind != JValue::Number) return "schema_version is not a number";
    if ((int)sv->num != 1) return "schema_version is not 1";

    const JValue* gr = root.get("git_revision");
    if (!gr) return "missing git_revision";
    if (gr->kind != JValue::Null && gr->kind != JValue::String) return "git_revision is neither null nor a string";

    const JValue* scene = nullptr;
    if (!(p = object(root, "scene", &scene)).empty()) return p;
    if (!(p = field(*scene, "scene", "name", JValue::String, "string")).empty()) return p;
    if (!(p = field(*scene, "scene", "hash", JValue::String, "string")).empty()) return p;
    if (!hex16(scene->get("hash")->str)) return "scene.hash is not 16 hex digits";

    const JValue* config = nullptr;
    if (!(p = object(root, "config", &config)).empty()) return p;
    for (const char* k : {"width", "height", "spp", "seed", "max_depth"})
        if (!(p = field(*config, "config", k, JValue::Number, "number")).empty()) return p;
    if (!(p = field(*config, "config", "mode", JValue::String, "string")).empty()) return p;

    const JValue* device = nullptr;
    if (!(p = object(root, "device", &device)).empty()) return p;
    for (const char* k : {"name", "driver_version", "cuda_runtime_version", "cuda_driver_version", "compute_capability"})
        if (!(p = field(*device, "device", k, JValue::String, "string")).empty()) return p;
    if (!(p = field(*device, "device", "total_memory_bytes", JValue::Number, "number")).empty()) return p;

    const JValue* t = nullptr;
    if (!(p = object(root, "timings_ms", &t)).empty()) return p;
    if (!(p = field(*t, "timings_ms", "runs", JValue::Array, "array")).empty()) return p;
    const JValue* runs = t->get("runs");
    if (runs->items.empty()) return "timings_ms.runs is empty";
    for (const JValue& r : runs->items)
        if (r.kind != JValue::Number) return "timings_ms.runs holds a value that is not a number";
    for (const char* k : {"min", "median", "p10", "p90", "max"})
        if (!(p = field(*t, "timings_ms", k, JValue::Number, "number")).empty()) return p;
    const double lo = t->get("min")->num, med = t->get("median")->num, hi = t->get("max")->num;
    if (!(lo <= med && med <= hi)) return "timings_ms does not satisfy min <= median <= max";

    const JValue* tp = nullptr;
    if (!(p = object(root, "throughput", &tp)).empty()) return p;
    if (!(p = field(*tp, "throughput", "samples_per_second", JValue::Number, "number")).empty()) return p;
    const JValue* rps = tp->get("rays_per_second");
    if (!rps) return "throughput has no rays_per_second";
    if (!is(rps, JValue::Null) && !is(rps, JValue::Number)) return "throughput.rays_per_second is neither null nor a number";

    const JValue* pv = root.get("peak_vram_bytes");
    if (!pv) return "missing peak_vram_bytes";
    if (pv->kind != JValue::Number) return "peak_vram_bytes is not a number";

    const JValue* q = nullptr;
    if (!(p = object(root, "queue_high_water", &q)).empty()) return p;
    for (const char* k : {"lambert", "metal", "dielectric", "shadow", "active"})
        if (!(p = field(*q, "queue_high_water", k, JValue::Number, "number")).empty()) return p;

    const JValue* out = nullptr;
    if (!(p = object(root, "outputs", &out)).empty()) return p;
    for (const char* k : {"png", "pfm", "pfm_hash"})
        if (!(p = field(*out, "outputs", k, JValue::String, "string")).empty()) return p;
    if (!hex16(out->get("pfm_hash")->str)) return "outputs.pfm_hash is not 16 hex digits";

    const JValue* val = nullptr;
    if (!(p = object(root, "validity", &val)).empty()) return p;
    for (const char* k : {"nonfinite_pixels", "queue_overflows", "invalid_hits"})
        if (!(p = field(*val, "validity", k, JValue::Number, "number")).empty()) return p;
    return "";
}

}  // namespace wfcli
