// gltf.cpp -- glTF 2.0 scene import (.glb binary container and .gltf JSON) into the flat
// scene records the renderer uploads: behaviour of GltfLoader::load (src/scene/GltfLoader.cpp:79-363).
//
// The reference hands the file to tinygltf; here a small JSON reader and the GLB chunk format
// are implemented directly, and every offset read from the file is bounds-checked (the file is
// untrusted input).  What is kept from the reference:
//   - .glb -> binary container, anything else -> JSON (GltfLoader.cpp:85-88); scene name = stem;
//   - one material per glTF material: baseColorFactor, metallicFactor, roughnessFactor,
//     emissiveFactor with glTF defaults (1,1,1,1 / 1 / 1 / 0), ior 1.5, texture -> image index
//     for base colour, metallic-roughness, normal and emissive (GltfLoader.cpp:101-155);
//   - one mesh per TRIANGLES primitive that has POSITION; NORMAL default (0,1,0), TEXCOORD_0
//     default (0,0), TANGENT default (1,0,0,1); indices u8 / u16 / u32, or 0..n-1 when absent;
//     material default 0; name = mesh name + "_<i>" when the mesh has several primitives
//     (GltfLoader.cpp:199-323);
//   - instances from a depth-first walk of the default scene (or scene 0), world = parent *
//     local, one instance per usable primitive of a node's mesh (GltfLoader.cpp:325-355);
//   - only the base-colour texture reaches the device material record (Scene.cpp:426-437);
//     no kernel samples textures, so images are described (size, type), not decoded.
// Where this differs, it follows the glTF 2.0 specification: node.matrix is column-major
// (GltfLoader.cpp:36-41 transposes it), accessor byteStride is honoured, a file without
// materials gets one default material so that material index 0 exists.
#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <functional>
#include <stdexcept>

#include "rr_host.h"

namespace rrh::scene {

namespace {

[[noreturn]] void fail(const std::string& msg) { throw std::runtime_error("glTF: " + msg); }

// ---- JSON ----------------------------------------------------------------------------------
struct Json {
    enum class T { Null, Bool, Num, Str, Arr, Obj } t = T::Null;
    bool b = false;
    double n = 0.0;
    std::string s;
    std::vector<Json> a;
    std::vector<std::pair<std::string, Json>> o;

    const Json* get(const char* key) const {
        if (t != T::Obj) return nullptr;
        for (const auto& kv : o)
            if (kv.first == key) return &kv.second;
        return nullptr;
    }
};

class JsonReader {
public:
    JsonReader(const char* p, size_t n) : p_(p), e_(p + n) {}
    Json parse() {
        if (e_ - p_ >= 3 && (unsigned char)p_[0] == 0xEF && (unsigned char)p_[1] == 0xBB && (unsigned char)p_[2] == 0xBF)
            p_ += 3;  // UTF-8 byte order mark
        Json v = value(0);
        ws();
        if (p_ != e_) fail("trailing characters after the JSON document");
        return v;
    }

private:
    const char* p_;
    const char* e_;

    void ws() {
        while (p_ < e_ && (*p_ == ' ' || *p_ == '\t' || *p_ == '\n' || *p_ == '\r')) ++p_;
    }
    bool lit(const char* w) {
        const size_t n = std::strlen(w);
        if ((size_t)(e_ - p_) < n || std::memcmp(p_, w, n) != 0) return false;
        p_ += n;
        return true;
    }
    Json value(int depth) {
        if (depth > 128) fail("JSON nested too deeply");
        ws();
        if (p_ >= e_) fail("unexpected end of JSON");
        Json v;
        const char c = *p_;
        if (c == '{') {
            v.t = Json::T::Obj;
            ++p_;
            ws();
            if (p_ < e_ && *p_ == '}') { ++p_; return v; }
            for (;;) {
                ws();
                if (p_ >= e_ || *p_ != '"') fail("expected an object key");
                std::string k = str();
                ws();
                if (p_ >= e_ || *p_ != ':') fail("expected ':' in object");
                ++p_;
                v.o.emplace_back(std::move(k), value(depth + 1));
                ws();
                if (p_ < e_ && *p_ == ',') { ++p_; continue; }
                if (p_ < e_ && *p_ == '}') { ++p_; return v; }
                fail("expected ',' or '}' in object");
            }
        }
        if (c == '[') {
            v.t = Json::T::Arr;
            ++p_;
            ws();
            if (p_ < e_ && *p_ == ']') { ++p_; return v; }
            for (;;) {
                v.a.push_back(value(depth + 1));
                ws();
                if (p_ < e_ && *p_ == ',') { ++p_; continue; }
                if (p_ < e_ && *p_ == ']') { ++p_; return v; }
                fail("expected ',' or ']' in array");
            }
        }
        if (c == '"') { v.t = Json::T::Str; v.s = str(); return v; }
        if (lit("true")) { v.t = Json::T::Bool; v.b = true; return v; }
        if (lit("false")) { v.t = Json::T::Bool; return v; }
        if (lit("null")) return v;
        v.t = Json::T::Num;
        v.n = number();
        return v;
    }
    double number() {
        const char* s = p_;
        if (p_ < e_ && *p_ == '-') ++p_;
        auto digits = [&] {
            const char* d = p_;
            while (p_ < e_ && *p_ >= '0' && *p_ <= '9') ++p_;
            return p_ > d;
        };
        if (!digits()) fail("bad JSON number");
        if (p_ < e_ && *p_ == '.') { ++p_; if (!digits()) fail("bad JSON number"); }
        if (p_ < e_ && (*p_ == 'e' || *p_ == 'E')) {
            ++p_;
            if (p_ < e_ && (*p_ == '+' || *p_ == '-')) ++p_;
            if (!digits()) fail("bad JSON number");
        }
        return std::strtod(std::string(s, p_).c_str(), nullptr);
    }
    static void utf8(std::string& out, uint32_t cp) {
        if (cp < 0x80) out += (char)cp;
        else if (cp < 0x800) { out += (char)(0xC0 | (cp >> 6)); out += (char)(0x80 | (cp & 0x3F)); }
        else if (cp < 0x10000) {
            out += (char)(0xE0 | (cp >> 12)); out += (char)(0x80 | ((cp >> 6) & 0x3F)); out += (char)(0x80 | (cp & 0x3F));
        } else {
            out += (char)(0xF0 | (cp >> 18)); out += (char)(0x80 | ((cp >> 12) & 0x3F));
            out += (char)(0x80 | ((cp >> 6) & 0x3F)); out += (char)(0x80 | (cp & 0x3F));
        }
    }
    uint32_t hex4() {
        if (e_ - p_ < 4) fail("bad \\u escape");
        uint32_t v = 0;
        for (int i = 0; i < 4; ++i) {
            const char h = *p_++;
            v <<= 4;
            if (h >= '0' && h <= '9') v |= (uint32_t)(h - '0');
            else if (h >= 'a' && h <= 'f') v |= (uint32_t)(h - 'a' + 10);
            else if (h >= 'A' && h <= 'F') v |= (uint32_t)(h - 'A' + 10);
            else fail("bad \\u escape");
        }
        return v;
    }
    std::string str() {
        ++p_;  // opening quote
        std::string out;
        while (p_ < e_ && *p_ != '"') {
            char c = *p_++;
            if ((unsigned char)c < 0x20) fail("control character in JSON string");
            if (c != '\\') { out += c; continue; }
            if (p_ >= e_) break;
            c = *p_++;
            switch (c) {
                case '"': out += '"'; break;
                case '\\': out += '\\'; break;
                case '/': out += '/'; break;
                case 'b': out += '\b'; break;
                case 'f': out += '\f'; break;
                case 'n': out += '\n'; break;
                case 'r': out += '\r'; break;
                case 't': out += '\t'; break;
                case 'u': {
                    uint32_t cp = hex4();
                    if (cp >= 0xD800 && cp < 0xDC00 && e_ - p_ >= 6 && p_[0] == '\\' && p_[1] == 'u') {
                        p_ += 2;
                        const uint32_t lo = hex4();
                        if (lo < 0xDC00 || lo > 0xDFFF) fail("bad surrogate pair");
                        cp = 0x10000 + ((cp - 0xD800) << 10) + (lo - 0xDC00);
                    }
                    utf8(out, cp);
                    break;
                }
                default: fail("bad escape in JSON string");
            }
        }
        if (p_ >= e_) fail("unterminated JSON string");
        ++p_;  // closing quote
        return out;
    }
};

// typed field access with glTF defaults
double num(const Json& j, const char* k, double def) {
    const Json* v = j.get(k);
    if (!v) return def;
    if (v->t != Json::T::Num) fail(std::string("'") + k + "' is not a number");
    return v->n;
}
int64_t idx(const Json& j, const char* k, int64_t def = -1) {
    const Json* v = j.get(k);
    if (!v) return def;
    if (v->t != Json::T::Num || v->n < 0 || v->n != std::floor(v->n) || v->n > 4294967295.0)
        fail(std::string("'") + k + "' is not a non-negative integer");
    return (int64_t)v->n;
}
int64_t req(const Json& j, const char* k) {
    const int64_t v = idx(j, k);
    if (v < 0) fail(std::string("missing '") + k + "'");
    return v;
}
std::string text(const Json& j, const char* k) {
    const Json* v = j.get(k);
    if (!v) return {};
    if (v->t != Json::T::Str) fail(std::string("'") + k + "' is not a string");
    return v->s;
}
const std::vector<Json>& list(const Json& j, const char* k) {
    static const std::vector<Json> empty;
    const Json* v = j.get(k);
    if (!v) return empty;
    if (v->t != Json::T::Arr) fail(std::string("'") + k + "' is not an array");
    return v->a;
}
std::vector<double> numbers(const Json& j, const char* k, size_t n) {
    const auto& a = list(j, k);
    if (a.empty()) return {};
    if (a.size() != n) fail(std::string("'") + k + "' must hold " + std::to_string(n) + " numbers");
    std::vector<double> out(n);
    for (size_t i = 0; i < n; ++i) {
        if (a[i].t != Json::T::Num) fail(std::string("'") + k + "' holds a non-number");
        out[i] = a[i].n;
    }
    return out;
}

// ---- files, URIs -----------------------------------------------------------------------------
std::vector<uint8_t> read_file(const std::string& path) {
    std::ifstream f(path, std::ios::binary);
    if (!f) fail("cannot open '" + path + "'");
    f.seekg(0, std::ios::end);
    const std::streamoff n = f.tellg();
    if (n < 0) fail("cannot read '" + path + "'");
    f.seekg(0, std::ios::beg);
    std::vector<uint8_t> out((size_t)n);
    if (n > 0 && !f.read(reinterpret_cast<char*>(out.data()), n)) fail("cannot read '" + path + "'");
    return out;
}

std::vector<uint8_t> base64(const std::string& s, size_t from) {
    auto val = [](char c) -> int {
        if (c >= 'A' && c <= 'Z') return c - 'A';
        if (c >= 'a' && c <= 'z') return c - 'a' + 26;
        if (c >= '0' && c <= '9') return c - '0' + 52;
        if (c == '+') return 62;
        if (c == '/') return 63;
        return -1;
    };
    std::vector<uint8_t> out;
    uint32_t acc = 0;
    int bits = 0;
    for (size_t i = from; i < s.size(); ++i) {
        if (s[i] == '=') break;
        const int v = val(s[i]);
        if (v < 0) fail("bad base64 data URI");
        acc = (acc << 6) | (uint32_t)v;
        bits += 6;
        if (bits >= 8) {
            bits -= 8;
            out.push_back((uint8_t)(acc >> bits));
        }
    }
    return out;
}

std::string percent_decode(const std::string& u) {
    std::string out;
    for (size_t i = 0; i < u.size(); ++i) {
        if (u[i] == '%' && i + 2 < u.size() && std::isxdigit((unsigned char)u[i + 1]) &&
            std::isxdigit((unsigned char)u[i + 2])) {
            out += (char)std::stoi(u.substr(i + 1, 2), nullptr, 16);
            i += 2;
        } else {
            out += u[i];
        }
    }
    return out;
}

// data: URI (base64) or a file relative to the .gltf
std::vector<uint8_t> load_uri(const std::string& uri, const std::string& dir) {
    if (uri.rfind("data:", 0) == 0) {
        const size_t comma = uri.find(',');
        if (comma == std::string::npos || uri.rfind(";base64", comma) == std::string::npos)
            fail("only base64 data URIs are supported");
        return base64(uri, comma + 1);
    }
    if (uri.find("://") != std::string::npos) fail("remote URI '" + uri + "' not supported");
    return read_file(dir + percent_decode(uri));
}

uint32_t rd32(const uint8_t* p) { uint32_t v; std::memcpy(&v, p, 4); return v; }
uint32_t be32(const uint8_t* p) { return (uint32_t)p[0] << 24 | (uint32_t)p[1] << 16 | (uint32_t)p[2] << 8 | p[3]; }

// width / height from a PNG IHDR or a JPEG SOFn marker; 0 when the bytes are neither
void image_size(const std::vector<uint8_t>& b, uint32_t& w, uint32_t& h) {
    w = h = 0;
    static const uint8_t png[8] = {0x89, 'P', 'N', 'G', '\r', '\n', 0x1A, '\n'};
    if (b.size() >= 24 && std::memcmp(b.data(), png, 8) == 0 && std::memcmp(b.data() + 12, "IHDR", 4) == 0) {
        w = be32(b.data() + 16);
        h = be32(b.data() + 20);
        return;
    }
    if (b.size() >= 4 && b[0] == 0xFF && b[1] == 0xD8) {
        size_t i = 2;
        while (i + 4 <= b.size()) {
            if (b[i] != 0xFF) return;
            const uint8_t m = b[i + 1];
            if (m == 0xFF) { ++i; continue; }
            const size_t len = (size_t)b[i + 2] << 8 | b[i + 3];
            const bool sof = m >= 0xC0 && m <= 0xCF && m != 0xC4 && m != 0xC8 && m != 0xCC;
            if (sof && i + 9 <= b.size()) {
                h = (uint32_t)b[i + 5] << 8 | b[i + 6];
                w = (uint32_t)b[i + 7] << 8 | b[i + 8];
                return;
            }
            if (len < 2) return;
            i += 2 + len;
        }
    }
}

// ---- accessors -------------------------------------------------------------------------------
struct View {
    const uint8_t* base = nullptr;  // first element
    size_t count = 0, stride = 0;
    int comp = 0, ncomp = 0;
    bool normalized = false;
};

size_t comp_bytes(int c) {
    switch (c) {
        case 5120: case 5121: return 1;   // BYTE, UNSIGNED_BYTE
        case 5122: case 5123: return 2;   // SHORT, UNSIGNED_SHORT
        case 5125: case 5126: return 4;   // UNSIGNED_INT, FLOAT
        default: return 0;
    }
}
int type_comps(const std::string& t) {
    if (t == "SCALAR") return 1;
    if (t == "VEC2") return 2;
    if (t == "VEC3") return 3;
    if (t == "VEC4") return 4;
    if (t == "MAT2") return 4;
    if (t == "MAT3") return 9;
    if (t == "MAT4") return 16;
    return 0;
}

struct Doc {
    Json root;
    std::vector<std::vector<uint8_t>> buffers;

    const std::vector<Json>& arr(const char* k) const { return list(root, k); }

    View accessor(int64_t ai) const {
        const auto& accs = arr("accessors");
        if (ai < 0 || (size_t)ai >= accs.size()) fail("accessor " + std::to_string(ai) + " out of range");
        const Json& a = accs[(size_t)ai];
        const std::string where = "accessor " + std::to_string(ai);
        if (a.get("sparse")) fail(where + ": sparse accessors are not supported");
        View v;
        v.comp = (int)req(a, "componentType");
        v.ncomp = type_comps(text(a, "type"));
        v.count = (size_t)req(a, "count");
        const Json* nrm = a.get("normalized");
        v.normalized = nrm && nrm->t == Json::T::Bool && nrm->b;
        const size_t cb = comp_bytes(v.comp);
        if (!cb || !v.ncomp) fail(where + ": bad componentType or type");
        const size_t elem = cb * (size_t)v.ncomp;
        const int64_t bvi = idx(a, "bufferView");
        if (bvi < 0) fail(where + ": accessors without a bufferView are not supported");
        const auto& bvs = arr("bufferViews");
        if ((size_t)bvi >= bvs.size()) fail(where + ": bufferView out of range");
        const Json& bv = bvs[(size_t)bvi];
        const int64_t bi = req(bv, "buffer");
        if ((size_t)bi >= buffers.size()) fail(where + ": buffer out of range");
        const size_t bv_off = (size_t)idx(bv, "byteOffset", 0), bv_len = (size_t)req(bv, "byteLength");
        const auto& buf = buffers[(size_t)bi];
        if (bv_off > buf.size() || bv_len > buf.size() - bv_off) fail(where + ": bufferView exceeds its buffer");
        const size_t stride = (size_t)idx(bv, "byteStride", 0);
        v.stride = stride ? stride : elem;
        if (v.stride < elem) fail(where + ": byteStride smaller than the element");
        const size_t off = (size_t)idx(a, "byteOffset", 0);
        if (v.count) {
            const unsigned __int128 end = (unsigned __int128)off + (unsigned __int128)(v.count - 1) * v.stride + elem;
            if (end > bv_len) fail(where + ": data exceeds its bufferView");
        }
        v.base = buf.data() + bv_off + off;
        return v;
    }
};

// component k of element i as float (floats as stored; normalized integers per glTF 2.0 3.11)
float comp_f(const View& v, size_t i, int k) {
    const uint8_t* p = v.base + i * v.stride;
    switch (v.comp) {
        case 5126: { float f; std::memcpy(&f, p + 4 * k, 4); return f; }
        case 5121: return p[k] / 255.0f;
        case 5123: { uint16_t u; std::memcpy(&u, p + 2 * k, 2); return u / 65535.0f; }
        case 5120: { int8_t s; std::memcpy(&s, p + k, 1); return std::max(s / 127.0f, -1.0f); }
        case 5122: { int16_t s; std::memcpy(&s, p + 2 * k, 2); return std::max(s / 32767.0f, -1.0f); }
        default: return 0.0f;
    }
}

View attribute(const Doc& d, const Json& attrs, const char* name, int ncomp, bool allow_normalized, size_t need,
               bool& present, const std::string& where) {
    present = false;
    const int64_t ai = idx(attrs, name);
    if (ai < 0) return {};
    View v = d.accessor(ai);
    if (v.ncomp != ncomp) fail(where + name + ": wrong accessor type");
    if (v.comp != 5126 && !(allow_normalized && v.normalized))
        fail(where + name + ": only FLOAT" + (allow_normalized ? " or normalized integer" : "") +
             " components are supported");
    if (need != (size_t)-1 && v.count < need) fail(where + name + ": fewer elements than POSITION");
    present = true;
    return v;
}

// ---- node transforms (double; row r, column c) ---------------------------------------------
using M4 = std::array<double, 16>;
M4 identity() { M4 m{}; m[0] = m[5] = m[10] = m[15] = 1.0; return m; }
M4 mul(const M4& a, const M4& b) {
    M4 r{};
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) {
            double s = 0.0;
            for (int k = 0; k < 4; ++k) s += a[i * 4 + k] * b[k * 4 + j];
            r[i * 4 + j] = s;
        }
    return r;
}

// node.matrix (column-major, glTF 2.0 5.25) or T * R * S (GltfLoader.cpp:33-61)
M4 local_transform(const Json& node) {
    const auto m = numbers(node, "matrix", 16);
    if (!m.empty()) {
        M4 r;
        for (int row = 0; row < 4; ++row)
            for (int col = 0; col < 4; ++col) r[row * 4 + col] = m[col * 4 + row];
        return r;
    }
    M4 t = identity();
    const auto tr = numbers(node, "translation", 3);
    if (!tr.empty()) { t[3] = tr[0]; t[7] = tr[1]; t[11] = tr[2]; }
    const auto q = numbers(node, "rotation", 4);  // (x, y, z, w)
    if (!q.empty()) {
        const double x = q[0], y = q[1], z = q[2], w = q[3];
        M4 r = identity();
        r[0] = 1 - 2 * (y * y + z * z); r[1] = 2 * (x * y - w * z);     r[2] = 2 * (x * z + w * y);
        r[4] = 2 * (x * y + w * z);     r[5] = 1 - 2 * (x * x + z * z); r[6] = 2 * (y * z - w * x);
        r[8] = 2 * (x * z - w * y);     r[9] = 2 * (y * z + w * x);     r[10] = 1 - 2 * (x * x + y * y);
        t = mul(t, r);
    }
    const auto s = numbers(node, "scale", 3);
    if (!s.empty()) {
        M4 sc = identity();
        sc[0] = s[0]; sc[5] = s[1]; sc[10] = s[2];
        t = mul(t, sc);
    }
    return t;
}

std::string file_name(const std::string& path) {
    const size_t slash = path.find_last_of('/');
    return slash == std::string::npos ? path : path.substr(slash + 1);
}

}  // namespace

std::string gltf_stem(const std::string& path) {
    const std::string f = file_name(path);
    const size_t dot = f.find_last_of('.');
    return (dot == std::string::npos || dot == 0) ? f : f.substr(0, dot);
}

void load_gltf(const std::string& path, GltfScene& out) {
    out = GltfScene{};
    const std::vector<uint8_t> bytes = read_file(path);
    const size_t slash = path.find_last_of('/');
    const std::string dir = slash == std::string::npos ? std::string() : path.substr(0, slash + 1);
    const bool binary = path.size() >= 4 && path.compare(path.size() - 4, 4, ".glb") == 0;

    Doc d;
    std::vector<uint8_t> bin;
    bool have_bin = false;
    if (binary) {
        // 12-byte header (magic "glTF", version 2, total length), then JSON and optional BIN chunks
        if (bytes.size() < 20 || rd32(bytes.data()) != 0x46546C67u) fail("'" + path + "' is not a GLB file");
        if (rd32(bytes.data() + 4) != 2) fail("GLB version " + std::to_string(rd32(bytes.data() + 4)) + " not supported");
        const size_t total = rd32(bytes.data() + 8);
        if (total > bytes.size() || total < 20) fail("GLB length field exceeds the file");
        size_t at = 12;
        bool have_json = false;
        while (at + 8 <= total) {
            const size_t len = rd32(bytes.data() + at), type = rd32(bytes.data() + at + 4);
            at += 8;
            if (len > total - at) fail("GLB chunk exceeds the file");
            if (!have_json) {
                if (type != 0x4E4F534Au) fail("first GLB chunk is not JSON");
                d.root = JsonReader(reinterpret_cast<const char*>(bytes.data() + at), len).parse();
                have_json = true;
            } else if (type == 0x004E4942u && !have_bin) {
                bin.assign(bytes.begin() + (ptrdiff_t)at, bytes.begin() + (ptrdiff_t)(at + len));
                have_bin = true;
            }
            at += (len + 3) & ~size_t(3);
        }
        if (!have_json) fail("GLB has no JSON chunk");
    } else {
        d.root = JsonReader(reinterpret_cast<const char*>(bytes.data()), bytes.size()).parse();
    }
    if (d.root.t != Json::T::Obj) fail("top level is not a JSON object");
    const Json* asset = d.root.get("asset");
    if (!asset || asset->t != Json::T::Obj) fail("missing 'asset'");
    const std::string ver = text(*asset, "version");
    if (ver.empty() || ver[0] != '2') fail("asset.version '" + ver + "' is not glTF 2.x");

    // buffers: the GLB BIN chunk for a buffer without uri, else the uri's bytes
    const auto& bufs = d.arr("buffers");
    for (size_t i = 0; i < bufs.size(); ++i) {
        const std::string uri = text(bufs[i], "uri");
        std::vector<uint8_t> data;
        if (uri.empty()) {
            if (!have_bin || i != 0) fail("buffer " + std::to_string(i) + " has no uri and no GLB BIN chunk");
            data = bin;
        } else {
            data = load_uri(uri, dir);
        }
        const size_t len = (size_t)req(bufs[i], "byteLength");
        if (len > data.size()) fail("buffer " + std::to_string(i) + " shorter than its byteLength");
        data.resize(len);
        d.buffers.push_back(std::move(data));
    }

    out.name = gltf_stem(path);

    // images: described, not decoded (no kernel samples them)
    const auto& imgs = d.arr("images");
    for (size_t i = 0; i < imgs.size(); ++i) {
        GltfImage im;
        im.name = text(imgs[i], "name");
        im.mime_type = text(imgs[i], "mimeType");
        im.uri = text(imgs[i], "uri");
        std::vector<uint8_t> data;
        const int64_t bvi = idx(imgs[i], "bufferView");
        try {
            if (bvi >= 0) {
                const auto& bvs = d.arr("bufferViews");
                if ((size_t)bvi >= bvs.size()) fail("image bufferView out of range");
                const Json& bv = bvs[(size_t)bvi];
                const int64_t bi = req(bv, "buffer");
                if ((size_t)bi >= d.buffers.size()) fail("image buffer out of range");
                const size_t off = (size_t)idx(bv, "byteOffset", 0), len = (size_t)req(bv, "byteLength");
                const auto& b = d.buffers[(size_t)bi];
                if (off > b.size() || len > b.size() - off) fail("image bufferView exceeds its buffer");
                data.assign(b.begin() + (ptrdiff_t)off, b.begin() + (ptrdiff_t)(off + len));
            } else if (!im.uri.empty()) {
                data = load_uri(im.uri, dir);
            }
        } catch (const std::runtime_error& e) {
            out.warnings.push_back(std::string("image ") + std::to_string(i) + ": " + e.what());
        }
        im.bytes = data.size();
        image_size(data, im.width, im.height);
        if (!im.width || !im.height)
            out.warnings.push_back("image " + std::to_string(i) + ": not a PNG / JPEG of known size, skipped");
        out.images.push_back(std::move(im));
    }

    // materials (GltfLoader.cpp:101-155; Scene.cpp:311-330)
    const auto& texs = d.arr("textures");
    auto tex_image = [&](const Json* info) -> int32_t {
        if (!info) return -1;
        const int64_t ti = idx(*info, "index");
        if (ti < 0) return -1;
        if ((size_t)ti >= texs.size()) fail("texture " + std::to_string(ti) + " out of range");
        const int64_t src = idx(texs[(size_t)ti], "source");
        if (src >= 0 && (size_t)src >= out.images.size()) fail("texture source out of range");
        return (int32_t)src;
    };
    const auto& mats = d.arr("materials");
    for (const Json& m : mats) {
        rrh_material r{};
        r.base_color[0] = r.base_color[1] = r.base_color[2] = r.base_color[3] = 1.0f;
        r.metallic = 1.0f;
        r.roughness = 1.0f;
        r.ior = 1.5f;
        GltfMaterialImages mi;
        if (const Json* pbr = m.get("pbrMetallicRoughness")) {
            const auto bc = numbers(*pbr, "baseColorFactor", 4);
            for (size_t k = 0; k < bc.size(); ++k) r.base_color[k] = (float)bc[k];
            r.metallic = (float)num(*pbr, "metallicFactor", 1.0);
            r.roughness = (float)num(*pbr, "roughnessFactor", 1.0);
            mi.base_color = tex_image(pbr->get("baseColorTexture"));
            mi.metallic_roughness = tex_image(pbr->get("metallicRoughnessTexture"));
        }
        const auto em = numbers(m, "emissiveFactor", 3);
        for (size_t k = 0; k < em.size(); ++k) r.emissive[k] = (float)em[k];
        mi.normal = tex_image(m.get("normalTexture"));
        mi.emissive = tex_image(m.get("emissiveTexture"));
        // only the base-colour texture is registered for the device record (Scene.cpp:426-437)
        const bool usable = mi.base_color >= 0 && out.images[(size_t)mi.base_color].width > 0;
        r.albedo_tex_idx = usable ? (uint32_t)mi.base_color : 0xFFFFFFFFu;
        r.normal_tex_idx = r.metallic_roughness_tex_idx = r.emissive_tex_idx = 0xFFFFFFFFu;
        r.sampler_idx = 0;
        out.materials.push_back(r);
        out.material_names.push_back(text(m, "name"));
        out.material_images.push_back(mi);
    }
    if (out.materials.empty()) {  // material index 0 must exist; Material defaults (Scene.h:36-48)
        rrh_material r{};
        r.base_color[0] = r.base_color[1] = r.base_color[2] = r.base_color[3] = 1.0f;
        r.roughness = 1.0f;
        r.ior = 1.5f;
        r.albedo_tex_idx = r.normal_tex_idx = r.metallic_roughness_tex_idx = r.emissive_tex_idx = 0xFFFFFFFFu;
        out.materials.push_back(r);
        out.material_names.push_back("default");
        out.material_images.push_back({});
        out.warnings.push_back("no materials: added a default material");
    }

    // meshes: one per TRIANGLES primitive with POSITION (GltfLoader.cpp:199-323)
    const auto& meshes = d.arr("meshes");
    std::vector<std::vector<uint32_t>> prim_mesh(meshes.size());
    for (size_t gi = 0; gi < meshes.size(); ++gi) {
        const auto& prims = list(meshes[gi], "primitives");
        prim_mesh[gi].assign(prims.size(), 0xFFFFFFFFu);
        for (size_t pi = 0; pi < prims.size(); ++pi) {
            const Json& p = prims[pi];
            const std::string where = "mesh " + std::to_string(gi) + " primitive " + std::to_string(pi) + ": ";
            if (idx(p, "mode", 4) != 4) continue;  // TRIANGLES only
            const Json* attrs = p.get("attributes");
            if (!attrs || attrs->t != Json::T::Obj) continue;
            bool has_pos, has_n, has_uv, has_t;
            View pos, nor, uv, tan;
            pos = attribute(d, *attrs, "POSITION", 3, false, (size_t)-1, has_pos, where);
            if (!has_pos) continue;
            nor = attribute(d, *attrs, "NORMAL", 3, false, pos.count, has_n, where);
            uv = attribute(d, *attrs, "TEXCOORD_0", 2, true, pos.count, has_uv, where);
            tan = attribute(d, *attrs, "TANGENT", 4, false, pos.count, has_t, where);
            const size_t nv = pos.count;
            std::vector<uint32_t> ind;
            const int64_t ia = idx(p, "indices");
            if (ia >= 0) {
                View iv = d.accessor(ia);
                if (iv.ncomp != 1 || (iv.comp != 5121 && iv.comp != 5123 && iv.comp != 5125))
                    fail(where + "indices must be SCALAR u8 / u16 / u32");
                ind.resize(iv.count);
                for (size_t k = 0; k < iv.count; ++k) {
                    const uint8_t* q = iv.base + k * iv.stride;
                    uint32_t x;
                    if (iv.comp == 5121) x = q[0];
                    else if (iv.comp == 5123) { uint16_t u; std::memcpy(&u, q, 2); x = u; }
                    else std::memcpy(&x, q, 4);
                    if (x >= nv) fail(where + "index " + std::to_string(x) + " >= vertex count " + std::to_string(nv));
                    ind[k] = x;
                }
            } else {
                ind.resize(nv);
                for (size_t k = 0; k < nv; ++k) ind[k] = (uint32_t)k;
            }
            if (ind.size() % 3) fail(where + "TRIANGLES index count not a multiple of 3");
            rrh_mesh me{};
            if (out.vertices.size() + nv > 0xFFFFFFFFull || (out.indices.size() + ind.size()) * 4 > 0xFFFFFFFFull)
                fail("scene too large for 32-bit mesh records");
            me.vertex_base = (uint32_t)out.vertices.size();
            me.vertex_count = (uint32_t)nv;
            me.index_byte_offset = (uint32_t)(out.indices.size() * 4);
            me.index_count = (uint32_t)ind.size();
            const int64_t mat = idx(p, "material");
            if (mat >= 0 && (size_t)mat >= mats.size()) fail(where + "material out of range");
            me.material_index = mat >= 0 ? (uint32_t)mat : 0u;
            for (size_t k = 0; k < nv; ++k) {
                rrh_vertex v{};
                for (int c = 0; c < 3; ++c) v.pos[c] = comp_f(pos, k, c);
                if (has_n) for (int c = 0; c < 3; ++c) v.normal[c] = comp_f(nor, k, c);
                else v.normal[1] = 1.0f;
                if (has_uv) for (int c = 0; c < 2; ++c) v.uv[c] = comp_f(uv, k, c);
                if (has_t) for (int c = 0; c < 4; ++c) v.tangent[c] = comp_f(tan, k, c);
                else { v.tangent[0] = 1.0f; v.tangent[3] = 1.0f; }
                out.vertices.push_back(v);
            }
            out.indices.insert(out.indices.end(), ind.begin(), ind.end());
            prim_mesh[gi][pi] = (uint32_t)out.meshes.size();
            out.meshes.push_back(me);
            out.mesh_names.push_back(text(meshes[gi], "name") + (prims.size() > 1 ? "_" + std::to_string(pi) : ""));
        }
    }

    // instances: depth-first over the default scene (GltfLoader.cpp:325-355)
    const auto& scenes = d.arr("scenes");
    const auto& nodes = d.arr("nodes");
    if (scenes.empty()) fail("no scenes");
    const int64_t si = idx(d.root, "scene", 0);
    if ((size_t)si >= scenes.size()) fail("default scene out of range");
    std::vector<char> on_path(nodes.size(), 0);
    std::function<void(int64_t, const M4&, int)> walk = [&](int64_t ni, const M4& parent, int depth) {
        if (ni < 0 || (size_t)ni >= nodes.size()) fail("node " + std::to_string(ni) + " out of range");
        if (on_path[(size_t)ni] || depth > 1024) fail("node hierarchy has a cycle");
        on_path[(size_t)ni] = 1;
        const Json& node = nodes[(size_t)ni];
        const M4 world = mul(parent, local_transform(node));
        const int64_t mi = idx(node, "mesh");
        if (mi >= 0) {
            if ((size_t)mi >= prim_mesh.size()) fail("node mesh out of range");
            for (uint32_t sm : prim_mesh[(size_t)mi]) {
                if (sm == 0xFFFFFFFFu) continue;
                rrh_instance inst{};
                for (int k = 0; k < 16; ++k) inst.transform[k] = (float)world[k];
                inst.mesh_index = sm;
                out.instances.push_back(inst);
            }
        }
        for (const Json& c : list(node, "children")) {
            if (c.t != Json::T::Num || c.n < 0 || c.n != std::floor(c.n)) fail("bad child index");
            walk((int64_t)c.n, world, depth + 1);
        }
        on_path[(size_t)ni] = 0;
    };
    for (const Json& r : list(scenes[(size_t)si], "nodes")) {
        if (r.t != Json::T::Num || r.n < 0 || r.n != std::floor(r.n)) fail("bad scene node index");
        walk((int64_t)r.n, identity(), 0);
    }
}

}  // namespace rrh::scene
