// Mesh loaders (see mesh_io.hpp for the behaviour contract).
#include "mesh_io.hpp"

#include <cctype>
#include <cmath>
#include <cstring>
#include <fstream>
#include <limits>
#include <map>
#include <memory>
#include <sstream>
#include <unordered_map>

namespace sbr {

namespace {

std::string lower(std::string s) {
    for (auto& c : s) c = static_cast<char>(std::tolower(static_cast<unsigned char>(c)));
    return s;
}
bool ends_with(const std::string& s, const std::string& suf) {
    return s.size() >= suf.size() && s.compare(s.size() - suf.size(), suf.size(), suf) == 0;
}
bool read_file(const std::string& path, std::string& out) {
    std::ifstream f(path, std::ios::binary);
    if (!f) return false;
    std::ostringstream ss;
    ss << f.rdbuf();
    out = ss.str();
    return true;
}

}  // namespace

// Number grammar and accumulation of tinyobjloader's tryParseDouble
// (reference include/model/tinyOBJLoader.hpp:897-1028): integer digits are
// accumulated as m = 10 m + d in double, fraction digits add d * 10^-k
// (k < 8 from a table of double constants, else pow(10, -k)), and an
// exponent e is applied as ldexp(m * 5^e, e).
bool obj_parse_number(const char* s, const char* end, double* out) {
    if (s >= end) return false;
    static const double frac[8] = {1.0, 0.1, 0.01, 0.001, 0.0001, 0.00001, 0.000001, 0.0000001};
    const char* c = s;
    bool neg = false;
    bool leadingDot = false;
    if (*c == '+' || *c == '-') {
        neg = *c == '-';
        ++c;
        if (c != end && *c == '.') leadingDot = true;
    } else if (std::isdigit(static_cast<unsigned char>(*c))) {
    } else if (*c == '.') {
        leadingDot = true;
    } else {
        return false;
    }
    double m = 0.0;
    int e = 0;
    if (!leadingDot) {
        int n = 0;
        while (c != end && std::isdigit(static_cast<unsigned char>(*c))) {
            m *= 10;
            m += static_cast<int>(*c - '0');
            ++c;
            ++n;
        }
        if (n == 0) return false;
    }
    if (c == end) goto done;
    if (*c == '.') {
        ++c;
        int k = 1;
        while (c != end && std::isdigit(static_cast<unsigned char>(*c))) {
            m += static_cast<int>(*c - '0') * (k < 8 ? frac[k] : std::pow(10.0, -k));
            ++k;
            ++c;
        }
    } else if (*c != 'e' && *c != 'E') {
        goto done;
    }
    if (c == end) goto done;
    if (*c == 'e' || *c == 'E') {
        ++c;
        bool eneg = false;
        if (c != end && (*c == '+' || *c == '-')) {
            eneg = *c == '-';
            ++c;
        } else if (c == end || !std::isdigit(static_cast<unsigned char>(*c))) {
            return false;
        }
        int n = 0;
        while (c != end && std::isdigit(static_cast<unsigned char>(*c))) {
            if (e > 2147483647 / 10) return false;
            e = e * 10 + static_cast<int>(*c - '0');
            ++c;
            ++n;
        }
        if (n == 0) return false;
        if (eneg) e = -e;
    }
done:
    *out = (neg ? -1 : 1) * (e ? std::ldexp(m * std::pow(5.0, e), e) : m);
    return true;
}

namespace {

// one whitespace-delimited number; default 0 on a parse failure
float obj_real(const char*& p) {
    p += std::strspn(p, " \t");
    const char* end = p + std::strcspn(p, " \t\r");
    double v = 0.0;
    obj_parse_number(p, end, &v);
    p = end;
    return static_cast<float>(v);
}

}  // namespace

namespace {

// One face corner: position, texcoord and normal index (0-based, -1 when
// absent).
struct Corner {
    int v = -1, vt = -1, vn = -1;
};

// OBJ index -> 0-based (1-based, negative = relative to the count so far;
// 0 is invalid), as tinyobj's fixIndex does.
bool fix_index(int idx, int n, int& out) {
    if (idx > 0) {
        out = idx - 1;
        return true;
    }
    if (idx == 0) return false;
    out = n + idx;
    return true;
}

// "v", "v/vt", "v//vn" or "v/vt/vn" (tinyobj's parseTriple).
bool parse_corner(const char* p, const char* end, int nv, int nvt, int nvn, Corner& c) {
    auto num = [&](const char*& q, int& val) -> bool {
        const char* s0 = q;
        val = std::atoi(q);
        if (*q == '+' || *q == '-') ++q;
        while (q < end && std::isdigit(static_cast<unsigned char>(*q))) ++q;
        return q > s0;
    };
    const char* q = p;
    int a = 0;
    if (!num(q, a) || !fix_index(a, nv, c.v)) return false;
    if (q < end && *q == '/') {
        ++q;
        if (q < end && *q != '/') {
            int b = 0;
            if (num(q, b) && !fix_index(b, nvt, c.vt)) return false;
        }
        if (q < end && *q == '/') {
            ++q;
            int d = 0;
            if (num(q, d) && !fix_index(d, nvn, c.vn)) return false;
        }
    }
    return true;
}

// W. R. Franklin's crossing-number point-in-polygon test, in float, as
// tinyobj applies it to a candidate ear (tinyOBJLoader.hpp:1440-1450).
bool inside_polygon(int n, const float* x, const float* y, float tx, float ty) {
    bool c = false;
    for (int a = 0, b = n - 1; a < n; b = a++)
        if (((y[a] > ty) != (y[b] > ty)) && (tx < (x[b] - x[a]) * (ty - y[a]) / (y[b] - y[a]) + x[a])) c = !c;
    return c;
}

// Triangulation of one face with 5+ corners by tinyobj's built-in ear
// clipping (tinyOBJLoader.hpp:1735-1948, its default without earcut):
// project onto the two axes chosen from the first corner with a
// non-degenerate turn; walk a guess index around the remaining polygon,
// skipping corners whose turn has the opposite sign to the fan term
// (x0 y1 - y0 x1) / 2 and corners whose triangle contains another remaining
// corner; emit each accepted ear and drop its middle corner; stop when the
// polygon is a triangle (emitted) or no corner was accepted in a full turn.
void clip_ears(const std::vector<Corner>& face, const std::vector<float>& v, std::vector<Corner>& tris) {
    const size_t nf = face.size();
    auto coord = [&](int vi, int axis) -> float {
        const size_t k = size_t(vi) * 3 + size_t(axis);
        return vi >= 0 && k < v.size() ? v[k] : 0.0f;
    };
    int ax0 = 1, ax1 = 2;
    for (size_t k = 0; k < nf; ++k) {
        const int i0 = face[k].v, i1 = face[(k + 1) % nf].v, i2 = face[(k + 2) % nf].v;
        if (i0 < 0 || i1 < 0 || i2 < 0 || size_t(i0) * 3 + 2 >= v.size() || size_t(i1) * 3 + 2 >= v.size() ||
            size_t(i2) * 3 + 2 >= v.size())
            continue;
        const float* p0 = &v[size_t(i0) * 3];
        const float* p1 = &v[size_t(i1) * 3];
        const float* p2 = &v[size_t(i2) * 3];
        const float e0x = p1[0] - p0[0], e0y = p1[1] - p0[1], e0z = p1[2] - p0[2];
        const float e1x = p2[0] - p1[0], e1y = p2[1] - p1[1], e1z = p2[2] - p1[2];
        const float cx = std::fabs(e0y * e1z - e0z * e1y);
        const float cy = std::fabs(e0z * e1x - e0x * e1z);
        const float cz = std::fabs(e0x * e1y - e0y * e1x);
        const float eps = std::numeric_limits<float>::epsilon();
        if (cx > eps || cy > eps || cz > eps) {
            if (!(cx > cy && cx > cz)) {
                ax0 = 0;
                if (cz > cx && cz > cy) ax1 = 1;
            }
            break;
        }
    }
    std::vector<Corner> rem = face;
    size_t guess = 0, budget = nf, lastSize = nf;
    while (rem.size() > 3 && budget > 0) {
        const size_t n = rem.size();
        if (guess >= n) guess -= n;
        if (lastSize != n) {
            lastSize = n;
            budget = n;
        } else {
            --budget;
        }
        Corner ear[3];
        float x[3], y[3];
        for (int k = 0; k < 3; ++k) {
            ear[k] = rem[(guess + k) % n];
            x[k] = coord(ear[k].v, ax0);
            y[k] = coord(ear[k].v, ax1);
        }
        const float turn = (x[1] - x[0]) * (y[2] - y[1]) - (y[1] - y[0]) * (x[2] - x[1]);
        const float fan = (x[0] * y[1] - y[0] * x[1]) * 0.5f;
        if (turn * fan < 0.0f) {
            ++guess;
            continue;
        }
        bool blocked = false;
        for (size_t o = 3; o < n && !blocked; ++o) {
            const int vi = rem[(guess + o) % n].v;
            if (vi < 0 || size_t(vi) * 3 + size_t(ax0) >= v.size() || size_t(vi) * 3 + size_t(ax1) >= v.size())
                continue;
            blocked = inside_polygon(3, x, y, v[size_t(vi) * 3 + size_t(ax0)], v[size_t(vi) * 3 + size_t(ax1)]);
        }
        if (blocked) {
            ++guess;
            continue;
        }
        tris.insert(tris.end(), ear, ear + 3);
        rem.erase(rem.begin() + static_cast<std::ptrdiff_t>((guess + 1) % n));
    }
    if (rem.size() == 3) tris.insert(tris.end(), rem.begin(), rem.end());
}

}  // namespace

bool load_obj(const std::string& path, MeshData& out, std::string& err) {
    std::ifstream f(path);
    if (!f) {
        err = "Cannot open file [" + path + "]";
        return false;
    }
    std::vector<float> v;  // all 'v' records
    int nvt = 0, nvn = 0;  // 'vt' and 'vn' records (for relative indices)
    std::vector<std::vector<Corner>> faces;
    std::string line, acc;
    while (std::getline(f, line)) {
        if (!line.empty() && line.back() == '\r') line.pop_back();
        // backslash line continuation
        if (!line.empty() && line.back() == '\\') {
            acc += line.substr(0, line.size() - 1);
            continue;
        }
        acc += line;
        std::string l;
        l.swap(acc);
        const char* p = l.c_str();
        p += std::strspn(p, " \t");
        if (p[0] == 'v' && (p[1] == ' ' || p[1] == '\t')) {
            p += 2;
            float x = obj_real(p), y = obj_real(p), z = obj_real(p);
            v.push_back(x);
            v.push_back(y);
            v.push_back(z);
        } else if (p[0] == 'v' && p[1] == 't' && (p[2] == ' ' || p[2] == '\t')) {
            ++nvt;
        } else if (p[0] == 'v' && p[1] == 'n' && (p[2] == ' ' || p[2] == '\t')) {
            ++nvn;
        } else if (p[0] == 'f' && (p[1] == ' ' || p[1] == '\t')) {
            p += 2;
            std::vector<Corner> fv;
            const int nv = static_cast<int>(v.size() / 3);
            while (true) {
                p += std::strspn(p, " \t");
                if (*p == '\0') break;
                const char* end = p + std::strcspn(p, " \t\r");
                Corner c;
                if (!parse_corner(p, end, nv, nvt, nvn, c)) {
                    err = "Failed to parse `f' line (invalid vertex index) in " + path;
                    return false;
                }
                fv.push_back(c);
                p = end;
            }
            if (fv.size() >= 3) faces.push_back(std::move(fv));
        }
    }
    out.positions.clear();
    out.indices.clear();
    const int nv = static_cast<int>(v.size() / 3);
    // vertices deduplicated by (position, normal, texcoord) in first-use order
    struct Key {
        int v, vn, vt;
        bool operator==(const Key& o) const { return v == o.v && vn == o.vn && vt == o.vt; }
    };
    struct KeyHash {
        size_t operator()(const Key& k) const {
            return std::hash<int>{}(k.v) ^ (std::hash<int>{}(k.vn) << 1) ^ (std::hash<int>{}(k.vt) << 2);
        }
    };
    std::unordered_map<Key, uint32_t, KeyHash> remap;
    auto add = [&](const Corner& c) -> bool {
        if (c.v < 0 || c.v >= nv) return false;
        const Key key{c.v, c.vn, c.vt};
        auto it = remap.find(key);
        uint32_t id;
        if (it == remap.end()) {
            id = static_cast<uint32_t>(out.positions.size() / 3);
            out.positions.push_back(v[3 * size_t(c.v)]);
            out.positions.push_back(v[3 * size_t(c.v) + 1]);
            out.positions.push_back(v[3 * size_t(c.v) + 2]);
            remap.emplace(key, id);
        } else {
            id = it->second;
        }
        out.indices.push_back(id);
        return true;
    };
    for (const auto& fv : faces) {
        std::vector<Corner> tri;
        if (fv.size() == 3) {
            tri = fv;
        } else if (fv.size() == 4) {
            // split on the shorter diagonal (tinyOBJLoader.hpp:1509-1600);
            // a quad with an index out of range is skipped
            bool ok = true;
            for (int k = 0; k < 4; ++k) ok = ok && fv[k].v >= 0 && fv[k].v < nv;
            if (!ok) continue;
            const float* a = &v[3 * size_t(fv[0].v)];
            const float* b = &v[3 * size_t(fv[1].v)];
            const float* c = &v[3 * size_t(fv[2].v)];
            const float* d = &v[3 * size_t(fv[3].v)];
            float e02x = c[0] - a[0], e02y = c[1] - a[1], e02z = c[2] - a[2];
            float e13x = d[0] - b[0], e13y = d[1] - b[1], e13z = d[2] - b[2];
            float s02 = e02x * e02x + e02y * e02y + e02z * e02z;
            float s13 = e13x * e13x + e13y * e13y + e13z * e13z;
            if (s02 < s13) tri = {fv[0], fv[1], fv[2], fv[0], fv[2], fv[3]};
            else tri = {fv[0], fv[1], fv[3], fv[1], fv[2], fv[3]};
        } else {
            clip_ears(fv, v, tri);
        }
        for (const Corner& c : tri)
            if (!add(c)) {
                err = "Face with invalid vertex index found in " + path;
                return false;
            }
    }
    return true;
}

// ------------------------------------------------------------------ glTF
namespace {

struct Json {
    enum Kind { Null, Bool, Num, Str, Arr, Obj } kind = Null;
    double num = 0;
    bool b = false;
    std::string str;
    std::vector<Json> arr;
    std::map<std::string, Json> obj;
    const Json* get(const std::string& k) const {
        if (kind != Obj) return nullptr;
        auto it = obj.find(k);
        return it == obj.end() ? nullptr : &it->second;
    }
    double num_or(const std::string& k, double d) const {
        const Json* j = get(k);
        return j && j->kind == Num ? j->num : d;
    }
};

struct JsonParser {
    const char* p;
    const char* e;
    std::string err;
    void ws() {
        while (p < e && (*p == ' ' || *p == '\t' || *p == '\n' || *p == '\r')) ++p;
    }
    bool value(Json& out, int depth = 0) {
        if (depth > 256) return fail("nesting too deep");
        ws();
        if (p >= e) return fail("unexpected end");
        char c = *p;
        if (c == '{') {
            out.kind = Json::Obj;
            ++p;
            ws();
            if (p < e && *p == '}') {
                ++p;
                return true;
            }
            while (true) {
                ws();
                Json key;
                if (p >= e || *p != '"' || !string(key.str)) return fail("bad object key");
                ws();
                if (p >= e || *p != ':') return fail("expected ':'");
                ++p;
                Json val;
                if (!value(val, depth + 1)) return false;
                out.obj[key.str] = std::move(val);
                ws();
                if (p < e && *p == ',') {
                    ++p;
                    continue;
                }
                if (p < e && *p == '}') {
                    ++p;
                    return true;
                }
                return fail("expected ',' or '}'");
            }
        }
        if (c == '[') {
            out.kind = Json::Arr;
            ++p;
            ws();
            if (p < e && *p == ']') {
                ++p;
                return true;
            }
            while (true) {
                Json val;
                if (!value(val, depth + 1)) return false;
                out.arr.push_back(std::move(val));
                ws();
                if (p < e && *p == ',') {
                    ++p;
                    continue;
                }
                if (p < e && *p == ']') {
                    ++p;
                    return true;
                }
                return fail("expected ',' or ']'");
            }
        }
        if (c == '"') {
            out.kind = Json::Str;
            return string(out.str);
        }
        if (c == 't' && e - p >= 4 && std::strncmp(p, "true", 4) == 0) {
            out.kind = Json::Bool;
            out.b = true;
            p += 4;
            return true;
        }
        if (c == 'f' && e - p >= 5 && std::strncmp(p, "false", 5) == 0) {
            out.kind = Json::Bool;
            p += 5;
            return true;
        }
        if (c == 'n' && e - p >= 4 && std::strncmp(p, "null", 4) == 0) {
            p += 4;
            return true;
        }
        char* q = nullptr;
        std::string tmp(p, std::min<size_t>(64, size_t(e - p)));
        double d = std::strtod(tmp.c_str(), &q);
        if (q == tmp.c_str()) return fail("bad value");
        out.kind = Json::Num;
        out.num = d;
        p += (q - tmp.c_str());
        return true;
    }
    bool string(std::string& s) {
        ++p;  // opening quote
        while (p < e && *p != '"') {
            if (*p == '\\') {
                ++p;
                if (p >= e) return fail("bad escape");
                char c = *p;
                if (c == 'u') {
                    if (e - p < 5) return fail("bad \\u escape");
                    unsigned cp = std::strtoul(std::string(p + 1, p + 5).c_str(), nullptr, 16);
                    if (cp < 0x80) s += char(cp);
                    else if (cp < 0x800) {
                        s += char(0xC0 | (cp >> 6));
                        s += char(0x80 | (cp & 0x3F));
                    } else {
                        s += char(0xE0 | (cp >> 12));
                        s += char(0x80 | ((cp >> 6) & 0x3F));
                        s += char(0x80 | (cp & 0x3F));
                    }
                    p += 5;
                    continue;
                }
                s += c == 'n' ? '\n' : c == 't' ? '\t' : c == 'r' ? '\r' : c == 'b' ? '\b' : c == 'f' ? '\f' : c;
                ++p;
            } else {
                s += *p++;
            }
        }
        if (p >= e) return fail("unterminated string");
        ++p;
        return true;
    }
    bool fail(const std::string& m) {
        if (err.empty()) err = m;
        return false;
    }
};

bool base64_decode(const std::string& in, std::string& out) {
    auto val = [](char c) -> int {
        if (c >= 'A' && c <= 'Z') return c - 'A';
        if (c >= 'a' && c <= 'z') return c - 'a' + 26;
        if (c >= '0' && c <= '9') return c - '0' + 52;
        if (c == '+' || c == '-') return 62;
        if (c == '/' || c == '_') return 63;
        return -1;
    };
    unsigned buf = 0;
    int bits = 0;
    for (char c : in) {
        if (c == '=' || c == '\n' || c == '\r') continue;
        int v = val(c);
        if (v < 0) return false;
        buf = (buf << 6) | unsigned(v);
        bits += 6;
        if (bits >= 8) {
            bits -= 8;
            out += char((buf >> bits) & 0xFF);
        }
    }
    return true;
}

int comp_count(const std::string& type) {
    if (type == "SCALAR") return 1;
    if (type == "VEC2") return 2;
    if (type == "VEC3") return 3;
    if (type == "VEC4") return 4;
    if (type == "MAT2") return 4;
    if (type == "MAT3") return 9;
    if (type == "MAT4") return 16;
    return 0;
}
int comp_size(int ct) {
    switch (ct) {
        case 5120: case 5121: return 1;
        case 5122: case 5123: return 2;
        case 5125: case 5126: return 4;
        default: return 0;
    }
}

struct View {
    const unsigned char* base = nullptr;
    size_t avail = 0;
    size_t stride = 0;
};

}  // namespace

bool load_gltf(const std::string& path, MeshData& out, std::string& err) {
    std::string raw;
    if (!read_file(path, raw)) {
        err = "File not found : " + path;
        return false;
    }
    std::string json, bin;
    bool glb = ends_with(lower(path), ".glb");
    if (glb) {
        if (raw.size() < 20 || raw.compare(0, 4, "glTF") != 0) {
            err = "Invalid GLB header";
            return false;
        }
        size_t off = 12;
        while (off + 8 <= raw.size()) {
            uint32_t len, type;
            std::memcpy(&len, raw.data() + off, 4);
            std::memcpy(&type, raw.data() + off + 4, 4);
            if (off + 8 + len > raw.size()) {
                err = "Truncated GLB chunk";
                return false;
            }
            if (type == 0x4E4F534Au) json.assign(raw.data() + off + 8, len);
            else if (type == 0x004E4942u) bin.assign(raw.data() + off + 8, len);
            off += 8 + ((len + 3) & ~3u);
        }
    } else {
        json = raw;
    }
    JsonParser jp{json.data(), json.data() + json.size(), ""};
    Json root;
    if (!jp.value(root) || root.kind != Json::Obj) {
        err = "JSON parse error: " + jp.err;
        return false;
    }
    std::string dir;
    size_t slash = path.find_last_of("/\\");
    if (slash != std::string::npos) dir = path.substr(0, slash + 1);
    // buffers
    std::vector<std::string> buffers;
    if (const Json* bs = root.get("buffers")) {
        for (size_t i = 0; i < bs->arr.size(); ++i) {
            const Json& b = bs->arr[i];
            const Json* uri = b.get("uri");
            std::string data;
            if (!uri) {
                if (!glb || i != 0) {
                    err = "buffer without uri";
                    return false;
                }
                data = bin;
            } else if (uri->str.compare(0, 5, "data:") == 0) {
                size_t comma = uri->str.find(',');
                if (comma == std::string::npos || !base64_decode(uri->str.substr(comma + 1), data)) {
                    err = "bad data URI";
                    return false;
                }
            } else if (!read_file(dir + uri->str, data)) {
                err = "cannot read buffer " + uri->str;
                return false;
            }
            buffers.push_back(std::move(data));
        }
    }
    const Json* views = root.get("bufferViews");
    const Json* accs = root.get("accessors");
    auto accessor_view = [&](int ai, View& v, int& ct, int& nc, size_t& count) -> bool {
        if (!accs || ai < 0 || size_t(ai) >= accs->arr.size()) return false;
        const Json& a = accs->arr[ai];
        int bvi = int(a.num_or("bufferView", -1));
        if (!views || bvi < 0 || size_t(bvi) >= views->arr.size()) return false;
        const Json& bv = views->arr[bvi];
        int bi = int(bv.num_or("buffer", -1));
        if (bi < 0 || size_t(bi) >= buffers.size()) return false;
        size_t off = size_t(bv.num_or("byteOffset", 0)) + size_t(a.num_or("byteOffset", 0));
        const std::string& buf = buffers[bi];
        if (off > buf.size()) return false;
        v.base = reinterpret_cast<const unsigned char*>(buf.data()) + off;
        v.avail = buf.size() - off;
        v.stride = size_t(bv.num_or("byteStride", 0));
        ct = int(a.num_or("componentType", 0));
        const Json* t = a.get("type");
        nc = t ? comp_count(t->str) : 0;
        count = size_t(a.num_or("count", 0));
        return true;
    };
    out.positions.clear();
    out.indices.clear();
    const Json* meshes = root.get("meshes");
    if (!meshes) return true;
    for (const Json& m : meshes->arr) {
        const Json* prims = m.get("primitives");
        if (!prims) continue;
        for (const Json& pr : prims->arr) {
            if (int(pr.num_or("mode", 4)) != 4) continue;
            const Json* attrs = pr.get("attributes");
            const Json* pos = attrs ? attrs->get("POSITION") : nullptr;
            if (!pos) continue;
            View pv;
            int ct = 0, nc = 0;
            size_t cnt = 0;
            if (!accessor_view(int(pos->num), pv, ct, nc, cnt)) {
                err = "bad POSITION accessor";
                return false;
            }
            if (ct != 5126 || cnt == 0) continue;  // FLOAT only
            size_t stride = pv.stride ? pv.stride : size_t(nc) * 4;
            if (cnt > 0 && (cnt - 1) * stride + size_t(std::min(nc, 3)) * 4 > pv.avail) {
                err = "POSITION accessor exceeds its buffer";
                return false;
            }
            const uint32_t base = static_cast<uint32_t>(out.positions.size() / 3);
            for (size_t i = 0; i < cnt; ++i) {
                const unsigned char* s = pv.base + i * stride;
                for (int c = 0; c < 3; ++c) {
                    float x = 0.0f;
                    if (c < nc) std::memcpy(&x, s + 4 * c, 4);
                    out.positions.push_back(x);
                }
            }
            const Json* ind = pr.get("indices");
            if (ind) {
                View iv;
                int ict = 0, inc = 0;
                size_t icnt = 0;
                if (!accessor_view(int(ind->num), iv, ict, inc, icnt)) {
                    err = "bad indices accessor";
                    return false;
                }
                size_t cs = size_t(comp_size(ict));
                size_t is = iv.stride ? iv.stride : cs;
                if (icnt > 0 && (cs == 0 || (icnt - 1) * is + cs > iv.avail)) {
                    err = "indices accessor exceeds its buffer";
                    return false;
                }
                for (size_t i = 0; i < icnt; ++i) {
                    const unsigned char* s = iv.base + i * is;
                    uint32_t x = 0;
                    if (ict == 5121) x = *s;
                    else if (ict == 5123) {
                        uint16_t h;
                        std::memcpy(&h, s, 2);
                        x = h;
                    } else if (ict == 5125) std::memcpy(&x, s, 4);
                    out.indices.push_back(base + x);
                }
            } else {
                for (size_t i = 0; i < cnt; ++i) out.indices.push_back(base + uint32_t(i));
            }
        }
    }
    return true;
}

bool load_model(const std::string& path, MeshData& out, std::string& err) {
    std::string l = lower(path);
    if (ends_with(l, ".obj")) return load_obj(path, out, err);
    if (ends_with(l, ".gltf") || ends_with(l, ".glb")) return load_gltf(path, out, err);
    err = "Unsupported file extension for \"" + path + "\"";
    return false;
}

}  // namespace sbr
