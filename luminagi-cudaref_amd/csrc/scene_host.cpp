// scene_host.cpp — scene.json reader/writer, procedural scene generator, PNG writer and
// batch-file parser (host side of liblgi.so).
#include "scene_host.h"

#include <zlib.h>

#include <algorithm>
#include <charconv>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iterator>
#include <sstream>

HostScene::HostScene() {
    // CameraDesc / DirectionalLight defaults (include/scene.h:6-23)
    camera.fov_y_deg = 60.0f;
    camera.aspect = 1.0f;
    camera.near_z = 0.1f;
    camera.far_z = 1000.0f;
    camera.width = 1920;
    camera.height = 1080;
    sun.dir[0] = 0; sun.dir[1] = 0; sun.dir[2] = -1;
    sun.color[0] = sun.color[1] = sun.color[2] = 1;
    sun.intensity = 1.0f;
}

// =============================================================================== JSON
// A single-pass reader for the scene.json v1 schema (src/scene_loader.cpp:46-128). Generic
// values are parsed into a small tree, except "triangles", whose elements are decoded
// straight into lgi_triangle records (no per-number allocation). Numbers are parsed as
// double (correctly rounded) then narrowed, matching json::get<float>() on a double.
namespace {

struct JVal {
    enum Kind { NUL, BOOL, NUM, STR, ARR, OBJ } kind = NUL;
    double num = 0;
    bool b = false;
    std::string str;
    std::vector<JVal> arr;
    std::vector<std::pair<std::string, JVal>> obj;
    const JVal* get(const char* key) const {
        for (auto& kv : obj) if (kv.first == key) return &kv.second;
        return nullptr;
    }
};

struct Parser {
    const char* p;
    const char* end;
    std::string err;
    bool fail(const char* what) {
        if (err.empty()) err = std::string(what) + " at byte " + std::to_string((long long)(end - p));
        return false;
    }
    void ws() {
        while (p < end && (*p == ' ' || *p == '\n' || *p == '\r' || *p == '\t')) ++p;
    }
    bool lit(const char* s) {
        size_t n = std::strlen(s);
        if ((size_t)(end - p) < n || std::memcmp(p, s, n) != 0) return fail("bad literal");
        p += n;
        return true;
    }
    bool number(double& out) {
        auto r = std::from_chars(p, end, out, std::chars_format::general);
        if (r.ec != std::errc()) return fail("bad number");
        p = r.ptr;
        return true;
    }
    bool string(std::string& out) {
        if (p >= end || *p != '"') return fail("expected string");
        ++p;
        out.clear();
        while (p < end && *p != '"') {
            char c = *p++;
            if (c == '\\') {
                if (p >= end) return fail("bad escape");
                char e = *p++;
                switch (e) {
                    case 'n': out += '\n'; break;
                    case 't': out += '\t'; break;
                    case 'r': out += '\r'; break;
                    case 'b': out += '\b'; break;
                    case 'f': out += '\f'; break;
                    case 'u': {
                        if (end - p < 4) return fail("bad \\u escape");
                        unsigned cp = 0;
                        for (int i = 0; i < 4; ++i) {
                            char h = *p++;
                            cp <<= 4;
                            if (h >= '0' && h <= '9') cp |= h - '0';
                            else if (h >= 'a' && h <= 'f') cp |= h - 'a' + 10;
                            else if (h >= 'A' && h <= 'F') cp |= h - 'A' + 10;
                            else return fail("bad \\u escape");
                        }
                        if (cp < 0x80) out += (char)cp;
                        else if (cp < 0x800) { out += (char)(0xC0 | (cp >> 6)); out += (char)(0x80 | (cp & 63)); }
                        else { out += (char)(0xE0 | (cp >> 12)); out += (char)(0x80 | ((cp >> 6) & 63)); out += (char)(0x80 | (cp & 63)); }
                        break;
                    }
                    default: out += e; break;
                }
            } else {
                out += c;
            }
        }
        if (p >= end) return fail("unterminated string");
        ++p;
        return true;
    }
    bool value(JVal& v, int depth = 0) {
        if (depth > 64) return fail("nesting too deep");
        ws();
        if (p >= end) return fail("unexpected end");
        char c = *p;
        if (c == '{') {
            v.kind = JVal::OBJ;
            ++p; ws();
            if (p < end && *p == '}') { ++p; return true; }
            for (;;) {
                ws();
                std::string k;
                if (!string(k)) return false;
                ws();
                if (p >= end || *p != ':') return fail("expected ':'");
                ++p;
                v.obj.emplace_back(std::move(k), JVal());
                if (!value(v.obj.back().second, depth + 1)) return false;
                ws();
                if (p < end && *p == ',') { ++p; continue; }
                if (p < end && *p == '}') { ++p; return true; }
                return fail("expected ',' or '}'");
            }
        }
        if (c == '[') {
            v.kind = JVal::ARR;
            ++p; ws();
            if (p < end && *p == ']') { ++p; return true; }
            for (;;) {
                v.arr.emplace_back();
                if (!value(v.arr.back(), depth + 1)) return false;
                ws();
                if (p < end && *p == ',') { ++p; continue; }
                if (p < end && *p == ']') { ++p; return true; }
                return fail("expected ',' or ']'");
            }
        }
        if (c == '"') { v.kind = JVal::STR; return string(v.str); }
        if (c == 't') { v.kind = JVal::BOOL; v.b = true; return lit("true"); }
        if (c == 'f') { v.kind = JVal::BOOL; v.b = false; return lit("false"); }
        if (c == 'n') { v.kind = JVal::NUL; return lit("null"); }
        v.kind = JVal::NUM;
        return number(v.num);
    }
    // Fast path for one triangles[] element: {"v":[[..],[..],[..]], "n":..., "uv":..., "mat":k}
    bool float_array(float* out, int n) {
        ws();
        if (p >= end || *p != '[') return fail("expected '['");
        ++p;
        for (int i = 0; i < n; ++i) {
            ws();
            double d;
            if (!number(d)) return false;
            out[i] = (float)d;
            ws();
            if (i + 1 < n) {
                if (p >= end || *p != ',') return fail("expected ','");
                ++p;
            }
        }
        ws();
        if (p >= end || *p != ']') return fail("expected ']'");
        ++p;
        return true;
    }
    bool vec_array(float* out, int rows, int cols) {
        ws();
        if (p >= end || *p != '[') return fail("expected '['");
        ++p;
        for (int r = 0; r < rows; ++r) {
            if (!float_array(out + r * cols, cols)) return false;
            ws();
            if (r + 1 < rows) {
                if (p >= end || *p != ',') return fail("expected ','");
                ++p;
            }
        }
        ws();
        // the reference reads v[0..2] and ignores further entries; accept and skip extras
        while (p < end && *p == ',') {
            ++p;
            JVal skip;
            if (!value(skip)) return false;
            ws();
        }
        if (p >= end || *p != ']') return fail("expected ']'");
        ++p;
        return true;
    }
    bool triangle(lgi_triangle& t) {
        ws();
        if (p >= end || *p != '{') return fail("expected triangle object");
        ++p;
        bool hv = false, hn = false, huv = false, hm = false;
        ws();
        if (p < end && *p == '}') return fail("triangle missing keys");
        for (;;) {
            ws();
            std::string k;
            if (!string(k)) return false;
            ws();
            if (p >= end || *p != ':') return fail("expected ':'");
            ++p;
            if (k == "v") { if (!vec_array(&t.v[0][0], 3, 3)) return false; hv = true; }
            else if (k == "n") { if (!vec_array(&t.n[0][0], 3, 3)) return false; hn = true; }
            else if (k == "uv") { if (!vec_array(&t.uv[0][0], 3, 2)) return false; huv = true; }
            else if (k == "mat") {
                ws();
                double d;
                if (!number(d)) return false;
                t.material = (int32_t)d;
                hm = true;
            } else {
                JVal skip;
                if (!value(skip)) return false;
            }
            ws();
            if (p < end && *p == ',') { ++p; continue; }
            if (p < end && *p == '}') { ++p; break; }
            return fail("expected ',' or '}'");
        }
        if (!(hv && hn && huv && hm)) return fail("triangle missing key v/n/uv/mat");
        return true;
    }
    bool triangles(std::vector<lgi_triangle>& out) {
        ws();
        if (p >= end || *p != '[') return fail("triangles must be an array");
        ++p;
        ws();
        if (p < end && *p == ']') { ++p; return true; }
        for (;;) {
            lgi_triangle t;
            std::memset(&t, 0, sizeof(t));
            if (!triangle(t)) return false;
            out.push_back(t);
            ws();
            if (p < end && *p == ',') { ++p; continue; }
            if (p < end && *p == ']') { ++p; return true; }
            return fail("expected ',' or ']'");
        }
    }
};

bool num_of(const JVal* v, float& out) {
    if (!v || v->kind != JVal::NUM) return false;
    out = (float)v->num;
    return true;
}
bool vec_of(const JVal* v, float* out, int n) {
    if (!v || v->kind != JVal::ARR || (int)v->arr.size() < n) return false;
    for (int i = 0; i < n; ++i)
        if (!num_of(&v->arr[i], out[i])) return false;
    return true;
}

}  // namespace

lgi_status scene_load_json(const char* path, HostScene& out, std::string& err, bool verbose) {
    std::ifstream f(path, std::ios::binary);
    if (!f.is_open()) {
        err = std::string("[scene] failed to open ") + path;
        if (verbose) std::fprintf(stderr, "%s\n", err.c_str());
        return LGI_ERR_IO;
    }
    std::string text((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
    Parser ps{text.data(), text.data() + text.size(), {}};
    HostScene s;
    // top level object: parse members; "triangles" takes the streaming path
    JVal top;
    top.kind = JVal::OBJ;
    ps.ws();
    bool ok = ps.p < ps.end && *ps.p == '{';
    if (ok) {
        ++ps.p;
        ps.ws();
        if (ps.p < ps.end && *ps.p == '}') ++ps.p;
        else
            for (;;) {
                ps.ws();
                std::string k;
                if (!ps.string(k)) { ok = false; break; }
                ps.ws();
                if (ps.p >= ps.end || *ps.p != ':') { ok = ps.fail("expected ':'"); break; }
                ++ps.p;
                if (k == "triangles") {
                    s.tris.clear();
                    if (!ps.triangles(s.tris)) { ok = false; break; }
                    top.obj.emplace_back(k, JVal());
                    top.obj.back().second.kind = JVal::ARR;
                } else {
                    top.obj.emplace_back(std::move(k), JVal());
                    if (!ps.value(top.obj.back().second)) { ok = false; break; }
                }
                ps.ws();
                if (ps.p < ps.end && *ps.p == ',') { ++ps.p; continue; }
                if (ps.p < ps.end && *ps.p == '}') { ++ps.p; break; }
                ok = ps.fail("expected ',' or '}'");
                break;
            }
    } else {
        ps.fail("expected top-level object");
    }
    if (ok) {
        ps.ws();
        if (ps.p != ps.end) ok = ps.fail("trailing characters");
    }
    if (!ok) {
        err = "[scene] JSON parse error: " + ps.err;
        if (verbose) std::fprintf(stderr, "%s\n", err.c_str());
        return LGI_ERR_PARSE;
    }
    // version (json::value("version", 0), src/scene_loader.cpp:46-51)
    int version = 0;
    if (const JVal* v = top.get("version"); v && v->kind == JVal::NUM) version = (int)v->num;
    if (version != 1) {
        err = "[scene] unsupported version " + std::to_string(version) + " (expected 1)";
        if (verbose) std::fprintf(stderr, "%s\n", err.c_str());
        return LGI_ERR_VERSION;
    }
    auto missing = [&](const char* what) {
        err = std::string("[scene] missing or malformed key: ") + what;
        if (verbose) std::fprintf(stderr, "%s\n", err.c_str());
        return LGI_ERR_PARSE;
    };
    const JVal* cam = top.get("camera");
    if (!cam || cam->kind != JVal::OBJ) return missing("camera");
    if (!vec_of(cam->get("pos"), s.camera.pos, 3)) return missing("camera.pos");
    if (!vec_of(cam->get("camera_to_world"), s.camera.camera_to_world, 16)) return missing("camera.camera_to_world");
    if (!num_of(cam->get("fov_y_deg"), s.camera.fov_y_deg)) return missing("camera.fov_y_deg");
    if (!num_of(cam->get("aspect"), s.camera.aspect)) return missing("camera.aspect");
    if (!num_of(cam->get("near"), s.camera.near_z)) return missing("camera.near");
    if (!num_of(cam->get("far"), s.camera.far_z)) return missing("camera.far");
    float wf, hf;
    if (!num_of(cam->get("image_width"), wf)) return missing("camera.image_width");
    if (!num_of(cam->get("image_height"), hf)) return missing("camera.image_height");
    s.camera.width = (int32_t)cam->get("image_width")->num;
    s.camera.height = (int32_t)cam->get("image_height")->num;
    const JVal* sun = top.get("sun");
    if (!sun || sun->kind != JVal::OBJ) return missing("sun");
    if (!vec_of(sun->get("direction"), s.sun.dir, 3)) return missing("sun.direction");
    if (!vec_of(sun->get("color"), s.sun.color, 3)) return missing("sun.color");
    if (!num_of(sun->get("intensity"), s.sun.intensity)) return missing("sun.intensity");
    const JVal* lights = top.get("lights");
    if (!lights || lights->kind != JVal::ARR) return missing("lights");
    for (const JVal& L : lights->arr) {
        const JVal* ty = L.get("type");
        if (!ty || ty->kind != JVal::STR) return missing("lights[].type");
        if (ty->str == "directional") {
            lgi_dir_light d{};
            if (!vec_of(L.get("direction"), d.dir, 3) || !vec_of(L.get("color"), d.color, 3) ||
                !num_of(L.get("intensity"), d.intensity))
                return missing("directional light");
            s.dirs.push_back(d);
        } else if (ty->str == "point") {
            lgi_point_light pl{};
            if (!vec_of(L.get("pos"), pl.pos, 3) || !vec_of(L.get("color"), pl.color, 3) ||
                !num_of(L.get("intensity"), pl.intensity) || !num_of(L.get("radius"), pl.radius))
                return missing("point light");
            s.points.push_back(pl);
        }
    }
    const JVal* mats = top.get("materials");
    if (!mats || mats->kind != JVal::ARR) return missing("materials");
    for (const JVal& M : mats->arr) {
        HostMaterial m;
        if (!vec_of(M.get("albedo"), m.albedo, 3) || !vec_of(M.get("emissive"), m.emissive, 3))
            return missing("materials[].albedo/emissive");
        if (const JVal* t = M.get("diffuse_tex")) {
            if (t->kind != JVal::STR) return missing("materials[].diffuse_tex");
            m.diffuse_tex = t->str;
        }
        s.materials.push_back(m);
    }
    if (!top.get("triangles")) return missing("triangles");
    out = std::move(s);
    if (verbose) {  // summary lines, src/scene_loader.cpp:130-136
        std::printf("[scene] loaded %s\n", path);
        std::printf("  camera: %dx%d, fov %.1f deg, pos (%.2f,%.2f,%.2f)\n", out.camera.width,
                    out.camera.height, out.camera.fov_y_deg, out.camera.pos[0], out.camera.pos[1],
                    out.camera.pos[2]);
        std::printf("  %zu directional + %zu point lights, %zu materials, %zu triangles\n",
                    out.dirs.size(), out.points.size(), out.materials.size(), out.tris.size());
    }
    return LGI_OK;
}

namespace {
void put_f(std::string& o, float v) {
    char b[32];
    int n = std::snprintf(b, sizeof b, "%.9g", (double)v);
    o.append(b, n);
}
void put_v(std::string& o, const float* v, int n) {
    o += '[';
    for (int i = 0; i < n; ++i) {
        if (i) o += ',';
        put_f(o, v[i]);
    }
    o += ']';
}
}  // namespace

lgi_status scene_save_json(const char* path, const HostScene& s, std::string& err) {
    std::string o;
    o.reserve(256 + s.tris.size() * 260);
    o += "{\n\"version\": 1,\n\"camera\": {\"pos\": ";
    put_v(o, s.camera.pos, 3);
    o += ", \"camera_to_world\": ";
    put_v(o, s.camera.camera_to_world, 16);
    o += ", \"fov_y_deg\": "; put_f(o, s.camera.fov_y_deg);
    o += ", \"aspect\": "; put_f(o, s.camera.aspect);
    o += ", \"near\": "; put_f(o, s.camera.near_z);
    o += ", \"far\": "; put_f(o, s.camera.far_z);
    o += ", \"image_width\": " + std::to_string(s.camera.width);
    o += ", \"image_height\": " + std::to_string(s.camera.height) + "},\n";
    o += "\"sun\": {\"direction\": "; put_v(o, s.sun.dir, 3);
    o += ", \"color\": "; put_v(o, s.sun.color, 3);
    o += ", \"intensity\": "; put_f(o, s.sun.intensity);
    o += "},\n\"lights\": [";
    bool first = true;
    for (auto& d : s.dirs) {
        if (!first) o += ",";
        first = false;
        o += "\n {\"type\": \"directional\", \"direction\": "; put_v(o, d.dir, 3);
        o += ", \"color\": "; put_v(o, d.color, 3);
        o += ", \"intensity\": "; put_f(o, d.intensity); o += "}";
    }
    for (auto& p : s.points) {
        if (!first) o += ",";
        first = false;
        o += "\n {\"type\": \"point\", \"pos\": "; put_v(o, p.pos, 3);
        o += ", \"color\": "; put_v(o, p.color, 3);
        o += ", \"intensity\": "; put_f(o, p.intensity);
        o += ", \"radius\": "; put_f(o, p.radius); o += "}";
    }
    o += "],\n\"materials\": [";
    for (size_t i = 0; i < s.materials.size(); ++i) {
        if (i) o += ",";
        o += "\n {\"albedo\": "; put_v(o, s.materials[i].albedo, 3);
        o += ", \"emissive\": "; put_v(o, s.materials[i].emissive, 3);
        if (!s.materials[i].diffuse_tex.empty()) o += ", \"diffuse_tex\": \"" + s.materials[i].diffuse_tex + "\"";
        o += "}";
    }
    o += "],\n\"triangles\": [";
    for (size_t i = 0; i < s.tris.size(); ++i) {
        const lgi_triangle& t = s.tris[i];
        if (i) o += ",";
        o += "\n{\"v\":[";
        for (int k = 0; k < 3; ++k) { if (k) o += ','; put_v(o, t.v[k], 3); }
        o += "],\"n\":[";
        for (int k = 0; k < 3; ++k) { if (k) o += ','; put_v(o, t.n[k], 3); }
        o += "],\"uv\":[";
        for (int k = 0; k < 3; ++k) { if (k) o += ','; put_v(o, t.uv[k], 2); }
        o += "],\"mat\":" + std::to_string(t.material) + "}";
    }
    o += "]\n}\n";
    FILE* f = std::fopen(path, "wb");
    if (!f) { err = std::string("cannot write ") + path; return LGI_ERR_IO; }
    bool ok = std::fwrite(o.data(), 1, o.size(), f) == o.size();
    ok = (std::fclose(f) == 0) && ok;
    if (!ok) { err = std::string("short write ") + path; return LGI_ERR_IO; }
    return LGI_OK;
}

// ====================================================================== camera
void camera_look_at(lgi_camera& cam, const float pos[3], const float tgt[3]) {
    // src/main.cpp:29-59: I = forward, J = left = cross(worldUp, I), K = up = cross(I, J),
    // stored as rows 0/1/2 of camera_to_world, translation in row 3.
    float fx = tgt[0] - pos[0], fy = tgt[1] - pos[1], fz = tgt[2] - pos[2];
    float fl = std::sqrt(fx * fx + fy * fy + fz * fz);
    if (fl < 1e-8f) return;
    fx /= fl; fy /= fl; fz /= fl;
    float lx = -fy, ly = fx, lz = 0.0f;
    float ll = std::sqrt(lx * lx + ly * ly + lz * lz);
    if (ll < 1e-6f) { lx = 0; ly = 1; lz = 0; ll = 1; }
    lx /= ll; ly /= ll; lz /= ll;
    float ux = fy * lz - fz * ly, uy = fz * lx - fx * lz, uz = fx * ly - fy * lx;
    float* m = cam.camera_to_world;
    m[0] = fx; m[1] = fy; m[2] = fz; m[3] = 0;
    m[4] = lx; m[5] = ly; m[6] = lz; m[7] = 0;
    m[8] = ux; m[9] = uy; m[10] = uz; m[11] = 0;
    m[12] = pos[0]; m[13] = pos[1]; m[14] = pos[2]; m[15] = 1;
    cam.pos[0] = pos[0]; cam.pos[1] = pos[1]; cam.pos[2] = pos[2];
}

// ================================================================= generator
namespace {

struct Prng {  // SplitMix64 stream (Steele, Lea, Flood 2014)
    uint64_t s;
    uint64_t next() {
        uint64_t z = (s += 0x9E3779B97F4A7C15ULL);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
        return z ^ (z >> 31);
    }
    double uni() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
    double uni(double a, double b) { return a + (b - a) * uni(); }
    double normal() {  // Box-Muller
        double u1 = uni(), u2 = uni();
        if (u1 < 1e-300) u1 = 1e-300;
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
    }
};

struct D3 { double x, y, z; };
D3 d3(double x, double y, double z) { return {x, y, z}; }
D3 operator+(D3 a, D3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
D3 operator*(D3 a, double s) { return {a.x * s, a.y * s, a.z * s}; }
D3 dcross(D3 a, D3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
double ddot(D3 a, D3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
D3 dnorm(D3 a) {
    double l = std::sqrt(ddot(a, a));
    return l > 0 ? a * (1.0 / l) : d3(0, 0, 1);
}

struct Mesh {
    std::vector<D3> p, n;
    std::vector<double> u, v;
    std::vector<int> idx;  // triangles
};

// Geodesic sphere: icosahedron faces subdivided at frequency f (20 f^2 triangles),
// vertices projected to the unit sphere. Frequency 2^L reproduces a level-L icosphere.
Mesh geosphere(int f) {
    const double t = (1.0 + std::sqrt(5.0)) / 2.0;
    D3 V[12] = {{-1, t, 0}, {1, t, 0}, {-1, -t, 0}, {1, -t, 0}, {0, -1, t}, {0, 1, t},
                {0, -1, -t}, {0, 1, -t}, {t, 0, -1}, {t, 0, 1}, {-t, 0, -1}, {-t, 0, 1}};
    int F[20][3] = {{0, 11, 5}, {0, 5, 1}, {0, 1, 7}, {0, 7, 10}, {0, 10, 11}, {1, 5, 9}, {5, 11, 4},
                    {11, 10, 2}, {10, 7, 6}, {7, 1, 8}, {3, 9, 4}, {3, 4, 2}, {3, 2, 6}, {3, 6, 8},
                    {3, 8, 9}, {4, 9, 5}, {2, 4, 11}, {6, 2, 10}, {8, 6, 7}, {9, 8, 1}};
    Mesh m;
    for (auto& face : F) {
        D3 a = V[face[0]], b = V[face[1]], c = V[face[2]];
        int base = (int)m.p.size();
        // barycentric lattice rows i = 0..f, j = 0..f-i
        for (int i = 0; i <= f; ++i)
            for (int j = 0; j <= f - i; ++j) {
                double wb = (double)i / f, wc = (double)j / f, wa = 1.0 - wb - wc;
                D3 q = dnorm(a * wa + b * wb + c * wc);
                m.p.push_back(q);
                m.n.push_back(q);
                m.u.push_back(0.5 + std::atan2(q.y, q.x) / 6.283185307179586);
                m.v.push_back(0.5 + std::asin(std::max(-1.0, std::min(1.0, q.z))) / 3.141592653589793);
            }
        auto id = [&](int i, int j) { return base + i * (f + 1) - (i * (i - 1)) / 2 + j; };
        for (int i = 0; i < f; ++i)
            for (int j = 0; j < f - i; ++j) {
                m.idx.insert(m.idx.end(), {id(i, j), id(i + 1, j), id(i, j + 1)});
                if (j + 1 < f - i) m.idx.insert(m.idx.end(), {id(i + 1, j), id(i + 1, j + 1), id(i, j + 1)});
            }
    }
    return m;
}

// Torus in the xy plane: a segments around the ring, b around the tube (2ab triangles).
Mesh torus(int a, int b, double R, double r) {
    Mesh m;
    for (int i = 0; i <= a; ++i)
        for (int j = 0; j <= b; ++j) {
            double th = 6.283185307179586 * i / a, ph = 6.283185307179586 * j / b;
            D3 c = d3(std::cos(th), std::sin(th), 0);
            D3 nn = c * std::cos(ph) + d3(0, 0, std::sin(ph));
            m.p.push_back(c * R + nn * r);
            m.n.push_back(nn);
            m.u.push_back((double)i / a);
            m.v.push_back((double)j / b);
        }
    for (int i = 0; i < a; ++i)
        for (int j = 0; j < b; ++j) {
            int p00 = i * (b + 1) + j, p10 = (i + 1) * (b + 1) + j, p01 = p00 + 1, p11 = p10 + 1;
            m.idx.insert(m.idx.end(), {p00, p10, p11, p00, p11, p01});
        }
    return m;
}

// Box [-1,1]^3 with g x g quads per face (12 g^2 triangles), flat face normals.
Mesh box(int g) {
    Mesh m;
    for (int f = 0; f < 6; ++f) {
        int ax = f / 2;
        double s = (f & 1) ? 1.0 : -1.0;
        D3 nrm = d3(ax == 0 ? s : 0, ax == 1 ? s : 0, ax == 2 ? s : 0);
        D3 t1 = d3(ax == 1 ? 1 : (ax == 2 ? 1 : 0), ax == 0 ? 1 : 0, 0);
        if (ax == 1) t1 = d3(0, 0, 1);
        D3 t2 = dcross(nrm, t1);
        int base = (int)m.p.size();
        for (int i = 0; i <= g; ++i)
            for (int j = 0; j <= g; ++j) {
                double a = -1.0 + 2.0 * i / g, b = -1.0 + 2.0 * j / g;
                m.p.push_back(nrm + t1 * a + t2 * b);
                m.n.push_back(nrm);
                m.u.push_back((double)i / g);
                m.v.push_back((double)j / g);
            }
        for (int i = 0; i < g; ++i)
            for (int j = 0; j < g; ++j) {
                int p00 = base + i * (g + 1) + j, p10 = base + (i + 1) * (g + 1) + j;
                int p01 = p00 + 1, p11 = p10 + 1;
                m.idx.insert(m.idx.end(), {p00, p10, p11, p00, p11, p01});
            }
    }
    return m;
}

// Value noise on an integer lattice with smoothstep interpolation, sum of 3 octaves.
double value_noise(uint64_t seed, double x, double y) {
    auto lat = [seed](int64_t i, int64_t j) {
        uint64_t h = seed ^ ((uint64_t)i * 0x9E3779B97F4A7C15ULL) ^ ((uint64_t)j * 0xC2B2AE3D27D4EB4FULL);
        h = (h ^ (h >> 30)) * 0xBF58476D1CE4E5B9ULL;
        h = (h ^ (h >> 27)) * 0x94D049BB133111EBULL;
        h ^= h >> 31;
        return (double)(h >> 11) * (1.0 / 9007199254740992.0) * 2.0 - 1.0;
    };
    double sum = 0, amp = 1, norm = 0, fx = x, fy = y;
    for (int o = 0; o < 3; ++o) {
        int64_t ix = (int64_t)std::floor(fx), iy = (int64_t)std::floor(fy);
        double tx = fx - ix, ty = fy - iy;
        tx = tx * tx * (3 - 2 * tx);
        ty = ty * ty * (3 - 2 * ty);
        double a = lat(ix, iy), b = lat(ix + 1, iy), c = lat(ix, iy + 1), d = lat(ix + 1, iy + 1);
        sum += amp * ((a * (1 - tx) + b * tx) * (1 - ty) + (c * (1 - tx) + d * tx) * ty);
        norm += amp;
        amp *= 0.5;
        fx *= 2;
        fy *= 2;
    }
    return sum / norm;
}

// Height-field patch over [-1,1]^2 with g x g quads, z = amplitude * noise.
Mesh heightfield(int g, double amplitude_local, uint64_t seed) {
    Mesh m;
    auto h = [&](double x, double y) { return amplitude_local * value_noise(seed, x * 3.0 + 17.0, y * 3.0 + 5.0); };
    for (int i = 0; i <= g; ++i)
        for (int j = 0; j <= g; ++j) {
            double x = -1.0 + 2.0 * i / g, y = -1.0 + 2.0 * j / g;
            double e = 1e-3;
            D3 nn = dnorm(d3(-(h(x + e, y) - h(x - e, y)) / (2 * e), -(h(x, y + e) - h(x, y - e)) / (2 * e), 1.0));
            m.p.push_back(d3(x, y, h(x, y)));
            m.n.push_back(nn);
            m.u.push_back((double)i / g);
            m.v.push_back((double)j / g);
        }
    for (int i = 0; i < g; ++i)
        for (int j = 0; j < g; ++j) {
            int p00 = i * (g + 1) + j, p10 = (i + 1) * (g + 1) + j, p01 = p00 + 1, p11 = p10 + 1;
            m.idx.insert(m.idx.end(), {p00, p10, p11, p00, p11, p01});
        }
    return m;
}

// Uniform random rotation from a unit quaternion (Shoemake 1992).
void random_rotation(Prng& r, double M[3][3]) {
    double u1 = r.uni(), u2 = r.uni(), u3 = r.uni();
    double a = std::sqrt(1 - u1), b = std::sqrt(u1);
    double qx = a * std::sin(6.283185307179586 * u2), qy = a * std::cos(6.283185307179586 * u2);
    double qz = b * std::sin(6.283185307179586 * u3), qw = b * std::cos(6.283185307179586 * u3);
    M[0][0] = 1 - 2 * (qy * qy + qz * qz); M[0][1] = 2 * (qx * qy - qz * qw); M[0][2] = 2 * (qx * qz + qy * qw);
    M[1][0] = 2 * (qx * qy + qz * qw); M[1][1] = 1 - 2 * (qx * qx + qz * qz); M[1][2] = 2 * (qy * qz - qx * qw);
    M[2][0] = 2 * (qx * qz - qy * qw); M[2][1] = 2 * (qy * qz + qx * qw); M[2][2] = 1 - 2 * (qx * qx + qy * qy);
}

void emit(HostScene& s, const Mesh& m, const double R[3][3], double scale, D3 pos, int mat) {
    auto xf = [&](D3 p) {
        return d3(R[0][0] * p.x + R[0][1] * p.y + R[0][2] * p.z, R[1][0] * p.x + R[1][1] * p.y + R[1][2] * p.z,
                  R[2][0] * p.x + R[2][1] * p.y + R[2][2] * p.z);
    };
    for (size_t t = 0; t + 2 < m.idx.size(); t += 3) {
        lgi_triangle tri{};
        for (int k = 0; k < 3; ++k) {
            int i = m.idx[t + k];
            D3 p = xf(m.p[i]) * scale + pos;
            D3 nn = dnorm(xf(m.n[i]));
            tri.v[k][0] = (float)p.x; tri.v[k][1] = (float)p.y; tri.v[k][2] = (float)p.z;
            tri.n[k][0] = (float)nn.x; tri.n[k][1] = (float)nn.y; tri.n[k][2] = (float)nn.z;
            tri.uv[k][0] = (float)m.u[i]; tri.uv[k][1] = (float)m.v[i];
        }
        tri.material = mat;
        s.tris.push_back(tri);
    }
}

void add_material(HostScene& s, Prng& r) {
    HostMaterial m;
    for (int c = 0; c < 3; ++c) m.albedo[c] = (float)r.uni(0.1, 0.9);
    s.materials.push_back(m);
}

}  // namespace

lgi_status scene_generate(int kind, uint64_t seed, int width, int height, HostScene& out, std::string& err) {
    if (kind != 1 && kind != 2) { err = "scene kind must be 1 or 2"; return LGI_ERR_INVALID_ARG; }
    if (width <= 0 || height <= 0) { err = "bad image size"; return LGI_ERR_INVALID_ARG; }
    HostScene s;
    Prng r{seed * 0x2545F4914F6CDD1DULL + (uint64_t)kind};
    // camera: pinhole, 60 deg vertical FOV, looking at the 20 m object cube
    s.camera.fov_y_deg = 60.0f;
    s.camera.width = width;
    s.camera.height = height;
    s.camera.aspect = (float)width / (float)height;
    s.camera.near_z = 0.1f;
    s.camera.far_z = 1000.0f;
    const float cpos[3] = {0.0f, -30.0f, 14.0f}, ctgt[3] = {0.0f, 0.0f, 6.0f};
    camera_look_at(s.camera, cpos, ctgt);
    // sun: direction to the light uniform on the upper hemisphere (elevation >= 10 deg)
    {
        double z = r.uni(std::sin(10.0 * 3.141592653589793 / 180.0), 1.0), ph = r.uni(0, 6.283185307179586);
        double rr = std::sqrt(1 - z * z);
        s.sun.dir[0] = (float)(-rr * std::cos(ph));
        s.sun.dir[1] = (float)(-rr * std::sin(ph));
        s.sun.dir[2] = (float)(-z);
        s.sun.color[0] = s.sun.color[1] = s.sun.color[2] = 1.0f;
        s.sun.intensity = 3.0f;
    }
    for (int i = 0; i < 2; ++i) {
        lgi_point_light p{};
        p.pos[0] = (float)r.uni(-10, 10); p.pos[1] = (float)r.uni(-10, 10); p.pos[2] = (float)r.uni(0, 20);
        p.color[0] = p.color[1] = p.color[2] = 1.0f;
        p.intensity = 50.0f;
        p.radius = 10.0f;
        s.points.push_back(p);
    }
    // ground quad 40 x 40 m at z = 0 (material 0)
    add_material(s, r);
    {
        lgi_triangle a{}, b{};
        const float P[4][3] = {{-20, -20, 0}, {20, -20, 0}, {20, 20, 0}, {-20, 20, 0}};
        const float UV[4][2] = {{0, 0}, {1, 0}, {1, 1}, {0, 1}};
        const int T[2][3] = {{0, 1, 2}, {0, 2, 3}};
        lgi_triangle* tt[2] = {&a, &b};
        for (int t = 0; t < 2; ++t) {
            for (int k = 0; k < 3; ++k) {
                std::memcpy(tt[t]->v[k], P[T[t][k]], 12);
                tt[t]->n[k][0] = 0; tt[t]->n[k][1] = 0; tt[t]->n[k][2] = 1;
                std::memcpy(tt[t]->uv[k], UV[T[t][k]], 8);
            }
            tt[t]->material = 0;
            s.tris.push_back(*tt[t]);
        }
    }
    const int n_meshes = kind == 1 ? 16 : 62;
    std::vector<int> types(n_meshes), counts(n_meshes);
    for (int i = 0; i < n_meshes; ++i) types[i] = (int)(r.next() % 3);
    if (kind == 2) {
        // log-uniform target counts in [2k, 80k], rescaled to 1.05M in total
        double tot = 0;
        std::vector<double> c(n_meshes);
        for (int i = 0; i < n_meshes; ++i) { c[i] = std::exp(r.uni(std::log(2000.0), std::log(80000.0))); tot += c[i]; }
        for (int i = 0; i < n_meshes; ++i) counts[i] = (int)std::lround(c[i] * (1050000.0 - 2.0) / tot);
    }
    for (int i = 0; i < n_meshes; ++i) {
        const double scale = std::exp(0.5 * r.normal());  // log-normal(0, 0.5)
        double R[3][3];
        random_rotation(r, R);
        const D3 pos = d3(r.uni(-10, 10), r.uni(-10, 10), r.uni(0, 20));
        const uint64_t noise_seed = r.next();
        Mesh m;
        double world_scale = scale;
        if (kind == 1) {
            if (types[i] == 0) m = geosphere(16);          // icosphere level 4: 5120 tris
            else if (types[i] == 1) m = box(8);            // 8x8 quads per face: 768 tris
            else m = torus(64, 32, 1.0, 0.35);             // 64x32: 4096 tris
        } else {
            const int n = counts[i];
            if (types[i] == 0) {
                m = geosphere(std::max(1, (int)std::lround(std::sqrt(n / 20.0))));
            } else if (types[i] == 1) {
                const int b = std::max(3, (int)std::lround(std::sqrt(n / 4.0)));
                m = torus(2 * b, b, 1.0, 0.35);
            } else {
                // 4 m x 4 m patch (times the mesh scale), value-noise relief of 0.2 m
                const int g = std::max(2, (int)std::lround(std::sqrt(n / 2.0)));
                world_scale = 2.0 * scale;
                m = heightfield(g, 0.2 / world_scale, noise_seed);
            }
        }
        add_material(s, r);
        emit(s, m, R, world_scale, pos, (int)s.materials.size() - 1);
    }
    out = std::move(s);
    return LGI_OK;
}

// ======================================================================= PNG
namespace {
void be32(std::vector<uint8_t>& o, uint32_t v) {
    o.push_back(v >> 24); o.push_back(v >> 16); o.push_back(v >> 8); o.push_back(v);
}
void chunk(std::vector<uint8_t>& o, const char* type, const uint8_t* data, size_t n) {
    be32(o, (uint32_t)n);
    size_t start = o.size();
    o.insert(o.end(), type, type + 4);
    o.insert(o.end(), data, data + n);
    uLong crc = crc32(0L, o.data() + start, (uInt)(n + 4));
    be32(o, (uint32_t)crc);
}
}  // namespace

lgi_status write_png_rgb(const char* path, const uint8_t* rgb, int w, int h) {
    if (!path || !rgb || w <= 0 || h <= 0) return LGI_ERR_INVALID_ARG;
    std::vector<uint8_t> raw((size_t)h * (1 + 3 * (size_t)w));
    for (int y = 0; y < h; ++y) {
        uint8_t* row = &raw[(size_t)y * (1 + 3 * (size_t)w)];
        row[0] = 0;  // filter: none
        std::memcpy(row + 1, rgb + (size_t)y * 3 * w, 3 * (size_t)w);
    }
    uLongf zlen = compressBound(raw.size());
    std::vector<uint8_t> z(zlen);
    if (compress2(z.data(), &zlen, raw.data(), raw.size(), 6) != Z_OK) return LGI_ERR_IO;
    std::vector<uint8_t> o = {0x89, 'P', 'N', 'G', '\r', '\n', 0x1A, '\n'};
    uint8_t ihdr[13];
    ihdr[0] = w >> 24; ihdr[1] = w >> 16; ihdr[2] = w >> 8; ihdr[3] = w;
    ihdr[4] = h >> 24; ihdr[5] = h >> 16; ihdr[6] = h >> 8; ihdr[7] = h;
    ihdr[8] = 8; ihdr[9] = 2; ihdr[10] = 0; ihdr[11] = 0; ihdr[12] = 0;
    chunk(o, "IHDR", ihdr, 13);
    chunk(o, "IDAT", z.data(), zlen);
    chunk(o, "IEND", nullptr, 0);
    FILE* f = std::fopen(path, "wb");
    if (!f) return LGI_ERR_IO;
    bool ok = std::fwrite(o.data(), 1, o.size(), f) == o.size();
    ok = (std::fclose(f) == 0) && ok;
    return ok ? LGI_OK : LGI_ERR_IO;
}

// ===================================================================== PNG reader
// Texture images (the reference loads them with stbi_load(..., 4), src/pathtracer.cu:553-561):
// non-interlaced PNG, bit depth 8 or 16 (16-bit samples keep their high byte), grey, grey +
// alpha, RGB, RGBA or palette, decoded to RGBA8 rows top to bottom.
namespace {

uint32_t be32(const uint8_t* p) { return (uint32_t)p[0] << 24 | (uint32_t)p[1] << 16 | (uint32_t)p[2] << 8 | p[3]; }

int paeth(int a, int b, int c) {
    const int p = a + b - c, pa = std::abs(p - a), pb = std::abs(p - b), pc = std::abs(p - c);
    if (pa <= pb && pa <= pc) return a;
    return pb <= pc ? b : c;
}

}  // namespace

lgi_status read_png_rgba(const char* path, int& w, int& h, std::vector<uint8_t>& rgba, std::string& err) {
    std::ifstream in(path, std::ios::binary);
    if (!in) { err = std::string("cannot open ") + path; return LGI_ERR_IO; }
    std::vector<uint8_t> f((std::istreambuf_iterator<char>(in)), std::istreambuf_iterator<char>());
    auto bad = [&](const char* what) { err = std::string(path) + ": " + what; return LGI_ERR_PARSE; };
    static const uint8_t sig[8] = {0x89, 'P', 'N', 'G', '\r', '\n', 0x1A, '\n'};
    if (f.size() < 8 || std::memcmp(f.data(), sig, 8) != 0) return bad("not a PNG file");
    int depth = 0, ctype = -1;
    bool have_hdr = false;
    std::vector<uint8_t> idat, pal;
    size_t pos = 8;
    while (pos + 12 <= f.size()) {
        const uint32_t len = be32(&f[pos]);
        if (len > f.size() - pos - 12) return bad("truncated chunk");
        const char* type = (const char*)&f[pos + 4];
        const uint8_t* d = &f[pos + 8];
        if (!std::memcmp(type, "IHDR", 4)) {
            if (len < 13) return bad("short IHDR");
            w = (int)be32(d);
            h = (int)be32(d + 4);
            depth = d[8];
            ctype = d[9];
            if (d[12] != 0) return bad("interlaced PNG not supported");
            have_hdr = true;
        } else if (!std::memcmp(type, "PLTE", 4)) {
            pal.assign(d, d + len);
        } else if (!std::memcmp(type, "IDAT", 4)) {
            idat.insert(idat.end(), d, d + len);
        } else if (!std::memcmp(type, "IEND", 4)) {
            break;
        }
        pos += 12 + len;
    }
    if (!have_hdr || w <= 0 || h <= 0 || w > (1 << 15) || h > (1 << 15)) return bad("bad or missing IHDR");
    int ch = 0;
    switch (ctype) {
        case 0: ch = 1; break;
        case 2: ch = 3; break;
        case 3: ch = 1; break;
        case 4: ch = 2; break;
        case 6: ch = 4; break;
        default: return bad("unsupported colour type");
    }
    if (!(depth == 8 || (depth == 16 && ctype != 3))) return bad("unsupported bit depth");
    if (ctype == 3 && pal.size() < 3) return bad("palette image without PLTE");
    const size_t bpp = (size_t)ch * depth / 8, stride = bpp * (size_t)w;
    // deflate expands at most ~1032:1, so a header promising far more pixels than the compressed
    // data can hold is corrupt; refusing it here avoids allocating for it
    if (((size_t)stride + 1) * h > idat.size() * 1032 + 1024) return bad("image data too short for its size");
    std::vector<uint8_t> raw(((size_t)stride + 1) * h);
    uLongf rlen = raw.size();
    if (uncompress(raw.data(), &rlen, idat.data(), idat.size()) != Z_OK || rlen != raw.size())
        return bad("corrupt image data");
    std::vector<uint8_t> img(stride * h), zero(stride, 0);
    for (int y = 0; y < h; ++y) {
        const uint8_t ft = raw[(size_t)y * (stride + 1)];
        const uint8_t* src = &raw[(size_t)y * (stride + 1) + 1];
        uint8_t* dst = &img[(size_t)y * stride];
        const uint8_t* up = y ? &img[(size_t)(y - 1) * stride] : zero.data();
        for (size_t x = 0; x < stride; ++x) {
            const int a = x >= bpp ? dst[x - bpp] : 0, b = up[x], c = x >= bpp ? up[x - bpp] : 0;
            int v = src[x];
            switch (ft) {
                case 0: break;
                case 1: v += a; break;
                case 2: v += b; break;
                case 3: v += (a + b) >> 1; break;
                case 4: v += paeth(a, b, c); break;
                default: return bad("bad filter type");
            }
            dst[x] = (uint8_t)v;
        }
    }
    rgba.assign((size_t)w * h * 4, 255);
    for (size_t i = 0; i < (size_t)w * h; ++i) {
        uint8_t s8[4];
        for (int k = 0; k < ch; ++k) s8[k] = img[i * bpp + (size_t)k * depth / 8];  // high byte of 16-bit
        uint8_t* o = &rgba[i * 4];
        switch (ctype) {
            case 0: o[0] = o[1] = o[2] = s8[0]; break;
            case 4: o[0] = o[1] = o[2] = s8[0]; o[3] = s8[1]; break;
            case 2: o[0] = s8[0]; o[1] = s8[1]; o[2] = s8[2]; break;
            case 6: o[0] = s8[0]; o[1] = s8[1]; o[2] = s8[2]; o[3] = s8[3]; break;
            case 3: {
                const size_t e = (size_t)s8[0] * 3;
                if (e + 2 >= pal.size()) return bad("palette index out of range");
                o[0] = pal[e]; o[1] = pal[e + 1]; o[2] = pal[e + 2];
                break;
            }
        }
    }
    return LGI_OK;
}

// ===================================================================== batch
lgi_status parse_batch_file(const char* path, std::vector<lgi_batch_job>& jobs, std::string& err) {
    // src/main.cpp:71-91: one job per line "out spp bounces cx cy cz tx ty tz", '#' starts a
    // comment, lines that do not parse are skipped.
    std::ifstream in(path);
    if (!in) { err = std::string("can't open batch file: ") + path; return LGI_ERR_IO; }
    std::string line;
    while (std::getline(in, line)) {
        auto hash = line.find('#');
        if (hash != std::string::npos) line.erase(hash);
        std::istringstream ss(line);
        std::string out;
        lgi_batch_job j{};
        if (!(ss >> out >> j.spp >> j.bounces >> j.cam_pos[0] >> j.cam_pos[1] >> j.cam_pos[2] >> j.cam_tgt[0] >>
              j.cam_tgt[1] >> j.cam_tgt[2]))
            continue;
        if (out.size() >= sizeof(j.out_png)) { err = "batch output path too long"; return LGI_ERR_INVALID_ARG; }
        std::memcpy(j.out_png, out.c_str(), out.size() + 1);
        jobs.push_back(j);
    }
    return LGI_OK;
}

// ------------------------------------------------------------------ binary scene file
// Little-endian container of the same scene model, written and read with bulk copies: a 64-byte
// header, the camera and sun records, the directional and point lights, the materials (albedo,
// emissive, texture path) and the 100-byte triangles, all as laid out in include/lgi.h.
namespace {

constexpr char kBinMagic[8] = {'L', 'G', 'I', 'S', 'C', 'N', '1', '\0'};
constexpr uint32_t kBinVersion = 1;

struct BinHeader {
    char magic[8];
    uint32_t version;
    uint32_t header_bytes;
    uint64_t n_tris;
    uint32_t n_mats, n_points, n_dirs, reserved;
    uint64_t file_bytes;  // total size, checked on load
    uint8_t pad[16];
};
static_assert(sizeof(BinHeader) == 64, "binary scene header");

struct FileCloser {
    FILE* f;
    ~FileCloser() {
        if (f) std::fclose(f);
    }
};

// fwrite/fread of n records; n = 0 passes (an empty vector's data() may be null, which the C
// library does not accept even for zero records).
bool put_records(FILE* f, const void* p, size_t size, size_t n) { return n == 0 || std::fwrite(p, size, n, f) == n; }
bool get_records(FILE* f, void* p, size_t size, size_t n) { return n == 0 || std::fread(p, size, n, f) == n; }

}  // namespace

lgi_status scene_save_bin(const char* path, const HostScene& s, std::string& err) {
    BinHeader h{};
    std::memcpy(h.magic, kBinMagic, 8);
    h.version = kBinVersion;
    h.header_bytes = sizeof(BinHeader);
    h.n_tris = s.tris.size();
    h.n_mats = (uint32_t)s.materials.size();
    h.n_points = (uint32_t)s.points.size();
    h.n_dirs = (uint32_t)s.dirs.size();
    uint64_t size = sizeof(BinHeader) + sizeof(lgi_camera) + sizeof(lgi_dir_light) + h.n_dirs * sizeof(lgi_dir_light) +
                    h.n_points * sizeof(lgi_point_light) + h.n_tris * sizeof(lgi_triangle);
    for (const HostMaterial& m : s.materials) size += 24 + 4 + m.diffuse_tex.size();
    h.file_bytes = size;
    FileCloser fc{std::fopen(path, "wb")};
    if (!fc.f) {
        err = std::string("cannot open ") + path + " for writing";
        return LGI_ERR_IO;
    }
    bool ok = std::fwrite(&h, sizeof h, 1, fc.f) == 1;
    ok = ok && std::fwrite(&s.camera, sizeof(lgi_camera), 1, fc.f) == 1;
    ok = ok && std::fwrite(&s.sun, sizeof(lgi_dir_light), 1, fc.f) == 1;
    ok = ok && put_records(fc.f, s.dirs.data(), sizeof(lgi_dir_light), s.dirs.size());
    ok = ok && put_records(fc.f, s.points.data(), sizeof(lgi_point_light), s.points.size());
    for (const HostMaterial& m : s.materials) {
        const uint32_t len = (uint32_t)m.diffuse_tex.size();
        ok = ok && std::fwrite(m.albedo, 4, 3, fc.f) == 3 && std::fwrite(m.emissive, 4, 3, fc.f) == 3;
        ok = ok && std::fwrite(&len, 4, 1, fc.f) == 1 && put_records(fc.f, m.diffuse_tex.data(), 1, len);
    }
    ok = ok && put_records(fc.f, s.tris.data(), sizeof(lgi_triangle), s.tris.size());
    ok = std::fclose(fc.f) == 0 && ok;
    fc.f = nullptr;
    if (!ok) {
        err = std::string("write to ") + path + " failed";
        return LGI_ERR_IO;
    }
    return LGI_OK;
}

lgi_status scene_load_bin(const char* path, HostScene& out, std::string& err, bool verbose) {
    FileCloser fc{std::fopen(path, "rb")};
    if (!fc.f) {
        err = std::string("cannot open ") + path;
        return LGI_ERR_IO;
    }
    std::fseek(fc.f, 0, SEEK_END);
    const long long fsize = std::ftell(fc.f);
    std::fseek(fc.f, 0, SEEK_SET);
    BinHeader h{};
    auto bad = [&](const char* what) {
        err = std::string(path) + ": " + what;
        return LGI_ERR_PARSE;
    };
    if (fsize < (long long)sizeof h || std::fread(&h, sizeof h, 1, fc.f) != 1) return bad("truncated header");
    if (std::memcmp(h.magic, kBinMagic, 8) != 0) return bad("not a binary scene file");
    if (h.version != kBinVersion) {
        err = std::string(path) + ": unsupported binary scene version " + std::to_string(h.version);
        return LGI_ERR_VERSION;
    }
    if (h.header_bytes != sizeof h || h.file_bytes != (uint64_t)fsize) return bad("size mismatch");
    const uint64_t fixed = sizeof h + sizeof(lgi_camera) + sizeof(lgi_dir_light);
    if (h.n_tris > (uint64_t)fsize / sizeof(lgi_triangle) || h.n_points > (uint64_t)fsize ||
        h.n_dirs > (uint64_t)fsize || h.n_mats > (uint64_t)fsize || fixed > (uint64_t)fsize)
        return bad("counts exceed the file size");
    HostScene s;
    s.dirs.resize(h.n_dirs);
    s.points.resize(h.n_points);
    s.materials.resize(h.n_mats);
    s.tris.resize(h.n_tris);
    bool ok = std::fread(&s.camera, sizeof(lgi_camera), 1, fc.f) == 1;
    ok = ok && std::fread(&s.sun, sizeof(lgi_dir_light), 1, fc.f) == 1;
    ok = ok && get_records(fc.f, s.dirs.data(), sizeof(lgi_dir_light), h.n_dirs);
    ok = ok && get_records(fc.f, s.points.data(), sizeof(lgi_point_light), h.n_points);
    for (HostMaterial& m : s.materials) {
        uint32_t len = 0;
        ok = ok && std::fread(m.albedo, 4, 3, fc.f) == 3 && std::fread(m.emissive, 4, 3, fc.f) == 3;
        ok = ok && std::fread(&len, 4, 1, fc.f) == 1 && len <= (uint64_t)fsize;
        if (ok) {
            m.diffuse_tex.resize(len);
            ok = get_records(fc.f, &m.diffuse_tex[0], 1, len);
        }
    }
    ok = ok && get_records(fc.f, s.tris.data(), sizeof(lgi_triangle), h.n_tris);
    if (!ok || std::ftell(fc.f) != fsize) return bad("truncated or inconsistent");
    out = std::move(s);
    if (verbose) {
        std::printf("[scene] loaded %s\n", path);
        std::printf("  camera: %dx%d, fov %.1f deg, pos (%.2f,%.2f,%.2f)\n", out.camera.width,
                    out.camera.height, out.camera.fov_y_deg, out.camera.pos[0], out.camera.pos[1],
                    out.camera.pos[2]);
        std::printf("  %zu directional + %zu point lights, %zu materials, %zu triangles\n",
                    out.dirs.size(), out.points.size(), out.materials.size(), out.tris.size());
    }
    return LGI_OK;
}
