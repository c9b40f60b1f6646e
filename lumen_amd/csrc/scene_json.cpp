// scene_json.cpp — Lumen JSON + OBJ scene loader (LumenScene.cpp:231-513).
//
// The JSON reader is a small recursive-descent parser (objects, arrays,
// numbers, strings, true/false/null); the OBJ reader restates the parts of
// tinyobjloader the reference relies on (libs/tiny_obj_loader.h): `v`/`vn`/`vt`
// records, `f` with v, v/vt, v//vn or v/vt/vn indices (1-based or negative),
// `o` and `g` starting a new shape (tiny_obj_loader.h:2770-2840), quads split
// along their shorter diagonal (:1427-1531).  Polygons with more than four
// vertices are fanned from their first vertex (tinyobjloader ear-clips them;
// none of the reference's scene files has such faces).  A material's
// "texture" (LumenScene.cpp:357-360, a path relative to the JSON file) is
// decoded as a baseline JPEG into 8-bit RGBA (image_jpeg.cpp) and stored as
// linear floats through the sRGB transfer function, the conversion that
// sampling a VK_FORMAT_R8G8B8A8_SRGB image applies before filtering
// (LumenScene.cpp:205-212); each reference adds its own texture, as there.
#include <algorithm>
#include <array>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iterator>
#include <map>
#include <memory>
#include <sstream>

#include "scene.h"

namespace lt {
namespace {

// ------------------------------------------------------------------ JSON ----
struct JVal {
    enum Kind { NUL, BOOL, NUM, STR, ARR, OBJ } kind = NUL;
    bool b = false;
    double num = 0.0;
    std::string str;
    std::vector<JVal> arr;
    std::vector<std::pair<std::string, JVal>> obj;

    const JVal* get(const char* key) const {
        if (kind != OBJ) return nullptr;
        for (auto& kv : obj)
            if (kv.first == key) return kv.second.kind == NUL ? nullptr : &kv.second;
        return nullptr;
    }
};

struct JParser {
    const char* p;
    const char* end;
    std::string err;

    void ws() {
        while (p < end && (*p == ' ' || *p == '\t' || *p == '\n' || *p == '\r')) p++;
    }
    bool fail(const std::string& m) {
        if (err.empty()) err = m;
        return false;
    }
    bool lit(const char* s) {
        size_t n = strlen(s);
        if ((size_t)(end - p) < n || strncmp(p, s, n) != 0) return false;
        p += n;
        return true;
    }
    bool str(std::string& out) {
        if (p >= end || *p != '"') return fail("expected string");
        p++;
        out.clear();
        while (p < end && *p != '"') {
            if (*p == '\\') {
                p++;
                if (p >= end) break;
                char c = *p;
                switch (c) {
                    case 'n': out += '\n'; break;
                    case 't': out += '\t'; break;
                    case 'r': out += '\r'; break;
                    case 'b': out += '\b'; break;
                    case 'f': out += '\f'; break;
                    case 'u':  // keep ASCII code points, replace others
                        if (end - p >= 5) {
                            unsigned cp = (unsigned)strtoul(std::string(p + 1, p + 5).c_str(), nullptr, 16);
                            out += cp < 128 ? (char)cp : '?';
                            p += 4;
                        }
                        break;
                    default: out += c;
                }
                p++;
            } else {
                out += *p++;
            }
        }
        if (p >= end) return fail("unterminated string");
        p++;
        return true;
    }
    bool value(JVal& v, int depth) {
        if (depth > 64) return fail("JSON nested too deeply");
        ws();
        if (p >= end) return fail("unexpected end of JSON");
        if (*p == '{') {
            p++;
            v.kind = JVal::OBJ;
            ws();
            if (p < end && *p == '}') {
                p++;
                return true;
            }
            for (;;) {
                ws();
                std::string k;
                if (!str(k)) return false;
                ws();
                if (p >= end || *p != ':') return fail("expected ':'");
                p++;
                JVal c;
                if (!value(c, depth + 1)) return false;
                v.obj.emplace_back(std::move(k), std::move(c));
                ws();
                if (p < end && *p == ',') {
                    p++;
                    continue;
                }
                if (p < end && *p == '}') {
                    p++;
                    return true;
                }
                return fail("expected ',' or '}'");
            }
        }
        if (*p == '[') {
            p++;
            v.kind = JVal::ARR;
            ws();
            if (p < end && *p == ']') {
                p++;
                return true;
            }
            for (;;) {
                JVal c;
                if (!value(c, depth + 1)) return false;
                v.arr.push_back(std::move(c));
                ws();
                if (p < end && *p == ',') {
                    p++;
                    continue;
                }
                if (p < end && *p == ']') {
                    p++;
                    return true;
                }
                return fail("expected ',' or ']'");
            }
        }
        if (*p == '"') {
            v.kind = JVal::STR;
            return str(v.str);
        }
        if (lit("true")) {
            v.kind = JVal::BOOL;
            v.b = true;
            return true;
        }
        if (lit("false")) {
            v.kind = JVal::BOOL;
            return true;
        }
        if (lit("null")) {
            v.kind = JVal::NUL;
            return true;
        }
        char* q = nullptr;
        std::string tmp(p, (size_t)std::min<ptrdiff_t>(end - p, 64));
        double d = strtod(tmp.c_str(), &q);
        if (q == tmp.c_str()) return fail("unexpected character in JSON");
        p += q - tmp.c_str();
        v.kind = JVal::NUM;
        v.num = d;
        return true;
    }
};

float num(const JVal* v, float def) {
    if (!v) return def;
    if (v->kind == JVal::NUM) return (float)v->num;
    if (v->kind == JVal::BOOL) return v->b ? 1.0f : 0.0f;
    return def;
}
bool vec3(const JVal* v, float out[3]) {
    if (!v || v->kind != JVal::ARR || v->arr.size() < 3) return false;
    for (int k = 0; k < 3; k++) out[k] = num(&v->arr[k], 0.0f);
    return true;
}

// ------------------------------------------------------------------- OBJ ----
struct ObjShape {
    std::string name;
    std::vector<float> pos, nrm, uv;  // 3 / 3 / 2 floats per face vertex
};

struct Idx {
    long v, vt, vn;  // 0-based, -1 = absent
};

bool parse_idx(const std::string& tok, size_t nv, size_t nvt, size_t nvn, Idx& o) {
    long a[3] = {0, 0, 0};
    bool has[3] = {false, false, false};
    size_t k = 0, i = 0;
    while (k < 3) {
        size_t j = tok.find('/', i);
        std::string part = tok.substr(i, j == std::string::npos ? std::string::npos : j - i);
        if (!part.empty()) {
            a[k] = strtol(part.c_str(), nullptr, 10);
            has[k] = true;
        }
        k++;
        if (j == std::string::npos) break;
        i = j + 1;
    }
    auto fix = [](long x, size_t n) -> long { return x > 0 ? x - 1 : (long)n + x; };  // 1-based / relative
    if (!has[0]) return false;
    o.v = fix(a[0], nv);
    o.vt = has[1] ? fix(a[1], nvt) : -1;
    o.vn = has[2] ? fix(a[2], nvn) : -1;
    return o.v >= 0 && (size_t)o.v < nv && o.vt < (long)nvt && o.vn < (long)nvn;
}

bool load_obj(const std::string& path, std::vector<ObjShape>& shapes, std::string& err) {
    std::ifstream f(path);
    if (!f) {
        err = "cannot open mesh file " + path;
        return false;
    }
    std::vector<float> v, vn, vt;
    ObjShape cur;
    auto flush = [&]() {
        if (!cur.pos.empty()) shapes.push_back(cur);
        cur = ObjShape();
    };
    std::string line;
    size_t line_no = 0;
    while (std::getline(f, line)) {
        line_no++;
        while (!line.empty() && (line.back() == '\r' || line.back() == ' ' || line.back() == '\t')) line.pop_back();
        size_t s0 = line.find_first_not_of(" \t");
        if (s0 == std::string::npos || line[s0] == '#') continue;
        std::istringstream ss(line.substr(s0));
        std::string tag;
        ss >> tag;
        if (tag == "v") {
            float x = 0, y = 0, z = 0;
            ss >> x >> y >> z;
            v.insert(v.end(), {x, y, z});
        } else if (tag == "vn") {
            float x = 0, y = 0, z = 0;
            ss >> x >> y >> z;
            vn.insert(vn.end(), {x, y, z});
        } else if (tag == "vt") {
            float x = 0, y = 0;
            ss >> x >> y;
            vt.insert(vt.end(), {x, y});
        } else if (tag == "o" || tag == "g") {
            flush();
            std::string rest;
            std::getline(ss, rest);
            size_t b = rest.find_first_not_of(" \t");
            rest = b == std::string::npos ? std::string() : rest.substr(b);
            if (tag == "g") {  // multiple group names are joined by one space
                std::istringstream gs(rest);
                std::string w, joined;
                while (gs >> w) joined += (joined.empty() ? "" : " ") + w;
                rest = joined;
            }
            cur.name = rest;
        } else if (tag == "f") {
            std::vector<Idx> face;
            std::string tok;
            while (ss >> tok) {
                Idx ix;
                if (!parse_idx(tok, v.size() / 3, vt.size() / 2, vn.size() / 3, ix)) {
                    err = path + ":" + std::to_string(line_no) + ": bad face index '" + tok + "'";
                    return false;
                }
                face.push_back(ix);
            }
            if (face.size() < 3) continue;  // degenerate face (skipped like tinyobjloader)
            std::vector<std::array<int, 3>> tris;
            if (face.size() == 4) {  // tiny_obj_loader.h:1474-1522: split along the shorter diagonal
                const float* p0 = &v[3 * face[0].v];
                const float* p1 = &v[3 * face[1].v];
                const float* p2 = &v[3 * face[2].v];
                const float* p3 = &v[3 * face[3].v];
                const float e02[3] = {p2[0] - p0[0], p2[1] - p0[1], p2[2] - p0[2]};
                const float e13[3] = {p3[0] - p1[0], p3[1] - p1[1], p3[2] - p1[2]};
                const float s02 = e02[0] * e02[0] + e02[1] * e02[1] + e02[2] * e02[2];
                const float s13 = e13[0] * e13[0] + e13[1] * e13[1] + e13[2] * e13[2];
                if (s02 < s13)
                    tris = {{0, 1, 2}, {0, 2, 3}};
                else
                    tris = {{0, 1, 3}, {1, 2, 3}};
            } else {
                for (size_t k = 2; k < face.size(); k++) tris.push_back({0, (int)k - 1, (int)k});
            }
            for (auto& t : tris)
                for (int c : t) {
                    const Idx& ix = face[c];
                    cur.pos.insert(cur.pos.end(), {v[3 * ix.v], v[3 * ix.v + 1], v[3 * ix.v + 2]});
                    if (ix.vn >= 0)
                        cur.nrm.insert(cur.nrm.end(), {vn[3 * ix.vn], vn[3 * ix.vn + 1], vn[3 * ix.vn + 2]});
                    else
                        cur.nrm.insert(cur.nrm.end(), {0.0f, 0.0f, 0.0f});
                    if (ix.vt >= 0)
                        cur.uv.insert(cur.uv.end(), {vt[2 * ix.vt], vt[2 * ix.vt + 1]});
                    else
                        cur.uv.insert(cur.uv.end(), {0.0f, 0.0f});
                }
        }
        // usemtl, mtllib, s, l, p and other records do not change the geometry
    }
    flush();
    return true;
}

// ------------------------------------------------------------- materials ----
// LumenScene.cpp:350-475
// sRGB transfer function (IEC 61966-2-1), 8-bit code -> linear
float srgb_to_linear(uint32_t c) {
    const double v = c / 255.0;
    return (float)(v <= 0.04045 ? v / 12.92 : std::pow((v + 0.055) / 1.055, 2.4));
}

bool load_texture(const std::string& path, lt_host_scene& s, int32_t& id, std::string& err) {
    std::ifstream f(path, std::ios::binary);
    if (!f) {
        err = "cannot open texture " + path;
        return false;
    }
    std::vector<uint8_t> bytes((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
    std::vector<uint8_t> rgba;
    uint32_t w = 0, h = 0;
    std::string derr;
    if (decode_jpeg(bytes.data(), bytes.size(), rgba, w, h, derr) != LT_OK) {
        err = path + ": " + derr;
        return false;
    }
    float lut[256];
    for (uint32_t c = 0; c < 256; c++) lut[c] = srgb_to_linear(c);
    std::vector<float> texels((size_t)w * h * 4);
    for (size_t k = 0; k < (size_t)w * h; k++) {
        for (int c = 0; c < 3; c++) texels[4 * k + c] = lut[rgba[4 * k + c]];
        texels[4 * k + 3] = rgba[4 * k + 3] / 255.0f;  // alpha is stored linear
    }
    s.tex_data.push_back(std::move(texels));
    s.tex_dims.push_back(w);
    s.tex_dims.push_back(h);
    id = (int32_t)s.tex_data.size() - 1;
    return true;
}

bool make_material(const JVal& b, lt_material& m, const std::string& dir, lt_host_scene& s, std::string& err) {
    m = lt_material{};
    m.texture_id = -1;
    if (const JVal* tex = b.get("texture")) {
        if (tex->kind != JVal::STR) {
            err = "\"texture\" is not a string";
            return false;
        }
        if (!load_texture(dir + tex->str, s, m.texture_id, err)) return false;
    }
    if (!vec3(b.get("albedo"), m.albedo)) m.albedo[0] = m.albedo[1] = m.albedo[2] = 1.0f;
    vec3(b.get("emissive_factor"), m.emissive_factor);
    const JVal* t = b.get("type");
    const std::string type = t && t->kind == JVal::STR ? t->str : "";
    if (type == "diffuse") {
        m.bsdf_type = LT_BSDF_DIFFUSE;
        m.bsdf_props = LT_BSDF_FLAG_DIFFUSE | LT_BSDF_FLAG_REFLECTION;
    } else if (type == "mirror") {
        m.bsdf_type = LT_BSDF_MIRROR;
        m.bsdf_props = LT_BSDF_FLAG_SPECULAR | LT_BSDF_FLAG_REFLECTION;
    } else if (type == "glass") {
        m.bsdf_type = LT_BSDF_GLASS;
        m.bsdf_props = LT_BSDF_FLAG_SPECULAR | LT_BSDF_FLAG_TRANSMISSION;
        m.ior = num(b.get("ior"), 0.0f);
    } else if (type == "dielectric") {
        m.bsdf_type = LT_BSDF_DIELECTRIC;
        m.ior = num(b.get("ior"), 1.0f);
        m.roughness = num(b.get("roughness"), 0.0f);
        const JVal* tr = b.get("transmission");
        const JVal* rf = b.get("reflection");
        if (!tr || num(tr, 1.0f) != 0.0f) m.bsdf_props |= LT_BSDF_FLAG_TRANSMISSION;
        if (!rf || num(rf, 1.0f) != 0.0f) m.bsdf_props |= LT_BSDF_FLAG_REFLECTION;
        if (m.ior != 1.0f && (double)m.roughness > 0.08)
            m.bsdf_props |= LT_BSDF_FLAG_GLOSSY;
        else
            m.bsdf_props |= LT_BSDF_FLAG_SPECULAR;
        m.thin = (uint32_t)num(b.get("thin"), 0.0f);
    } else if (type == "conductor") {
        m.bsdf_type = LT_BSDF_CONDUCTOR;
        m.roughness = num(b.get("roughness"), 0.0f);
        float r[3];
        if (vec3(b.get("reflectance"), r)) {  // LumenScene.cpp:47-50
            for (int k = 0; k < 3; k++) {
                const float c = fminf(fmaxf(r[k], 0.0f), 0.9999f);
                m.albedo[k] = 1.0f;
                m.k[k] = 2.0f * sqrtf(c) / sqrtf(fmaxf(1.0f - c, 0.001f));
            }
        }
        float et[3], rv[3];
        if (vec3(b.get("edge_tint"), et)) {  // Gulbrandsen 2014, LumenScene.cpp:424-435
            if (!vec3(b.get("reflectivity"), rv)) {
                err = "conductor edge_tint needs reflectivity";
                return false;
            }
            for (int k = 0; k < 3; k++) {
                const float sr = sqrtf(rv[k]);
                m.albedo[k] = et[k] * (1.0f - rv[k]) / (1.0f + rv[k]) + (1.0f - et[k]) * (1.0f + sr) / (1.0f - sr);
                const float i1 = rv[k] * (m.albedo[k] + 1.0f), i2 = m.albedo[k] - 1.0f;
                m.k[k] = sqrtf(1.0f / (1.0f - rv[k]) * (i1 * i1 - i2 * i2));
            }
        }
        m.bsdf_props = LT_BSDF_FLAG_REFLECTION;
        m.bsdf_props |= (double)m.roughness > 0.08 ? LT_BSDF_FLAG_GLOSSY : LT_BSDF_FLAG_SPECULAR;
    } else if (type == "principled") {
        m.bsdf_type = LT_BSDF_PRINCIPLED;
        float a[3] = {1.0f, 1.0f, 1.0f};
        vec3(b.get("albedo"), a);
        memcpy(m.albedo, a, 12);
        m.ior = num(b.get("ior"), 1.0f);
        m.roughness = num(b.get("roughness"), 0.5f);
        m.diffuse_trans = num(b.get("diffuse_transmission"), 0.0f);
        m.spec_trans = num(b.get("specular_transmission"), 0.0f);
        m.metallic = num(b.get("metallic"), 0.0f);
        m.specular_tint = num(b.get("specular_tint"), 0.0f);
        m.sheen_tint = num(b.get("sheen_tint"), 0.5f);
        m.clearcoat = num(b.get("clearcoat"), 0.0f);
        m.clearcoat_gloss = num(b.get("clearcoat_gloss"), 1.0f);
        m.subsurface = num(b.get("subsurface"), 0.0f);
        m.flatness = num(b.get("flatness"), 0.0f);
        m.sheen = num(b.get("sheen"), 0.0f);
        m.anisotropy = num(b.get("anisotropy"), 0.0f);
        m.thin = (uint32_t)num(b.get("thin"), 0.0f);
        if (m.roughness < 1.0f) m.bsdf_props |= LT_BSDF_FLAG_REFLECTION;
        if (m.spec_trans > 0.0f) m.bsdf_props |= LT_BSDF_FLAG_TRANSMISSION;
        m.bsdf_props |= (double)m.roughness > 0.08 ? LT_BSDF_FLAG_GLOSSY : LT_BSDF_FLAG_SPECULAR;
    }
    // Any other type (e.g. "disney" in scenes/cornell_box/cornell_box_disney.json)
    // falls through the reference's chain of type tests (LumenScene.cpp:372-475)
    // and keeps the zero-initialised bsdf_type and bsdf_props with the albedo,
    // emission and texture read above: a surface that emits but never scatters
    // (every BSDF switch of the shaders leaves f and the pdf at zero for it).
    return true;
}

}  // namespace

lt_status load_json_scene(const char* path, lt_host_scene** out, std::string& err) {
    *out = nullptr;
    std::ifstream f(path, std::ios::binary);
    if (!f) {
        err = std::string("cannot open ") + path;
        return LT_ERR_INVALID_ARG;
    }
    std::stringstream buf;
    buf << f.rdbuf();
    const std::string text = buf.str();
    JParser jp{text.data(), text.data() + text.size(), {}};
    JVal root;
    if (!jp.value(root, 0) || root.kind != JVal::OBJ) {
        err = "JSON parse error: " + (jp.err.empty() ? std::string("top level is not an object") : jp.err);
        return LT_ERR_PARSE;
    }
    std::string dir = path;
    const size_t slash = dir.find_last_of("/\\");
    dir = slash == std::string::npos ? std::string() : dir.substr(0, slash + 1);

    std::unique_ptr<lt_host_scene> s(new lt_host_scene);
    // integrator block (LumenScene.cpp:237-268)
    const JVal* integ = root.get("integrator");
    s->icfg.kind = -1;
    s->icfg.path_length = 6;
    s->icfg.base_radius = 0.03f;
    if (integ) {
        const JVal* ty = integ->get("type");
        const std::string type = ty && ty->kind == JVal::STR ? ty->str : "";
        s->icfg.kind = type == "path" ? LT_INTEGRATOR_PATH
                       : type == "bdpt" ? LT_INTEGRATOR_BDPT
                       : type == "vcm"  ? LT_INTEGRATOR_VCM
                       : type == "sppm" ? LT_INTEGRATOR_SPPM
                                        : -1;
        s->icfg.path_length = (int32_t)num(integ->get("path_length"), 6.0f);
        vec3(integ->get("sky_col"), s->icfg.sky_col);
        s->icfg.enable_vm = num(integ->get("enable_vm"), 0.0f) == 1.0f ? 1 : 0;
        s->icfg.radius_factor = num(integ->get("radius_factor"), 0.025f);
        s->icfg.base_radius = num(integ->get("base_radius"), 0.03f);  // SceneConfig.h:34
    }
    // meshes (LumenScene.cpp:270-334)
    const JVal* mf = root.get("mesh_file");
    if (!mf || mf->kind != JVal::STR) {
        err = "scene has no mesh_file";
        return LT_ERR_PARSE;
    }
    std::vector<ObjShape> shapes;
    if (!load_obj(dir + mf->str, shapes, err)) return LT_ERR_PARSE;
    // materials and their refs (LumenScene.cpp:336-485): a shape takes the
    // last bsdf naming it, material 0 otherwise
    std::vector<uint32_t> shape_mat(shapes.size(), 0u);
    const JVal* bsdfs = root.get("bsdfs");
    if (bsdfs && bsdfs->kind == JVal::ARR) {
        for (size_t i = 0; i < bsdfs->arr.size(); i++) {
            lt_material m;
            if (!make_material(bsdfs->arr[i], m, dir, *s, err)) {
                err = "bsdf " + std::to_string(i) + ": " + err;
                return LT_ERR_PARSE;
            }
            s->materials.push_back(m);
            const JVal* refs = bsdfs->arr[i].get("refs");
            if (refs && refs->kind == JVal::ARR)
                for (auto& r : refs->arr)
                    for (size_t k = 0; k < shapes.size(); k++)
                        if (r.kind == JVal::STR && r.str == shapes[k].name) shape_mat[k] = (uint32_t)i;
        }
    }
    if (s->materials.empty()) {
        err = "scene has no bsdfs";
        return LT_ERR_PARSE;
    }
    for (size_t k = 0; k < shapes.size(); k++)
        add_mesh(*s, shapes[k].name, shapes[k].pos, shapes[k].nrm, shapes[k].uv, shape_mat[k]);
    // camera (LumenScene.cpp:487-491)
    const JVal* cam = root.get("camera");
    if (cam) {
        s->cam_fov = num(cam->get("fov"), 45.0f);
        vec3(cam->get("position"), s->cam_pos);
        vec3(cam->get("dir"), s->cam_dir);
    }
    // analytic lights (LumenScene.cpp:493-512)
    std::vector<AnalyticLight> al;
    const JVal* lights = root.get("lights");
    if (lights && lights->kind == JVal::ARR) {
        for (auto& l : lights->arr) {
            AnalyticLight a{};
            vec3(l.get("pos"), a.pos);
            vec3(l.get("dir"), a.to);
            vec3(l.get("L"), a.L);
            const JVal* ty = l.get("type");
            const std::string type = ty && ty->kind == JVal::STR ? ty->str : "";
            if (type == "spot")
                a.flags = LT_LIGHT_SPOT | LT_LIGHT_FINITE | LT_LIGHT_DELTA;
            else if (type == "directional")
                a.flags = LT_LIGHT_DIRECTIONAL | LT_LIGHT_DELTA;
            a.cos_width = cosf(3.14159265359f / 6.0f);
            a.cos_falloff = cosf(25.0f * 3.14159265359f / 180.0f);
            al.push_back(a);
        }
    }
    finalize_scene(*s, al);
    texture_views(*s);
    *out = s.release();
    return LT_OK;
}

}  // namespace lt
