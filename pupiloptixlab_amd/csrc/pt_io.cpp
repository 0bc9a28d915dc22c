// Scene file input (mitsuba-subset XML, Wavefront OBJ) and EXR output.
//
// XML semantics follow resource/xml/visitor.h:18-194 and
// resource/xml/object.cpp:9-62 (objects vs. properties, <default> with $name
// substitution, <ref id>), resource/scene.cpp:27-226 (integrator, sensor,
// film, emitters), resource/shape.cpp:85-266 (shapes) and
// resource/material.cpp:26-190 (BSDFs).  Transforms compose
// scale -> rotate -> translate regardless of element order
// (resource/xml/util_loader.cpp:128-192).
#include <algorithm>
#include <array>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <map>
#include <memory>
#include <sstream>

#include "pt_scene.h"

namespace pupil {
namespace {

// ---------------------------------------------------------------------------
// Minimal XML DOM.
struct XmlNode {
    std::string name;
    std::vector<std::pair<std::string, std::string>> attrs;
    std::vector<std::unique_ptr<XmlNode>> children;
    std::string attr(const std::string &k) const {
        for (auto &a : attrs)
            if (a.first == k) return a.second;
        return "";
    }
    bool has_attr(const std::string &k) const {
        for (auto &a : attrs)
            if (a.first == k) return true;
        return false;
    }
};

class XmlReader {
public:
    explicit XmlReader(const std::string &s) : s_(s) {}
    std::unique_ptr<XmlNode> parse(std::string &err) {
        std::unique_ptr<XmlNode> root;
        skip_misc();
        if (!element(root, err)) return nullptr;
        return root;
    }

private:
    void skip_ws() {
        while (i_ < s_.size() && std::isspace((unsigned char)s_[i_])) ++i_;
    }
    void skip_misc() {
        for (;;) {
            skip_ws();
            if (s_.compare(i_, 4, "<!--") == 0) {
                size_t e = s_.find("-->", i_ + 4);
                i_ = e == std::string::npos ? s_.size() : e + 3;
            } else if (s_.compare(i_, 2, "<?") == 0) {
                size_t e = s_.find("?>", i_ + 2);
                i_ = e == std::string::npos ? s_.size() : e + 2;
            } else if (s_.compare(i_, 2, "<!") == 0) {
                size_t e = s_.find('>', i_ + 2);
                i_ = e == std::string::npos ? s_.size() : e + 1;
            } else {
                return;
            }
        }
    }
    static std::string unescape(const std::string &v) {
        std::string r;
        for (size_t k = 0; k < v.size(); ++k) {
            if (v[k] == '&') {
                size_t e = v.find(';', k);
                if (e != std::string::npos) {
                    std::string ent = v.substr(k + 1, e - k - 1);
                    if (ent == "lt") r += '<';
                    else if (ent == "gt") r += '>';
                    else if (ent == "amp") r += '&';
                    else if (ent == "quot") r += '"';
                    else if (ent == "apos") r += '\'';
                    else r += v.substr(k, e - k + 1);
                    k = e;
                    continue;
                }
            }
            r += v[k];
        }
        return r;
    }
    bool element(std::unique_ptr<XmlNode> &out, std::string &err) {
        if (i_ >= s_.size() || s_[i_] != '<') {
            err = "parse: expected '<'";
            return false;
        }
        ++i_;
        auto node = std::make_unique<XmlNode>();
        while (i_ < s_.size() && !std::isspace((unsigned char)s_[i_]) && s_[i_] != '>' && s_[i_] != '/')
            node->name += s_[i_++];
        for (;;) {
            skip_ws();
            if (i_ >= s_.size()) {
                err = "parse: unterminated tag <" + node->name + ">";
                return false;
            }
            if (s_[i_] == '/') {
                if (s_.compare(i_, 2, "/>") != 0) {
                    err = "parse: bad tag end";
                    return false;
                }
                i_ += 2;
                out = std::move(node);
                return true;
            }
            if (s_[i_] == '>') {
                ++i_;
                break;
            }
            std::string key;
            while (i_ < s_.size() && s_[i_] != '=' && !std::isspace((unsigned char)s_[i_])) key += s_[i_++];
            skip_ws();
            if (i_ >= s_.size() || s_[i_] != '=') {
                err = "parse: attribute without value in <" + node->name + ">";
                return false;
            }
            ++i_;
            skip_ws();
            if (i_ >= s_.size() || (s_[i_] != '"' && s_[i_] != '\'')) {
                err = "parse: unquoted attribute";
                return false;
            }
            char q = s_[i_++];
            size_t e = s_.find(q, i_);
            if (e == std::string::npos) {
                err = "parse: unterminated attribute";
                return false;
            }
            node->attrs.emplace_back(key, unescape(s_.substr(i_, e - i_)));
            i_ = e + 1;
        }
        // children until </name>
        for (;;) {
            while (i_ < s_.size() && s_[i_] != '<') ++i_;  // text is ignored
            if (i_ >= s_.size()) {
                err = "parse: missing </" + node->name + ">";
                return false;
            }
            if (s_.compare(i_, 4, "<!--") == 0 || s_.compare(i_, 2, "<?") == 0) {
                skip_misc();
                continue;
            }
            if (s_.compare(i_, 2, "</") == 0) {
                size_t e = s_.find('>', i_);
                if (e == std::string::npos) {
                    err = "parse: bad closing tag";
                    return false;
                }
                std::string nm = s_.substr(i_ + 2, e - i_ - 2);
                while (!nm.empty() && std::isspace((unsigned char)nm.back())) nm.pop_back();
                if (nm != node->name) {
                    err = "parse: mismatched </" + nm + "> for <" + node->name + ">";
                    return false;
                }
                i_ = e + 1;
                out = std::move(node);
                return true;
            }
            std::unique_ptr<XmlNode> child;
            if (!element(child, err)) return false;
            node->children.push_back(std::move(child));
        }
    }
    const std::string &s_;
    size_t i_ = 0;
};

// ---------------------------------------------------------------------------
// Generic object tree of resource/xml/object.h.
struct Obj {
    std::string obj_name;   // element name (bsdf, shape, ...)
    std::string type;       // type attribute
    std::string id;
    std::string var_name;   // name attribute
    std::vector<std::pair<std::string, std::string>> props;
    std::vector<Obj *> subs;
    std::string prop(const std::string &k) const {
        for (auto &p : props)
            if (p.first == k) return p.second;
        return "";
    }
    Obj *sub(const std::string &n) const {
        for (Obj *o : subs)
            if (o->obj_name == n) return o;
        return nullptr;
    }
    Obj *param(const std::string &n) const {
        for (Obj *o : subs)
            if (o->var_name == n) return o;
        return nullptr;
    }
};

struct ObjTree {
    std::vector<std::unique_ptr<Obj>> pool;
    std::map<std::string, Obj *> ids;
    std::vector<std::pair<std::string, std::string>> defaults;
    Obj *root = nullptr;

    std::string subst(const std::string &v) const {
        if (v.find('$') == std::string::npos) return v;
        // Longer names first so "$resx" is not consumed by "$res".
        std::vector<std::pair<std::string, std::string>> d = defaults;
        std::stable_sort(d.begin(), d.end(), [](const auto &a, const auto &b) { return a.first.size() > b.first.size(); });
        std::string r = v;
        for (auto &kv : d) {
            std::string key = "$" + kv.first;
            size_t pos = 0;
            while ((pos = r.find(key, pos)) != std::string::npos) {
                r.replace(pos, key.size(), kv.second);
                pos += kv.second.size();
            }
        }
        return r;
    }
    std::string a(const XmlNode &n, const std::string &k) const { return subst(n.attr(k)); }

    Obj *make(const std::string &obj_name, const std::string &type) {
        pool.push_back(std::make_unique<Obj>());
        Obj *o = pool.back().get();
        o->obj_name = obj_name;
        o->type = type;
        return o;
    }

    void visit(const XmlNode &n, Obj *cur) {
        static const char *objects[] = {"bsdf", "emitter", "film", "integrator", "sensor", "shape", "texture", "transform"};
        static const char *props[] = {"integer", "string", "float", "rgb", "boolean", "matrix"};
        const std::string &nm = n.name;
        if (nm == "scene") {
            root = make("scene", a(n, "version"));
            for (auto &c : n.children) visit(*c, root);
            return;
        }
        if (nm == "default") {
            defaults.emplace_back(a(n, "name"), a(n, "value"));
            return;
        }
        if (nm == "ref") {
            auto it = ids.find(a(n, "id"));
            if (it != ids.end() && cur) cur->subs.push_back(it->second);
            return;
        }
        if (nm == "lookat") {
            Obj *o = make("lookat", "");
            o->props.emplace_back("origin", a(n, "origin"));
            o->props.emplace_back("target", a(n, "target"));
            o->props.emplace_back("up", a(n, "up"));
            if (cur) cur->subs.push_back(o);
            return;
        }
        if (nm == "rotate") {
            Obj *o = make("rotate", "");
            std::string axis;
            if (n.has_attr("value")) axis = a(n, "value");
            else if (n.has_attr("x")) axis = "1, 0, 0";
            else if (n.has_attr("y")) axis = "0, 1, 0";
            else if (n.has_attr("z")) axis = "0, 0, 1";
            o->props.emplace_back("axis", axis);
            o->props.emplace_back("angle", a(n, "angle"));
            if (cur) cur->subs.push_back(o);
            return;
        }
        if (nm == "scale" || nm == "point" || nm == "translate") {
            std::string name = a(n, "name");
            if (name.empty()) name = nm;
            std::string value = a(n, "value");
            if (value.empty()) {
                const char *dflt = nm == "scale" ? "1" : "0";
                std::string x = a(n, "x"), y = a(n, "y"), z = a(n, "z");
                value = (x.empty() ? dflt : x) + "," + (y.empty() ? dflt : y) + "," + (z.empty() ? dflt : z);
            }
            if (cur) cur->props.emplace_back(name, value);
            return;
        }
        for (const char *p : props) {
            if (nm == p) {
                std::string name = a(n, "name");
                if (name.empty()) name = nm;
                if (cur) cur->props.emplace_back(name, a(n, "value"));
                return;
            }
        }
        for (const char *ob : objects) {
            if (nm == ob) {
                Obj *o = make(nm, a(n, "type"));
                o->id = a(n, "id");
                o->var_name = a(n, "name");
                if (!o->id.empty()) ids[o->id] = o;
                if (cur) cur->subs.push_back(o);
                for (auto &c : n.children) visit(*c, o);
                return;
            }
        }
        // Unknown element: skipped together with its children (visitor.h:81-87).
    }
};

std::vector<std::string> split(const std::string &s, char sep) {
    std::vector<std::string> r;
    std::string cur;
    for (char c : s) {
        if (c == sep) {
            r.push_back(cur);
            cur.clear();
        } else {
            cur += c;
        }
    }
    r.push_back(cur);
    return r;
}
std::vector<std::string> split_ws(const std::string &s) {
    std::vector<std::string> r;
    std::istringstream is(s);
    std::string t;
    while (is >> t) r.push_back(t);
    return r;
}
bool to_float(const std::string &s, float &v) {
    try {
        size_t k = 0;
        v = std::stof(s, &k);
        return true;
    } catch (...) {
        return false;
    }
}
bool load_float(const Obj *o, const std::string &k, float &v, float dflt) {
    std::string s = o->prop(k);
    if (s.empty() || !to_float(s, v)) {
        v = dflt;
        return false;
    }
    return true;
}
bool load_int(const Obj *o, const std::string &k, int &v, int dflt) {
    std::string s = o->prop(k);
    if (s.empty()) {
        v = dflt;
        return false;
    }
    try {
        v = std::stoi(s);
    } catch (...) {
        v = dflt;
        return false;
    }
    return true;
}
bool load_float3(const std::string &s, float v[3], float d0, float d1, float d2, bool allow_scalar = true) {
    v[0] = d0;
    v[1] = d1;
    v[2] = d2;
    if (s.empty()) return false;
    auto xyz = split(s, ',');
    if (xyz.size() == 3) {
        for (int k = 0; k < 3; ++k)
            if (!to_float(xyz[k], v[k])) return false;
        return true;
    }
    if (xyz.size() == 1 && allow_scalar) {
        float f;
        if (!to_float(xyz[0], f)) return false;
        v[0] = v[1] = v[2] = f;
        return true;
    }
    return false;
}
bool load_bool(const Obj *o, const std::string &k, bool dflt) {
    std::string s = o->prop(k);
    if (s == "true") return true;
    if (s == "false") return false;
    return dflt;
}

struct Loader {
    ObjTree &tree;
    HostScene &scene;
    std::string dir;
    std::string err;
    std::map<std::string, std::shared_ptr<MeshData>> meshes;
    std::map<std::string, std::shared_ptr<BitmapData>> images;
    std::map<std::string, uint32_t> bitmap_index;
    // rectangle, cube and sphere shapes share one flip_normals flag per kind:
    // the last parsed instance wins (resource/shape.cpp:85-132, 441-497).
    std::vector<std::pair<uint32_t, int>> shared_flip_users;  // (shape index, 0 rect / 1 cube / 2 sphere)
    bool shared_flip[3] = {false, false, false};

    Mat4 load_transform(const Obj *t, bool &is_lookat) {
        is_lookat = false;
        Transform tr;
        if (!t) return tr.m;
        std::string mv = t->prop("matrix");
        if (!mv.empty()) {
            auto e = split_ws(mv);
            Mat4 m = Mat4::identity();
            if (e.size() == 16) {
                for (int i = 0; i < 16; ++i) to_float(e[i], m.e[i]);
            } else if (e.size() == 9) {
                for (int r = 0; r < 3; ++r)
                    for (int c = 0; c < 3; ++c) to_float(e[3 * r + c], m.e[4 * r + c]);
            } else {
                for (size_t i = 0; i < e.size() && i < 16; ++i) to_float(e[i], m.e[i]);
            }
            return m;
        }
        if (Obj *la = t->sub("lookat")) {
            float o[3], tg[3], up[3];
            load_float3(la->prop("origin"), o, 1.f, 0.f, 0.f, false);
            load_float3(la->prop("target"), tg, 0.f, 0.f, 0.f, false);
            load_float3(la->prop("up"), up, 0.f, 1.f, 0.f, false);
            tr.look_at(o, tg, up);
            for (int r = 0; r < 3; ++r) {  // mitsuba axis convention (util_loader.cpp:153-160)
                tr.m.e[4 * r + 0] *= -1.f;
                tr.m.e[4 * r + 2] *= -1.f;
            }
            is_lookat = true;
            return tr.m;
        }
        float s[3];
        if (load_float3(t->prop("scale"), s, 1.f, 1.f, 1.f)) tr.scale(s[0], s[1], s[2]);
        if (Obj *ro = t->sub("rotate")) {
            float ax[3], ang;
            if (load_float3(ro->prop("axis"), ax, 0.f, 0.f, 0.f, false) && load_float(ro, "angle", ang, 0.f))
                tr.rotate(ax[0], ax[1], ax[2], ang);
        }
        float tt[3];
        if (load_float3(t->prop("translate"), tt, 0.f, 0.f, 0.f, false)) tr.translate(tt[0], tt[1], tt[2]);
        return tr.m;
    }

    bool load_texture(const Obj *o, const std::string &name, pt_texture_desc &tex, float d0, float d1, float d2) {
        Obj *t = o->param(name);
        if (!t) {
            float c[3];
            std::string v = o->prop(name);
            load_float3(v, c, d0, d1, d2);
            tex = make_rgb_texture(c[0], c[1], c[2]);
            return true;
        }
        if (t->type == "checkerboard") {
            float c0[3], c1[3];
            load_float3(t->prop("color0"), c0, 0.4f, 0.4f, 0.4f);
            load_float3(t->prop("color1"), c1, 0.2f, 0.2f, 0.2f);
            tex = make_checker_texture(c0, c1);
        } else if (t->type == "bitmap") {
            // resource/scene.cpp:144-165: filter "bilinear" or point (also when
            // absent), wrap "repeat" (default), "mirror" or "clamp".
            const uint32_t filter = t->prop("filter_type") == "bilinear" ? PT_FILTER_LINEAR : PT_FILTER_POINT;
            const std::string wm = t->prop("wrap_mode");
            const uint32_t wrap = wm == "mirror" ? PT_WRAP_MIRROR : (wm == "clamp" ? PT_WRAP_CLAMP : PT_WRAP_REPEAT);
            uint32_t index;
            if (!bitmap(t->prop("filename"), filter, wrap, index)) return false;
            tex = make_bitmap_texture(index);
        } else {
            tex = make_rgb_texture(0.f, 0.f, 0.f);  // unknown texture type reads black
            return true;
        }
        // Texture transform (Texture::Sample applies rows 0 and 1 to (u, v, 0, 1)).
        if (Obj *tr = t->sub("transform")) {
            bool la = false;
            const Mat4 m = load_transform(tr, la);
            for (int k = 0; k < 4; ++k) {
                tex.uv_row0[k] = m.e[k];
                tex.uv_row1[k] = m.e[4 + k];
            }
        }
        return true;
    }

    // util::BitmapTexture loads through the texture manager's cache
    // (resource/texture.cpp): one image per file, one bitmap per sampling mode.
    bool bitmap(const std::string &fn, uint32_t filter, uint32_t wrap, uint32_t &index) {
        const std::string path = fn.size() && fn[0] == '/' ? fn : dir + "/" + fn;
        const std::string key = path + "|" + std::to_string(filter) + "|" + std::to_string(wrap);
        auto it = bitmap_index.find(key);
        if (it != bitmap_index.end()) {
            index = it->second;
            return true;
        }
        auto img = images.find(path);
        if (img == images.end()) {
            auto data = std::make_shared<BitmapData>();
            if (!load_image(path, data->width, data->height, data->rgba, err)) return false;
            img = images.emplace(path, data).first;
        }
        index = scene.add_bitmap(img->second, filter, wrap);
        bitmap_index[key] = index;
        return true;
    }

    float dielectric_ior(const Obj *o, const std::string &k, float dflt) {
        std::string v = o->prop(k);
        if (v.empty()) return dflt;
        float f;
        if (to_float(v, f)) return f;
        if (lookup_dielectric_ior(v, f)) return f;
        return dflt;
    }

    bool load_material(const Obj *o, pt_material_desc &m) {
        std::memset(&m, 0, sizeof(m));
        for (int i = 0; i < 4; ++i) m.tex[i] = make_rgb_texture(0.f, 0.f, 0.f);
        m.int_ior = 1.5046f;
        m.ext_ior = 1.000277f;
        if (!o) {  // no bsdf: Material{} of type Unknown; use the loader's diffuse default
            m.type = PT_MAT_DIFFUSE;
            m.tex[0] = make_rgb_texture(0.5f, 0.5f, 0.5f);
            return true;
        }
        const std::string &t = o->type;
        if (t == "twosided") {
            if (!load_material(o->sub("bsdf"), m)) return false;
            m.twosided = 1;
            return true;
        }
        if (t == "diffuse") {
            m.type = PT_MAT_DIFFUSE;
            return load_texture(o, "reflectance", m.tex[0], 0.5f, 0.5f, 0.5f);
        }
        if (t == "dielectric" || t == "roughdielectric") {
            m.type = t == "dielectric" ? PT_MAT_DIELECTRIC : PT_MAT_ROUGH_DIELECTRIC;
            m.int_ior = dielectric_ior(o, "int_ior", 1.5046f);
            m.ext_ior = dielectric_ior(o, "ext_ior", 1.000277f);
            if (!load_texture(o, "specular_reflectance", m.tex[0], 1.f, 1.f, 1.f)) return false;
            if (!load_texture(o, "specular_transmittance", m.tex[1], 1.f, 1.f, 1.f)) return false;
            if (m.type == PT_MAT_ROUGH_DIELECTRIC && !load_texture(o, "alpha", m.tex[2], 0.1f, 0.1f, 0.1f)) return false;
            return true;
        }
        if (t == "conductor" || t == "roughconductor") {
            m.type = t == "conductor" ? PT_MAT_CONDUCTOR : PT_MAT_ROUGH_CONDUCTOR;
            float eta[3] = {0.f, 0.f, 0.f}, k[3] = {1.f, 1.f, 1.f};
            const std::string metal = o->prop("material");
            if (!lookup_conductor_ior(metal, eta, k)) {
                // material.cpp:72-80, :92-96: an absent or unknown name is the
                // perfect mirror (eta 0, k 1).
                eta[0] = eta[1] = eta[2] = 0.f;
                k[0] = k[1] = k[2] = 1.f;
                if (!metal.empty())
                    scene.warnings.push_back("unknown conductor material '" + metal + "': using eta 0, k 1");
            }
            if (!load_texture(o, "specular_reflectance", m.tex[0], 1.f, 1.f, 1.f)) return false;
            if (!load_texture(o, "eta", m.tex[1], eta[0], eta[1], eta[2])) return false;
            if (!load_texture(o, "k", m.tex[2], k[0], k[1], k[2])) return false;
            if (m.type == PT_MAT_ROUGH_CONDUCTOR && !load_texture(o, "alpha", m.tex[3], 0.1f, 0.1f, 0.1f)) return false;
            return true;
        }
        if (t == "plastic" || t == "roughplastic") {
            m.type = t == "plastic" ? PT_MAT_PLASTIC : PT_MAT_ROUGH_PLASTIC;
            m.int_ior = dielectric_ior(o, "int_ior", 1.49f);
            m.ext_ior = dielectric_ior(o, "ext_ior", 1.000277f);
            m.nonlinear = o->prop("nonlinear") == "true" ? 1u : 0u;
            if (!load_texture(o, "diffuse_reflectance", m.tex[0], 0.5f, 0.5f, 0.5f)) return false;
            if (!load_texture(o, "specular_reflectance", m.tex[1], 1.f, 1.f, 1.f)) return false;
            if (m.type == PT_MAT_ROUGH_PLASTIC && !load_texture(o, "alpha", m.tex[2], 0.1f, 0.1f, 0.1f)) return false;
            return true;
        }
        // Unknown bsdf: the reference warns and uses an Unknown material that
        // shades black; map it to a black diffuse surface.
        scene.warnings.push_back("unknown bsdf type '" + t + "': shaded black");
        m.type = PT_MAT_DIFFUSE;
        m.tex[0] = make_rgb_texture(0.f, 0.f, 0.f);
        return true;
    }

    bool load_shape(const Obj *o) {
        std::shared_ptr<MeshData> mesh;
        uint32_t flags = 0;
        int shared_kind = -1;
        if (o->type == "rectangle" || o->type == "cube") {
            shared_kind = o->type == "rectangle" ? 0 : 1;
            static const MeshData rect = make_rectangle_mesh(), cube = make_cube_mesh();
            mesh = std::make_shared<MeshData>(shared_kind == 0 ? rect : cube);
            shared_flip[shared_kind] = load_bool(o, "flip_normals", false);
        } else if (o->type == "obj") {
            std::string fn = o->prop("filename");
            std::string path = fn.size() && fn[0] == '/' ? fn : dir + "/" + fn;
            auto it = meshes.find(path);
            if (it == meshes.end()) {
                auto m = std::make_shared<MeshData>();
                if (!load_obj_mesh(path, *m, err)) return false;
                meshes[path] = m;
                mesh = m;
            } else {
                mesh = it->second;
            }
            if (load_bool(o, "flip_tex_coords", true)) flags |= PT_SHAPE_FLIP_TEXCOORDS;
            if (load_bool(o, "flip_normals", false)) flags |= PT_SHAPE_FLIP_NORMALS;
        } else if (o->type == "sphere") {
            shared_kind = 2;
            shared_flip[2] = load_bool(o, "flip_normals", false);
        } else {
            scene.warnings.push_back("unknown shape type '" + o->type + "': skipped");
            return true;  // skipped with a warning in the reference
        }
        pt_material_desc mat;
        if (!load_material(o->sub("bsdf"), mat)) return false;
        uint32_t mi = scene.add_material(mat);
        bool lookat = false;
        Mat4 to_world = load_transform(o->sub("transform"), lookat);
        bool is_emitter = false;
        pt_texture_desc rad = make_rgb_texture(0.f, 0.f, 0.f);
        if (Obj *e = o->sub("emitter")) {
            if (e->type == "area") {
                is_emitter = true;
                if (!load_texture(e, "radiance", rad, 0.f, 0.f, 0.f)) return false;
            }
        }
        uint32_t si;
        if (shared_kind == 2) {
            // ShapeLoader<_sphere>: to_world * T(center) * S(radius) applied to
            // the unit sphere.
            float center[3], radius;
            load_float3(o->prop("center"), center, 0.f, 0.f, 0.f, false);
            load_float(o, "radius", radius, 1.f);
            Transform local;
            local.scale(radius, radius, radius);
            local.translate(center[0], center[1], center[2]);
            si = scene.add_sphere(to_world * local.m, mi, 0u, is_emitter, &rad);
        } else {
            si = scene.add_instance(*mesh, to_world, mi, flags, is_emitter, &rad);
        }
        if (shared_kind >= 0) shared_flip_users.emplace_back(si, shared_kind);
        return true;
    }

    bool run() {
        Obj *root = tree.root;
        if (!root) {
            err = "parse: no <scene> root";
            return false;
        }
        scene.max_depth = 1;
        int w = 768, h = 576;
        float fov = 90.f, near_clip = 0.01f, far_clip = 10000.f;
        Mat4 cam = Mat4::identity();
        bool have_sensor = false;
        for (Obj *o : root->subs) {
            if (o->obj_name == "integrator") {
                int md;
                load_int(o, "max_depth", md, 1);
                scene.max_depth = (uint32_t)std::max(1, md);
            } else if (o->obj_name == "sensor") {
                if (o->type != "perspective") continue;
                have_sensor = true;
                load_float(o, "fov", fov, 90.f);
                load_float(o, "near_clip", near_clip, 0.01f);
                load_float(o, "far_clip", far_clip, 10000.f);
                if (Obj *film = o->sub("film")) {
                    if (film->type == "hdrfilm") {
                        load_int(film, "width", w, 768);
                        load_int(film, "height", h, 576);
                    }
                }
                std::string axis = o->prop("fov_axis");
                char fov_axis = 'x';
                if (axis == "y" || axis == "Y") fov_axis = 'y';
                if (fov_axis == 'x') {  // resource/scene.cpp:122-127
                    float aspect = (float)h / (float)w;
                    float radian = fov * 3.14159265358979323846f / 180.f * 0.5f;
                    float t = std::tan(radian) * aspect;
                    fov = 2.f * std::atan(t) * 180.f / 3.14159265358979323846f;
                }
                bool lookat = false;
                cam = load_transform(o->sub("transform"), lookat);
                for (int r = 0; r < 3; ++r) {  // resource/scene.cpp:132-139
                    cam.e[4 * r + 0] *= -1.f;
                    cam.e[4 * r + 2] *= -1.f;
                }
            } else if (o->obj_name == "shape") {
                if (!load_shape(o)) return false;
            } else if (o->obj_name == "emitter") {
                if (o->type == "constant") {
                    float c[3];
                    load_float3(o->prop("radiance"), c, 0.f, 0.f, 0.f);
                    scene.env_type = PT_ENV_CONSTANT;
                    for (int k = 0; k < 3; ++k) scene.env_radiance[k] = c[k];
                } else if (o->type == "envmap") {
                    // resource/scene.cpp:209-222: bilinear, repeat; scale; the
                    // 3x3 part of to_world orients the map.
                    uint32_t index;
                    if (!bitmap(o->prop("filename"), PT_FILTER_LINEAR, PT_WRAP_REPEAT, index)) return false;
                    float sc;
                    load_float(o, "scale", sc, 1.f);
                    bool la = false;
                    const Mat4 m = load_transform(o->sub("transform"), la);
                    scene.env_type = PT_ENV_MAP;
                    scene.env_bitmap = index;
                    scene.env_scale = sc;
                    for (int r = 0; r < 3; ++r)
                        for (int c = 0; c < 3; ++c) scene.env_to_world[3 * r + c] = m.e[4 * r + c];
                }
            }
        }
        for (auto &u : shared_flip_users) {
            if (shared_flip[u.second]) scene.shapes[u.first].flags |= PT_SHAPE_FLIP_NORMALS;
        }
        if (w <= 0 || h <= 0) {
            err = "parse: film size must be positive";
            return false;
        }
        scene.width = (uint32_t)w;
        scene.height = (uint32_t)h;
        (void)have_sensor;
        Mat4 s2c = sample_to_camera_matrix(fov, (float)w / (float)h, near_clip, far_clip);
        std::memcpy(scene.camera.sample_to_camera, s2c.e, sizeof(s2c.e));
        std::memcpy(scene.camera.camera_to_world, cam.e, sizeof(cam.e));
        return true;
    }
};

}  // namespace

bool load_xml_scene(const std::string &path, HostScene &out, std::string &err) {
    std::ifstream f(path, std::ios::binary);
    if (!f) {
        err = "io: cannot open " + path;
        return false;
    }
    std::stringstream ss;
    ss << f.rdbuf();
    std::string text = ss.str();
    XmlReader rd(text);
    auto dom = rd.parse(err);
    if (!dom) return false;
    if (dom->name != "scene") {
        err = "parse: root element is <" + dom->name + ">, expected <scene>";
        return false;
    }
    ObjTree tree;
    tree.visit(*dom, nullptr);
    out = HostScene();
    Loader ld{tree, out, "", ""};
    size_t slash = path.find_last_of('/');
    ld.dir = slash == std::string::npos ? "." : path.substr(0, slash);
    if (!ld.run()) {
        err = ld.err;
        return false;
    }
    return true;
}

// Ear clipping of a polygon with five or more corners (assimp's
// aiProcess_Triangulate clips ears of larger polygons for the reference,
// resource/shape.cpp:273).  The corners are projected onto the coordinate
// plane most perpendicular to the Newell normal; an ear is a convex corner
// whose triangle with its two neighbours holds no other remaining corner.
// The search starts at corner 1 and resumes at the corner after the one
// clipped, so a convex polygon gives the fan from corner 0 -- (0, k, k + 1)
// -- and every triangle keeps the polygon's winding.  When no ear is left
// (degenerate or self-intersecting input) the remaining corners are fanned.
static std::vector<std::array<int, 3>> triangulate_face(const std::vector<std::array<double, 3>> &P) {
    const size_t n = P.size();
    double nrm[3] = {0.0, 0.0, 0.0};
    for (size_t a = 0; a < n; ++a) {
        const std::array<double, 3> &p0 = P[a], &p1 = P[(a + 1) % n];
        nrm[0] += (p0[1] - p1[1]) * (p0[2] + p1[2]);
        nrm[1] += (p0[2] - p1[2]) * (p0[0] + p1[0]);
        nrm[2] += (p0[0] - p1[0]) * (p0[1] + p1[1]);
    }
    int k = 0;
    for (int c = 1; c < 3; ++c)
        if (std::fabs(nrm[c]) > std::fabs(nrm[k])) k = c;
    const int iu = (k + 1) % 3, iv = (k + 2) % 3;  // cyclic: 2D winding sign = sign of nrm[k]
    const double orient = nrm[k] >= 0.0 ? 1.0 : -1.0;
    auto cross2 = [&](size_t a, size_t b, size_t c) {
        return ((P[b][iu] - P[a][iu]) * (P[c][iv] - P[a][iv]) - (P[b][iv] - P[a][iv]) * (P[c][iu] - P[a][iu])) *
               orient;
    };
    std::vector<size_t> rem(n);
    for (size_t a = 0; a < n; ++a) rem[a] = a;
    std::vector<std::array<int, 3>> out;
    size_t i = 1;
    while (rem.size() > 3) {
        const size_t m = rem.size();
        bool clipped = false;
        for (size_t tries = 0; tries < m; ++tries, ++i) {
            const size_t c = i % m, pr = (c + m - 1) % m, nx = (c + 1) % m;
            const size_t a = rem[pr], b = rem[c], d = rem[nx];
            if (!(cross2(a, b, d) > 0.0)) continue;  // reflex or degenerate corner
            bool inside = false;
            for (size_t q = 0; q < m && !inside; ++q) {
                if (q == pr || q == c || q == nx) continue;
                const size_t e = rem[q];
                inside = cross2(a, b, e) >= 0.0 && cross2(b, d, e) >= 0.0 && cross2(d, a, e) >= 0.0;
            }
            if (inside) continue;
            out.push_back({(int)a, (int)b, (int)d});
            rem.erase(rem.begin() + (std::ptrdiff_t)c);
            i = c;  // the corner after the clipped one now sits at position c
            clipped = true;
            break;
        }
        if (!clipped) {
            for (size_t q = 1; q + 1 < rem.size(); ++q) out.push_back({(int)rem[0], (int)rem[q], (int)rem[q + 1]});
            return out;
        }
    }
    out.push_back({(int)rem[0], (int)rem[1], (int)rem[2]});
    return out;
}

// Wavefront OBJ: v / vt / vn / f (polygons triangulated as below); one vertex
// per face corner.
bool load_obj_mesh(const std::string &path, MeshData &mesh, std::string &err) {
    std::ifstream f(path);
    if (!f) {
        err = "io: cannot open " + path;
        return false;
    }
    std::vector<float> v, vt, vn;
    mesh = MeshData();
    bool any_n = false, any_t = false;
    struct Corner { int v, t, n; };
    std::vector<Corner> corners;
    std::vector<uint32_t> idx;
    std::string line;
    while (std::getline(f, line)) {
        std::istringstream is(line);
        std::string tag;
        if (!(is >> tag)) continue;
        if (tag == "v") {
            float x, y, z;
            is >> x >> y >> z;
            v.insert(v.end(), {x, y, z});
        } else if (tag == "vt") {
            float a, b;
            is >> a >> b;
            vt.insert(vt.end(), {a, b});
        } else if (tag == "vn") {
            float x, y, z;
            is >> x >> y >> z;
            vn.insert(vn.end(), {x, y, z});
        } else if (tag == "f") {
            std::vector<Corner> face;
            std::string tok;
            while (is >> tok) {
                Corner c{0, 0, 0};
                auto parts = split(tok, '/');
                auto rd = [](const std::string &s, int count) {
                    if (s.empty()) return 0;
                    int i = std::stoi(s);
                    return i < 0 ? count + i + 1 : i;
                };
                try {
                    c.v = rd(parts[0], (int)v.size() / 3);
                    if (parts.size() > 1) c.t = rd(parts[1], (int)vt.size() / 2);
                    if (parts.size() > 2) c.n = rd(parts[2], (int)vn.size() / 3);
                } catch (...) {
                    err = "parse: bad face in " + path;
                    return false;
                }
                if (c.v <= 0 || c.v > (int)v.size() / 3) {
                    err = "parse: vertex index out of range in " + path;
                    return false;
                }
                face.push_back(c);
            }
            // Quads fan out from their concave (reflex) corner, if they have
            // one (a quad has at most one), the triangulation assimp's
            // aiProcess_Triangulate gives the reference (resource/shape.cpp:273)
            // for quads; convex quads keep the fan from the first corner.
            // Reflex: the turn at the corner opposes the polygon normal
            // (Newell).  Larger polygons are ear-clipped (triangulate_face).
            if (face.size() >= 5) {
                std::vector<std::array<double, 3>> P(face.size());
                for (size_t a = 0; a < face.size(); ++a)
                    for (int c = 0; c < 3; ++c) P[a][(size_t)c] = v[3 * (size_t)(face[a].v - 1) + c];
                for (const std::array<int, 3> &tri : triangulate_face(P)) {
                    for (int k : tri) {
                        const Corner &c = face[(size_t)k];
                        idx.push_back((uint32_t)corners.size());
                        corners.push_back(c);
                        any_n = any_n || c.n > 0;
                        any_t = any_t || c.t > 0;
                    }
                }
                continue;
            }
            if (face.size() == 4) {
                double P[4][3];
                for (int a = 0; a < 4; ++a)
                    for (int c = 0; c < 3; ++c) P[a][c] = v[3 * (size_t)(face[(size_t)a].v - 1) + c];
                double n[3] = {0.0, 0.0, 0.0};
                for (int a = 0; a < 4; ++a) {
                    const double *p0 = P[a], *p1 = P[(a + 1) % 4];
                    n[0] += (p0[1] - p1[1]) * (p0[2] + p1[2]);
                    n[1] += (p0[2] - p1[2]) * (p0[0] + p1[0]);
                    n[2] += (p0[0] - p1[0]) * (p0[1] + p1[1]);
                }
                int start = 0;
                for (int a = 0; a < 4; ++a) {
                    const double *pp = P[(a + 3) % 4], *pc = P[a], *pn = P[(a + 1) % 4];
                    const double e0[3] = {pc[0] - pp[0], pc[1] - pp[1], pc[2] - pp[2]};
                    const double e1[3] = {pn[0] - pc[0], pn[1] - pc[1], pn[2] - pc[2]};
                    const double t[3] = {e0[1] * e1[2] - e0[2] * e1[1], e0[2] * e1[0] - e0[0] * e1[2],
                                         e0[0] * e1[1] - e0[1] * e1[0]};
                    if (t[0] * n[0] + t[1] * n[1] + t[2] * n[2] < 0.0) {
                        start = a;
                        break;
                    }
                }
                std::rotate(face.begin(), face.begin() + start, face.end());
            }
            for (size_t k = 1; k + 1 < face.size(); ++k) {
                for (const Corner &c : {face[0], face[k], face[k + 1]}) {
                    idx.push_back((uint32_t)corners.size());
                    corners.push_back(c);
                    any_n = any_n || c.n > 0;
                    any_t = any_t || c.t > 0;
                }
            }
        }
    }
    for (const Corner &c : corners) {
        mesh.positions.insert(mesh.positions.end(), v.begin() + 3 * (c.v - 1), v.begin() + 3 * c.v);
        if (any_n) {
            if (c.n > 0 && c.n <= (int)vn.size() / 3)
                mesh.normals.insert(mesh.normals.end(), vn.begin() + 3 * (c.n - 1), vn.begin() + 3 * c.n);
            else
                mesh.normals.insert(mesh.normals.end(), {0.f, 0.f, 0.f});
        }
        if (any_t) {
            if (c.t > 0 && c.t <= (int)vt.size() / 2)
                mesh.texcoords.insert(mesh.texcoords.end(), vt.begin() + 2 * (c.t - 1), vt.begin() + 2 * c.t);
            else
                mesh.texcoords.insert(mesh.texcoords.end(), {0.f, 0.f});
        }
    }
    mesh.indices = idx;
    return true;
}

// Single-part scanline EXR, no compression, FLOAT channels B, G, R.  Buffer
// row 0 is the bottom of the image, so rows are written in reverse
// (util/texture.cpp:37-44).
bool write_exr_rgb(const std::string &path, const float *rgba, uint32_t w, uint32_t h, std::string &err) {
    std::vector<unsigned char> out;
    auto put = [&](const void *p, size_t n) {
        const unsigned char *c = static_cast<const unsigned char *>(p);
        out.insert(out.end(), c, c + n);
    };
    auto put_str = [&](const char *s) { put(s, std::strlen(s) + 1); };
    auto put_i32 = [&](int32_t v) { put(&v, 4); };
    auto put_f32 = [&](float v) { put(&v, 4); };
    const unsigned char magic[4] = {0x76, 0x2f, 0x31, 0x01};
    put(magic, 4);
    put_i32(2);
    put_str("channels");
    put_str("chlist");
    put_i32(3 * (2 + 16) + 1);
    for (const char *ch : {"B", "G", "R"}) {
        put_str(ch);
        put_i32(2);  // FLOAT
        const unsigned char z[4] = {0, 0, 0, 0};
        put(z, 4);
        put_i32(1);
        put_i32(1);
    }
    out.push_back(0);
    put_str("compression");
    put_str("compression");
    put_i32(1);
    out.push_back(0);
    for (const char *nm : {"dataWindow", "displayWindow"}) {
        put_str(nm);
        put_str("box2i");
        put_i32(16);
        put_i32(0);
        put_i32(0);
        put_i32((int32_t)w - 1);
        put_i32((int32_t)h - 1);
    }
    put_str("lineOrder");
    put_str("lineOrder");
    put_i32(1);
    out.push_back(0);
    put_str("pixelAspectRatio");
    put_str("float");
    put_i32(4);
    put_f32(1.f);
    put_str("screenWindowCenter");
    put_str("v2f");
    put_i32(8);
    put_f32(0.f);
    put_f32(0.f);
    put_str("screenWindowWidth");
    put_str("float");
    put_i32(4);
    put_f32(1.f);
    out.push_back(0);
    const uint64_t line_bytes = 8 + (uint64_t)w * 3 * 4;
    const uint64_t table_start = out.size();
    for (uint32_t y = 0; y < h; ++y) {
        uint64_t off = table_start + (uint64_t)h * 8 + (uint64_t)y * line_bytes;
        put(&off, 8);
    }
    std::vector<float> row(w);
    for (uint32_t y = 0; y < h; ++y) {
        put_i32((int32_t)y);
        put_i32((int32_t)(w * 3 * 4));
        const float *src = rgba + (size_t)(h - 1 - y) * w * 4;
        for (int ch : {2, 1, 0}) {
            for (uint32_t x = 0; x < w; ++x) row[x] = src[4 * x + ch];
            put(row.data(), (size_t)w * 4);
        }
    }
    FILE *fp = std::fopen(path.c_str(), "wb");
    if (!fp) {
        err = "io: cannot write " + path;
        return false;
    }
    size_t wr = std::fwrite(out.data(), 1, out.size(), fp);
    std::fclose(fp);
    if (wr != out.size()) {
        err = "io: short write " + path;
        return false;
    }
    return true;
}

}  // namespace pupil
