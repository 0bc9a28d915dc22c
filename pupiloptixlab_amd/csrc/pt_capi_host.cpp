// Host-side C-ABI entry points: scene loading/generation, EXR output and a
// CPU traversal of the same compressed wide BVH the kernels use.
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "../../include/pupil_hip.h"
#include "pt_bvh_build.h"
#include "pt_scene.h"

struct pt_host_scene {
    pupil::HostScene scene;
};

namespace {
struct HostStack {
    uint32_t a[pupil::kMaxBvhDepth + 2], b[pupil::kMaxBvhDepth + 2];
    uint32_t sp = 0;
    void push(uint32_t x, uint32_t y) {
        a[sp] = x;
        b[sp] = y;
        ++sp;
    }
    void pop(uint32_t &x, uint32_t &y) {
        --sp;
        x = a[sp];
        y = b[sp];
    }
    bool empty() const { return sp == 0; }
};
void copy_err(const std::string &s, char *err, size_t len) {
    if (!err || len == 0) return;
    std::strncpy(err, s.c_str(), len - 1);
    err[len - 1] = 0;
}
}  // namespace

namespace {
// A host allocation that fails inside an entry point is reported as
// PT_ERR_OOM instead of escaping the C interface.
template <typename F>
int guard_oom(std::nullptr_t, const char *, F &&body) {
    try {
        return body();
    } catch (const std::bad_alloc &) {
        return PT_ERR_OOM;
    }
}
}  // namespace

extern "C" {

static int pt_scene_load_xml_unguarded(const char *path, pt_host_scene **out, char *err, size_t err_len);
int pt_scene_load_xml(const char *path, pt_host_scene **out, char *err, size_t err_len) {
    return guard_oom(nullptr, "pt_scene_load_xml", [&] {
        return pt_scene_load_xml_unguarded(path, out, err, err_len);
    });
}
static int pt_scene_load_xml_unguarded(const char *path, pt_host_scene **out, char *err, size_t err_len) {
    if (!path || !out) return PT_ERR_INVALID_ARG;
    *out = nullptr;
    pt_host_scene *hs = new (std::nothrow) pt_host_scene();
    if (!hs) return PT_ERR_OOM;
    std::string e;
    if (!pupil::load_xml_scene(path, hs->scene, e)) {
        copy_err(e, err, err_len);
        delete hs;
        return e.rfind("io:", 0) == 0 ? PT_ERR_IO : PT_ERR_PARSE;
    }
    *out = hs;
    return PT_OK;
}

int pt_scene_get_warnings(const pt_host_scene *scene, char *buf, size_t len) {
    if (!scene) return PT_ERR_INVALID_ARG;
    std::string all;
    for (const auto &w : scene->scene.warnings) all += w + "\n";
    copy_err(all, buf, len);
    return (int)scene->scene.warnings.size();
}

static int pt_scene_generate_unguarded(int kind, const pt_gen_params *params, pt_host_scene **out);
int pt_scene_generate(int kind, const pt_gen_params *params, pt_host_scene **out) {
    return guard_oom(nullptr, "pt_scene_generate", [&] { return pt_scene_generate_unguarded(kind, params, out); });
}
static int pt_scene_generate_unguarded(int kind, const pt_gen_params *params, pt_host_scene **out) {
    if (!out) return PT_ERR_INVALID_ARG;
    *out = nullptr;
    pt_gen_params p;
    std::memset(&p, 0, sizeof(p));
    if (params) p = *params;
    pt_host_scene *hs = new (std::nothrow) pt_host_scene();
    if (!hs) return PT_ERR_OOM;
    std::string e;
    if (!pupil::generate_scene(kind, p, hs->scene, e)) {
        delete hs;
        return PT_ERR_INVALID_ARG;
    }
    *out = hs;
    return PT_OK;
}

int pt_scene_get_desc(const pt_host_scene *scene, pt_scene_desc *out) {
    if (!scene || !out) return PT_ERR_INVALID_ARG;
    *out = scene->scene.desc();
    return PT_OK;
}

void pt_scene_destroy(pt_host_scene *scene) { delete scene; }

static int pt_load_image_unguarded(const char *path, uint32_t *width, uint32_t *height, float *rgba, size_t capacity, char *err,
                  size_t err_len);
int pt_load_image(const char *path, uint32_t *width, uint32_t *height, float *rgba, size_t capacity, char *err,
                  size_t err_len) {
    return guard_oom(nullptr, "pt_load_image", [&] {
        return pt_load_image_unguarded(path, width, height, rgba, capacity, err, err_len);
    });
}
static int pt_load_image_unguarded(const char *path, uint32_t *width, uint32_t *height, float *rgba, size_t capacity, char *err,
                  size_t err_len) {
    if (!path || !width || !height) return PT_ERR_INVALID_ARG;
    uint32_t w = 0, h = 0;
    std::vector<float> data;
    std::string e;
    if (!pupil::load_image(path, w, h, data, e)) {
        copy_err(e, err, err_len);
        return e.rfind("io:", 0) == 0 ? PT_ERR_IO : (e.rfind("unsupported", 0) == 0 ? PT_ERR_UNSUPPORTED : PT_ERR_PARSE);
    }
    *width = w;
    *height = h;
    if (rgba) {
        if (capacity < data.size()) return PT_ERR_CAPACITY;
        std::memcpy(rgba, data.data(), data.size() * sizeof(float));
    }
    return PT_OK;
}

static int pt_write_exr_unguarded(const char *path, const float *rgba, uint32_t width, uint32_t height);
int pt_write_exr(const char *path, const float *rgba, uint32_t width, uint32_t height) {
    return guard_oom(nullptr, "pt_write_exr", [&] { return pt_write_exr_unguarded(path, rgba, width, height); });
}
static int pt_write_exr_unguarded(const char *path, const float *rgba, uint32_t width, uint32_t height) {
    if (!path || !rgba || width == 0 || height == 0) return PT_ERR_INVALID_ARG;
    std::string e;
    return pupil::write_exr_rgb(path, rgba, width, height, e) ? PT_OK : PT_ERR_IO;
}

static int pt_host_trace_rays_unguarded(const pt_scene_desc *s, const float *rays, uint32_t count, int any_hit, void *hits,
                       uint64_t *node_visits, uint64_t *tri_tests);
int pt_host_trace_rays(const pt_scene_desc *s, const float *rays, uint32_t count, int any_hit, void *hits,
                       uint64_t *node_visits, uint64_t *tri_tests) {
    return guard_oom(nullptr, "pt_host_trace_rays", [&] {
        return pt_host_trace_rays_unguarded(s, rays, count, any_hit, hits, node_visits, tri_tests);
    });
}
static int pt_host_trace_rays_unguarded(const pt_scene_desc *s, const float *rays, uint32_t count, int any_hit, void *hits,
                       uint64_t *node_visits, uint64_t *tri_tests) {
    if (!s || (!rays && count) || (!hits && count)) return PT_ERR_INVALID_ARG;
    if (s->num_triangles && (!s->positions || !s->indices)) return PT_ERR_INVALID_ARG;
    if (s->num_spheres && !s->spheres) return PT_ERR_INVALID_ARG;
    std::vector<float> sph4((size_t)s->num_spheres * 4, 0.f);
    for (uint32_t k = 0; k < s->num_spheres; ++k) pupil::sphere_center_radius(s->spheres[k].to_world, &sph4[4 * k]);
    std::vector<float> pos4((size_t)s->num_vertices * 4, 0.f);
    for (uint32_t i = 0; i < s->num_vertices; ++i)
        for (int k = 0; k < 3; ++k) pos4[4 * i + k] = s->positions[3 * i + k];
    std::vector<uint32_t> tri4((size_t)s->num_triangles * 4, 0u);
    for (uint32_t t = 0; t < s->num_triangles; ++t)
        for (int k = 0; k < 3; ++k) tri4[4 * t + k] = s->indices[3 * t + k];
    std::vector<pupil::WideNode> nodes;
    std::vector<pupil::LeafTri> leaf;
    pupil::BvhBuildStats st;
    std::string err;
    if (!pupil::build_wide_bvh(pos4.data(), s->num_vertices, tri4.data(), s->num_triangles, sph4.data(),
                               s->num_spheres, pupil::BvhBuildOptions{}, nodes, leaf, st, err))
        return PT_ERR_INVALID_ARG;
    uint32_t *h = static_cast<uint32_t *>(hits);
    uint64_t nv = 0, tv = 0;
    for (uint32_t i = 0; i < count; ++i) {
        const float *r = rays + 8 * (size_t)i;
        HostStack stack;
        pupil::Hit hit;
        pupil::TraverseStats ts{0, 0};
        auto ln = [&](uint32_t k, pupil::NodeWords &n) { std::memcpy(&n, &nodes[k], sizeof(n)); };
        auto lt = [&](uint32_t k, pupil::LeafTri &t) { t = leaf[k]; };
        if (any_hit)
            pupil::traverse_bvh<true>(ln, lt, pupil::mk3(r[0], r[1], r[2]), pupil::mk3(r[4], r[5], r[6]), r[3], r[7],
                                      stack, hit, &ts);
        else
            pupil::traverse_bvh<false>(ln, lt, pupil::mk3(r[0], r[1], r[2]), pupil::mk3(r[4], r[5], r[6]), r[3], r[7],
                                       stack, hit, &ts);
        nv += ts.nodes;
        tv += ts.tris;
        h[4 * i] = pupil::f2u(hit.t);
        h[4 * i + 1] = hit.prim;
        h[4 * i + 2] = pupil::f2u(hit.b1);
        h[4 * i + 3] = pupil::f2u(hit.b2);
    }
    if (node_visits) *node_visits = nv;
    if (tri_tests) *tri_tests = tv;
    return PT_OK;
}

}  // extern "C"
