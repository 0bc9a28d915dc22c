#include "pt_scene.h"

#include <cmath>
#include <cstring>
#include <map>

namespace pupil {

// ---------------------------------------------------------------------------
// Mat4 / Transform
Mat4 Mat4::identity() {
    Mat4 m;
    for (int i = 0; i < 16; ++i) m.e[i] = (i % 5 == 0) ? 1.f : 0.f;
    return m;
}
Mat4 Mat4::operator*(const Mat4 &b) const {
    Mat4 r;
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) {
            float s = e[i * 4 + 0] * b.e[0 * 4 + j];
            s = s + e[i * 4 + 1] * b.e[1 * 4 + j];
            s = s + e[i * 4 + 2] * b.e[2 * 4 + j];
            s = s + e[i * 4 + 3] * b.e[3 * 4 + j];
            r.e[i * 4 + j] = s;
        }
    return r;
}
Mat4 Mat4::transpose() const {
    Mat4 r;
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) r.e[i * 4 + j] = e[j * 4 + i];
    return r;
}
static bool invert4(const double *m, double *inv) {
    double a[4][8];
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 8; ++j) a[i][j] = j < 4 ? m[i * 4 + j] : (j - 4 == i ? 1.0 : 0.0);
    for (int c = 0; c < 4; ++c) {
        int piv = c;
        for (int r = c + 1; r < 4; ++r)
            if (std::fabs(a[r][c]) > std::fabs(a[piv][c])) piv = r;
        if (a[piv][c] == 0.0) return false;
        if (piv != c)
            for (int j = 0; j < 8; ++j) std::swap(a[c][j], a[piv][j]);
        double d = a[c][c];
        for (int j = 0; j < 8; ++j) a[c][j] /= d;
        for (int r = 0; r < 4; ++r) {
            if (r == c) continue;
            double f = a[r][c];
            if (f == 0.0) continue;
            for (int j = 0; j < 8; ++j) a[r][j] -= f * a[c][j];
        }
    }
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) inv[i * 4 + j] = a[i][j + 4];
    return true;
}
Mat4 Mat4::inverse() const {
    double m[16], inv[16];
    for (int i = 0; i < 16; ++i) m[i] = e[i];
    Mat4 r = identity();
    if (!invert4(m, inv)) return r;
    for (int i = 0; i < 16; ++i) r.e[i] = (float)inv[i];
    return r;
}

void Transform::scale(float x, float y, float z) {
    Mat4 s = Mat4::identity();
    s.e[0] = x;
    s.e[5] = y;
    s.e[10] = z;
    m = s * m;
}
void Transform::translate(float x, float y, float z) {
    Mat4 t = Mat4::identity();
    t.e[3] = x;
    t.e[7] = y;
    t.e[11] = z;
    m = t * m;
}
// Quaternion rotation about a unit axis (util/transform.cpp:7-50).
void Transform::rotate(float ux, float uy, float uz, float angle) {
    float len = std::sqrt(ux * ux + uy * uy + uz * uz);
    ux /= len;
    uy /= len;
    uz /= len;
    float theta = angle / 180.f * 3.14159265358979323846f;
    float a = std::cos(0.5f * theta);
    float b = std::sin(0.5f * theta) * ux;
    float c = std::sin(0.5f * theta) * uy;
    float d = std::sin(0.5f * theta) * uz;
    Mat4 r;
    const float v[16] = {1.f - 2.f * c * c - 2.f * d * d, 2.f * b * c - 2.f * a * d, 2.f * a * c + 2.f * b * d, 0.f,
                         2.f * b * c + 2.f * a * d, 1.f - 2.f * b * b - 2.f * d * d, 2.f * c * d - 2.f * a * b, 0.f,
                         2.f * b * d - 2.f * a * c, 2.f * a * b + 2.f * c * d, 1.f - 2.f * b * b - 2.f * c * c, 0.f,
                         0.f, 0.f, 0.f, 1.f};
    std::memcpy(r.e, v, sizeof(v));
    m = r * m;
}
// Camera-to-world of a right-handed look-at: z axis = normalize(eye - target),
// x = normalize(cross(up, z)), y = cross(z, x) (util/transform.cpp:87-98).
void Transform::look_at(const float o[3], const float t[3], const float up[3]) {
    double z[3] = {(double)o[0] - t[0], (double)o[1] - t[1], (double)o[2] - t[2]};
    double zl = std::sqrt(z[0] * z[0] + z[1] * z[1] + z[2] * z[2]);
    for (double &v : z) v /= zl;
    double x[3] = {up[1] * z[2] - up[2] * z[1], up[2] * z[0] - up[0] * z[2], up[0] * z[1] - up[1] * z[0]};
    double xl = std::sqrt(x[0] * x[0] + x[1] * x[1] + x[2] * x[2]);
    for (double &v : x) v /= xl;
    double y[3] = {z[1] * x[2] - z[2] * x[1], z[2] * x[0] - z[0] * x[2], z[0] * x[1] - z[1] * x[0]};
    Mat4 r = Mat4::identity();
    for (int i = 0; i < 3; ++i) {
        r.e[i * 4 + 0] = (float)x[i];
        r.e[i * 4 + 1] = (float)y[i];
        r.e[i * 4 + 2] = (float)z[i];
        r.e[i * 4 + 3] = o[i];
    }
    m = r;
}
void transform_point(const Mat4 &mm, const float p[3], float out[3]) {
    const float *m = mm.e;
    float x = m[0] * p[0] + m[1] * p[1] + m[2] * p[2] + m[3];
    float y = m[4] * p[0] + m[5] * p[1] + m[6] * p[2] + m[7];
    float z = m[8] * p[0] + m[9] * p[1] + m[10] * p[2] + m[11];
    float w = m[12] * p[0] + m[13] * p[1] + m[14] * p[2] + m[15];
    out[0] = x / w;
    out[1] = y / w;
    out[2] = z / w;
}
void transform_normal_unnormalized(const Mat4 &inv_t, const float n[3], float out[3]) {
    const float *m = inv_t.e;
    out[0] = m[0] * n[0] + m[1] * n[1] + m[2] * n[2];
    out[1] = m[4] * n[0] + m[5] * n[1] + m[6] * n[2];
    out[2] = m[8] * n[0] + m[9] * n[1] + m[10] * n[2];
}

// Inverse of (P_rh * T(1,1,0) * S(0.5,0.5,1)) in the row-vector convention of
// the projection, transposed to the column-vector convention of the kernels.
Mat4 sample_to_camera_matrix(float fov_y_deg, float aspect, float n, float f) {
    double half = 0.5 * (double)(fov_y_deg / 180.f * 3.14159265358979323846f);
    double h = std::cos(half) / std::sin(half);
    double w = h / aspect;
    double fr = (double)f / ((double)n - (double)f);
    double P[16] = {w, 0, 0, 0, 0, h, 0, 0, 0, 0, fr, -1, 0, 0, fr * n, 0};
    double T[16] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 1, 1, 0, 1};
    double S[16] = {0.5, 0, 0, 0, 0, 0.5, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1};
    double PT[16], PTS[16], inv[16];
    auto mul = [](const double *a, const double *b, double *r) {
        for (int i = 0; i < 4; ++i)
            for (int j = 0; j < 4; ++j) {
                double s = 0;
                for (int k = 0; k < 4; ++k) s += a[i * 4 + k] * b[k * 4 + j];
                r[i * 4 + j] = s;
            }
    };
    mul(P, T, PT);
    mul(PT, S, PTS);
    Mat4 r = Mat4::identity();
    if (!invert4(PTS, inv)) return r;
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) r.e[i * 4 + j] = (float)inv[j * 4 + i];
    return r;
}

// ---------------------------------------------------------------------------
// Textures / materials
pt_texture_desc make_rgb_texture(float r, float g, float b) {
    pt_texture_desc t;
    std::memset(&t, 0, sizeof(t));
    t.type = PT_TEX_RGB;
    t.color0[0] = r;
    t.color0[1] = g;
    t.color0[2] = b;
    t.uv_row0[0] = 1.f;
    t.uv_row1[1] = 1.f;
    return t;
}
pt_texture_desc make_bitmap_texture(uint32_t bitmap) {
    pt_texture_desc t = make_rgb_texture(0.f, 0.f, 0.f);
    t.type = PT_TEX_BITMAP;
    t.bitmap = bitmap;
    return t;
}
pt_texture_desc make_checker_texture(const float c0[3], const float c1[3]) {
    pt_texture_desc t = make_rgb_texture(c0[0], c0[1], c0[2]);
    t.type = PT_TEX_CHECKER;
    for (int i = 0; i < 3; ++i) t.color1[i] = c1[i];
    return t;
}
pt_material_desc make_diffuse(float r, float g, float b, bool twosided) {
    pt_material_desc m;
    std::memset(&m, 0, sizeof(m));
    m.type = PT_MAT_DIFFUSE;
    m.twosided = twosided ? 1u : 0u;
    m.int_ior = 1.5046f;
    m.ext_ior = 1.000277f;
    m.tex[0] = make_rgb_texture(r, g, b);
    for (int i = 1; i < 4; ++i) m.tex[i] = make_rgb_texture(0.f, 0.f, 0.f);
    return m;
}

// ---------------------------------------------------------------------------
// Meshes
MeshData make_rectangle_mesh() {
    MeshData m;
    m.positions = {-1.f, -1.f, 0.f, 1.f, -1.f, 0.f, 1.f, 1.f, 0.f, -1.f, 1.f, 0.f};
    m.normals = {0.f, 0.f, 1.f, 0.f, 0.f, 1.f, 0.f, 0.f, 1.f, 0.f, 0.f, 1.f};
    m.texcoords = {0.f, 0.f, 1.f, 0.f, 1.f, 1.f, 0.f, 1.f};
    m.indices = {0, 1, 2, 0, 2, 3};
    return m;
}
MeshData make_cube_mesh() {
    MeshData m;
    // Six faces, four vertices each, counter-clockwise seen from outside.
    const int axes[6] = {0, 0, 1, 1, 2, 2};
    const float signs[6] = {-1.f, 1.f, -1.f, 1.f, -1.f, 1.f};
    for (int f = 0; f < 6; ++f) {
        int a = axes[f];
        float s = signs[f];
        int u = (a + 1) % 3, v = (a + 2) % 3;
        const float cu[4] = {-1.f, 1.f, 1.f, -1.f};
        const float cv[4] = {-1.f, -1.f, 1.f, 1.f};
        uint32_t base = (uint32_t)(m.positions.size() / 3);
        for (int k = 0; k < 4; ++k) {
            float p[3];
            p[a] = s;
            p[u] = cu[k] * (s > 0.f ? 1.f : -1.f);
            p[v] = cv[k];
            m.positions.insert(m.positions.end(), p, p + 3);
            float n[3] = {0.f, 0.f, 0.f};
            n[a] = s;
            m.normals.insert(m.normals.end(), n, n + 3);
            m.texcoords.push_back(k == 1 || k == 2 ? 1.f : 0.f);
            m.texcoords.push_back(k >= 2 ? 1.f : 0.f);
        }
        m.indices.insert(m.indices.end(), {base, base + 1, base + 2, base, base + 2, base + 3});
    }
    return m;
}
MeshData make_icosphere_mesh(int subdivisions) {
    const float t = (1.f + std::sqrt(5.f)) / 2.f;
    std::vector<float> v = {-1, t, 0, 1, t, 0, -1, -t, 0, 1, -t, 0, 0, -1, t, 0, 1, t,
                            0, -1, -t, 0, 1, -t, t, 0, -1, t, 0, 1, -t, 0, -1, -t, 0, 1};
    std::vector<uint32_t> f = {0, 11, 5, 0, 5, 1, 0, 1, 7, 0, 7, 10, 0, 10, 11, 1, 5, 9, 5, 11, 4,
                               11, 10, 2, 10, 7, 6, 7, 1, 8, 3, 9, 4, 3, 4, 2, 3, 2, 6, 3, 6, 8,
                               3, 8, 9, 4, 9, 5, 2, 4, 11, 6, 2, 10, 8, 6, 7, 9, 8, 1};
    auto normalize_vertex = [&](uint32_t i) {
        double x = v[3 * i], y = v[3 * i + 1], z = v[3 * i + 2];
        double l = std::sqrt(x * x + y * y + z * z);
        v[3 * i] = (float)(x / l);
        v[3 * i + 1] = (float)(y / l);
        v[3 * i + 2] = (float)(z / l);
    };
    for (uint32_t i = 0; i < 12; ++i) normalize_vertex(i);
    for (int s = 0; s < subdivisions; ++s) {
        std::map<uint64_t, uint32_t> mid;
        auto midpoint = [&](uint32_t a, uint32_t b) {
            uint64_t key = a < b ? ((uint64_t)a << 32 | b) : ((uint64_t)b << 32 | a);
            auto it = mid.find(key);
            if (it != mid.end()) return it->second;
            uint32_t idx = (uint32_t)(v.size() / 3);
            for (int k = 0; k < 3; ++k) v.push_back(0.5f * (v[3 * a + k] + v[3 * b + k]));
            normalize_vertex(idx);
            mid[key] = idx;
            return idx;
        };
        std::vector<uint32_t> nf;
        for (size_t i = 0; i < f.size(); i += 3) {
            uint32_t a = f[i], b = f[i + 1], c = f[i + 2];
            uint32_t ab = midpoint(a, b), bc = midpoint(b, c), ca = midpoint(c, a);
            nf.insert(nf.end(), {a, ab, ca, b, bc, ab, c, ca, bc, ab, bc, ca});
        }
        f.swap(nf);
    }
    MeshData m;
    m.positions = v;
    m.normals = v;
    m.indices = f;
    return m;
}

// ---------------------------------------------------------------------------
// HostScene
uint32_t HostScene::add_material(const pt_material_desc &m) {
    materials.push_back(m);
    return (uint32_t)materials.size() - 1;
}

uint32_t HostScene::add_instance(const MeshData &mesh, const Mat4 &to_world, uint32_t material, uint32_t flags,
                                 bool is_emitter, const pt_texture_desc *radiance) {
    const uint32_t base = (uint32_t)(positions.size() / 3);
    const uint32_t nv = (uint32_t)(mesh.positions.size() / 3);
    const bool has_n = mesh.normals.size() == mesh.positions.size() && !mesh.normals.empty();
    const bool has_uv = mesh.texcoords.size() / 2 == nv && nv > 0;
    Mat4 inv_t = to_world.inverse().transpose();
    double det;
    {
        const float *m = to_world.e;
        det = (double)m[0] * ((double)m[5] * m[10] - (double)m[6] * m[9]) -
              (double)m[1] * ((double)m[4] * m[10] - (double)m[6] * m[8]) +
              (double)m[2] * ((double)m[4] * m[9] - (double)m[5] * m[8]);
    }
    for (uint32_t i = 0; i < nv; ++i) {
        float p[3];
        transform_point(to_world, &mesh.positions[3 * i], p);
        positions.insert(positions.end(), p, p + 3);
        float n[3] = {0.f, 0.f, 0.f};
        if (has_n) transform_normal_unnormalized(inv_t, &mesh.normals[3 * i], n);
        normals.insert(normals.end(), n, n + 3);
        if (has_uv) {
            texcoords.push_back(mesh.texcoords[2 * i]);
            texcoords.push_back(mesh.texcoords[2 * i + 1]);
        } else {
            texcoords.push_back(0.f);
            texcoords.push_back(0.f);
        }
    }
    any_normals = any_normals || has_n;
    any_texcoords = any_texcoords || has_uv;
    pt_shape_desc s;
    std::memset(&s, 0, sizeof(s));
    s.material = material;
    s.flags = flags | (has_n ? PT_SHAPE_HAS_NORMALS : 0u) | (has_uv ? PT_SHAPE_HAS_TEXCOORDS : 0u) |
              (det < 0.0 ? PT_SHAPE_NEG_FACE_NORMAL : 0u);
    s.first_tri = (uint32_t)(indices.size() / 3);
    s.num_tris = (uint32_t)(mesh.indices.size() / 3);
    s.is_emitter = is_emitter ? 1 : 0;
    s.radiance = radiance ? *radiance : make_rgb_texture(0.f, 0.f, 0.f);
    for (uint32_t idx : mesh.indices) indices.push_back(base + idx);
    shapes.push_back(s);
    return (uint32_t)shapes.size() - 1;
}

uint32_t HostScene::add_bitmap(std::shared_ptr<BitmapData> data, uint32_t filter, uint32_t wrap) {
    pt_bitmap_desc b;
    std::memset(&b, 0, sizeof(b));
    b.width = data->width;
    b.height = data->height;
    b.filter = filter;
    b.wrap = wrap;
    b.rgba = data->rgba.data();  // owned by the shared image; the buffer does not move
    bitmap_data.push_back(std::move(data));
    bitmaps.push_back(b);
    return (uint32_t)bitmaps.size() - 1;
}

uint32_t HostScene::add_sphere(const Mat4 &to_world, uint32_t material, uint32_t flags, bool is_emitter,
                               const pt_texture_desc *radiance) {
    pt_sphere_desc sp;
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 4; ++c) sp.to_world[4 * r + c] = to_world.e[4 * r + c];
    pt_shape_desc s;
    std::memset(&s, 0, sizeof(s));
    s.material = material;
    s.flags = (flags & PT_SHAPE_FLIP_NORMALS) | PT_SHAPE_SPHERE;
    s.first_tri = (uint32_t)spheres.size();
    s.num_tris = 1;
    s.is_emitter = is_emitter ? 1 : 0;
    s.radiance = radiance ? *radiance : make_rgb_texture(0.f, 0.f, 0.f);
    spheres.push_back(sp);
    shapes.push_back(s);
    return (uint32_t)shapes.size() - 1;
}

void HostScene::set_camera(float fov_y_deg, const Mat4 &c2w) {
    Mat4 s2c = sample_to_camera_matrix(fov_y_deg, (float)width / (float)height, 0.01f, 10000.f);
    std::memcpy(camera.sample_to_camera, s2c.e, sizeof(s2c.e));
    std::memcpy(camera.camera_to_world, c2w.e, sizeof(c2w.e));
}

pt_scene_desc HostScene::desc() const {
    pt_scene_desc d;
    std::memset(&d, 0, sizeof(d));
    d.num_vertices = (uint32_t)(positions.size() / 3);
    d.num_triangles = (uint32_t)(indices.size() / 3);
    d.num_shapes = (uint32_t)shapes.size();
    d.num_materials = (uint32_t)materials.size();
    d.positions = positions.data();
    d.normals = any_normals ? normals.data() : nullptr;
    d.texcoords = any_texcoords ? texcoords.data() : nullptr;
    d.indices = indices.data();
    d.shapes = shapes.data();
    d.materials = materials.data();
    d.env_type = env_type;
    std::memcpy(d.env_radiance, env_radiance, sizeof(env_radiance));
    d.camera = camera;
    d.width = width;
    d.height = height;
    d.max_depth = max_depth;
    d.num_spheres = (uint32_t)spheres.size();
    d.spheres = spheres.empty() ? nullptr : spheres.data();
    d.num_bitmaps = (uint32_t)bitmaps.size();
    d.bitmaps = bitmaps.empty() ? nullptr : bitmaps.data();
    d.env_bitmap = env_bitmap;
    d.env_scale = env_scale;
    std::memcpy(d.env_to_world, env_to_world, sizeof(env_to_world));
    return d;
}

// ---------------------------------------------------------------------------
// Procedural scenes
namespace {

struct SplitMix {
    uint64_t s;
    uint64_t next() {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
    float range(float a, float b) { return (float)(a + (b - a) * uniform()); }
};

uint64_t mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// Value noise on an integer lattice with smoothstep interpolation.
double lattice(int64_t ix, int64_t iz, uint64_t salt) {
    uint64_t h = mix64((uint64_t)ix * 0x9E3779B97F4A7C15ull ^ mix64((uint64_t)iz + salt));
    return (double)(h >> 11) * (2.0 / 9007199254740992.0) - 1.0;
}
double value_noise(double x, double z, uint64_t salt) {
    double fx = std::floor(x), fz = std::floor(z);
    int64_t ix = (int64_t)fx, iz = (int64_t)fz;
    double tx = x - fx, tz = z - fz;
    double sx = tx * tx * (3 - 2 * tx), sz = tz * tz * (3 - 2 * tz);
    double a = lattice(ix, iz, salt), b = lattice(ix + 1, iz, salt);
    double c = lattice(ix, iz + 1, salt), d = lattice(ix + 1, iz + 1, salt);
    return (a + (b - a) * sx) + ((c + (d - c) * sx) - (a + (b - a) * sx)) * sz;
}
double fbm_height(double x, double z, uint64_t seed) {
    double sum = 0.0, amp = 1.0, freq = 2.0, norm = 0.0;
    for (int o = 0; o < 6; ++o) {
        sum += amp * value_noise(x * freq, z * freq, seed * 1315423911ull + (uint64_t)o * 2654435761ull);
        norm += amp;
        amp *= 0.5;
        freq *= 2.0;
    }
    return 0.2 * sum / norm;
}

// Conductor eta/k of gold (the Au entry of the conductor table).
const float kAuEta[3] = {0.14282f, 0.37414f, 1.43944f};
const float kAuK[3] = {3.97472f, 2.38066f, 1.59981f};

// Opaque types are two-sided: the icospheres sink into the heightfield, so
// paths reach their inner faces, where a one-sided conductor would evaluate
// its Fresnel term at cos < 0 (reflectance above 1; conductor.h:27-34 with
// fresnel.h:31-49 of the reference) and plastics would go black.
pt_material_desc random_material(SplitMix &rng, int type) {
    pt_material_desc m;
    std::memset(&m, 0, sizeof(m));
    m.type = (uint32_t)type;
    m.twosided = (type == PT_MAT_DIELECTRIC || type == PT_MAT_ROUGH_DIELECTRIC) ? 0u : 1u;
    m.int_ior = 1.5f;
    m.ext_ior = 1.0f;
    for (int i = 0; i < 4; ++i) m.tex[i] = make_rgb_texture(0.f, 0.f, 0.f);
    float alpha = rng.range(0.05f, 0.5f);
    float r = rng.range(0.2f, 0.9f), g = rng.range(0.2f, 0.9f), b = rng.range(0.2f, 0.9f);
    switch (type) {
        case PT_MAT_DIFFUSE:
            m.tex[0] = make_rgb_texture(r, g, b);
            break;
        case PT_MAT_DIELECTRIC:
            m.tex[0] = make_rgb_texture(1.f, 1.f, 1.f);
            m.tex[1] = make_rgb_texture(1.f, 1.f, 1.f);
            break;
        case PT_MAT_ROUGH_DIELECTRIC:
            m.tex[0] = make_rgb_texture(1.f, 1.f, 1.f);
            m.tex[1] = make_rgb_texture(1.f, 1.f, 1.f);
            m.tex[2] = make_rgb_texture(alpha, alpha, alpha);
            break;
        case PT_MAT_CONDUCTOR:
            m.tex[0] = make_rgb_texture(1.f, 1.f, 1.f);
            m.tex[1] = make_rgb_texture(kAuEta[0], kAuEta[1], kAuEta[2]);
            m.tex[2] = make_rgb_texture(kAuK[0], kAuK[1], kAuK[2]);
            break;
        case PT_MAT_ROUGH_CONDUCTOR:
            m.tex[0] = make_rgb_texture(1.f, 1.f, 1.f);
            m.tex[1] = make_rgb_texture(kAuEta[0], kAuEta[1], kAuEta[2]);
            m.tex[2] = make_rgb_texture(kAuK[0], kAuK[1], kAuK[2]);
            m.tex[3] = make_rgb_texture(alpha, alpha, alpha);
            break;
        case PT_MAT_PLASTIC:
            m.tex[0] = make_rgb_texture(r, g, b);
            m.tex[1] = make_rgb_texture(1.f, 1.f, 1.f);
            break;
        default:  // rough plastic
            m.tex[0] = make_rgb_texture(r, g, b);
            m.tex[1] = make_rgb_texture(1.f, 1.f, 1.f);
            m.tex[2] = make_rgb_texture(alpha, alpha, alpha);
            break;
    }
    return m;
}

Mat4 trs(float sx, float sy, float sz, float ax, float ay, float az, float angle, float tx, float ty, float tz) {
    Transform t;
    t.scale(sx, sy, sz);
    if (angle != 0.f) t.rotate(ax, ay, az, angle);
    t.translate(tx, ty, tz);
    return t.m;
}

// Cornell-box style scene: five walls, two rotated boxes, one area light.
void build_cornell(HostScene &s, const pt_gen_params &p) {
    s.width = p.width ? p.width : 512;
    s.height = p.height ? p.height : 512;
    s.max_depth = p.max_depth ? p.max_depth : 8;
    uint32_t white = s.add_material(make_diffuse(0.73f, 0.73f, 0.73f, true));
    uint32_t red = s.add_material(make_diffuse(0.65f, 0.05f, 0.05f, true));
    uint32_t green = s.add_material(make_diffuse(0.12f, 0.45f, 0.15f, true));
    uint32_t black = s.add_material(make_diffuse(0.f, 0.f, 0.f, true));
    MeshData rect = make_rectangle_mesh(), cube = make_cube_mesh();
    // Box interior: x in [-1, 1], y in [0, 2], z in [-1, 1]; open towards +z.
    s.add_instance(rect, trs(1, 1, 1, 1, 0, 0, -90.f, 0, 0, 0), white, 0, false, nullptr);     // floor
    s.add_instance(rect, trs(1, 1, 1, 1, 0, 0, 90.f, 0, 2, 0), white, 0, false, nullptr);      // ceiling
    s.add_instance(rect, trs(1, 1, 1, 0, 0, 0, 0.f, 0, 1, -1), white, 0, false, nullptr);      // back
    s.add_instance(rect, trs(1, 1, 1, 0, 1, 0, 90.f, -1, 1, 0), red, 0, false, nullptr);       // left
    s.add_instance(rect, trs(1, 1, 1, 0, 1, 0, -90.f, 1, 1, 0), green, 0, false, nullptr);     // right
    s.add_instance(cube, trs(0.3f, 0.3f, 0.3f, 0, 1, 0, -17.f, 0.33f, 0.3f, 0.37f), white, 0, false, nullptr);
    s.add_instance(cube, trs(0.3f, 0.6f, 0.3f, 0, 1, 0, 17.f, -0.335f, 0.6f, -0.29f), white, 0, false, nullptr);
    pt_texture_desc le = make_rgb_texture(17.f, 12.f, 4.f);
    s.add_instance(rect, trs(0.235f, 0.19f, 1, 1, 0, 0, 90.f, -0.005f, 1.98f, -0.03f), black, 0, true, &le);
    float eye[3] = {0.f, 1.f, 3.9f}, target[3] = {0.f, 1.f, 0.f}, up[3] = {0.f, 1.f, 0.f};
    Transform cam;
    cam.look_at(eye, target, up);
    s.set_camera(40.f, cam.m);
}

void add_grid_patch(HostScene &s, uint32_t material, int i0, int i1, int j0, int j1, int grid, uint64_t seed) {
    MeshData m;
    const int nx = i1 - i0 + 1, nz = j1 - j0 + 1;
    const double step = 2.0 / (grid - 1);
    for (int j = j0; j <= j1; ++j) {
        for (int i = i0; i <= i1; ++i) {
            double x = -1.0 + i * step, z = -1.0 + j * step;
            double y = fbm_height(x, z, seed);
            double e = 0.5 * step;
            double dx = (fbm_height(x + e, z, seed) - fbm_height(x - e, z, seed)) / (2 * e);
            double dz = (fbm_height(x, z + e, seed) - fbm_height(x, z - e, seed)) / (2 * e);
            m.positions.insert(m.positions.end(), {(float)x, (float)y, (float)z});
            double nl = std::sqrt(dx * dx + 1.0 + dz * dz);
            m.normals.insert(m.normals.end(), {(float)(-dx / nl), (float)(1.0 / nl), (float)(-dz / nl)});
            m.texcoords.insert(m.texcoords.end(), {(float)(0.5 * (x + 1.0)), (float)(0.5 * (z + 1.0))});
        }
    }
    for (int j = 0; j < nz - 1; ++j)
        for (int i = 0; i < nx - 1; ++i) {
            uint32_t a = (uint32_t)(j * nx + i), b = a + 1, c = a + (uint32_t)nx, d = c + 1;
            m.indices.insert(m.indices.end(), {a, c, b, b, c, d});
        }
    s.add_instance(m, Mat4::identity(), material, 0, false, nullptr);
}

// Fractal heightfield, icospheres with materials of all seven types, four
// area-light quads and a constant environment.
void build_terrain(HostScene &s, const pt_gen_params &p) {
    SplitMix rng{p.seed * 0x2545F4914F6CDD1Dull + 17};
    s.width = p.width ? p.width : 1920;
    s.height = p.height ? p.height : 1080;
    s.max_depth = p.max_depth ? p.max_depth : 16;
    const int grid = p.grid ? (int)p.grid : 1024;
    const int nsph = p.spheres ? (int)p.spheres : 256;
    const int cells = grid - 1;
    const int bounds[5] = {0, cells / 4, cells / 2, (3 * cells) / 4, cells};
    for (int bj = 0; bj < 4; ++bj)
        for (int bi = 0; bi < 4; ++bi) {
            int type = 1 + (int)(rng.next() % 7);
            uint32_t mat = s.add_material(random_material(rng, type));
            add_grid_patch(s, mat, bounds[bi], bounds[bi + 1], bounds[bj], bounds[bj + 1], grid, p.seed);
        }
    MeshData ico = make_icosphere_mesh(3);
    for (int k = 0; k < nsph; ++k) {
        float r = rng.range(0.02f, 0.1f);
        float x = rng.range(-0.9f, 0.9f), z = rng.range(-0.9f, 0.9f);
        float y = (float)fbm_height(x, z, p.seed) + 0.8f * r;
        int type = 1 + (int)(rng.next() % 7);
        uint32_t mat = s.add_material(random_material(rng, type));
        s.add_instance(ico, trs(r, r, r, 0, 0, 0, 0.f, x, y, z), mat, 0, false, nullptr);
    }
    uint32_t black = s.add_material(make_diffuse(0.f, 0.f, 0.f, false));
    MeshData rect = make_rectangle_mesh();
    for (int l = 0; l < 4; ++l) {
        float le = rng.range(5.f, 20.f);
        pt_texture_desc rad = make_rgb_texture(le, le, le);
        float x = rng.range(-0.7f, 0.7f), z = rng.range(-0.7f, 0.7f), y = rng.range(1.2f, 1.6f);
        s.add_instance(rect, trs(0.125f, 0.125f, 1, 1, 0, 0, 90.f, x, y, z), black, 0, true, &rad);
    }
    s.env_type = PT_ENV_CONSTANT;
    s.env_radiance[0] = s.env_radiance[1] = s.env_radiance[2] = 0.1f;
    float eye[3] = {0.f, 0.9f, 2.0f}, target[3] = {0.f, 0.f, 0.f}, up[3] = {0.f, 1.f, 0.f};
    Transform cam;
    cam.look_at(eye, target, up);
    s.set_camera(45.f, cam.m);
}

// Axis-aligned wall split into a grid of quads.  `axis` is the constant
// coordinate; (u0,u1) x (v0,v1) span the other two in cyclic order.
void add_wall(HostScene &s, uint32_t material, int axis, float c, float u0, float u1, float v0, float v1, float cell,
              bool is_emitter = false, const pt_texture_desc *rad = nullptr) {
    int nu = std::max(1, (int)std::ceil((u1 - u0) / cell)), nv = std::max(1, (int)std::ceil((v1 - v0) / cell));
    MeshData m;
    int ua = (axis + 1) % 3, va = (axis + 2) % 3;
    for (int j = 0; j <= nv; ++j)
        for (int i = 0; i <= nu; ++i) {
            float p[3];
            p[axis] = c;
            p[ua] = u0 + (u1 - u0) * (float)i / (float)nu;
            p[va] = v0 + (v1 - v0) * (float)j / (float)nv;
            m.positions.insert(m.positions.end(), p, p + 3);
            float n[3] = {0.f, 0.f, 0.f};
            n[axis] = 1.f;
            m.normals.insert(m.normals.end(), n, n + 3);
        }
    for (int j = 0; j < nv; ++j)
        for (int i = 0; i < nu; ++i) {
            uint32_t a = (uint32_t)(j * (nu + 1) + i), b = a + 1, cc = a + (uint32_t)(nu + 1), d = cc + 1;
            m.indices.insert(m.indices.end(), {a, b, d, a, d, cc});
        }
    s.add_instance(m, Mat4::identity(), material, 0, is_emitter, rad);
}

// 4 x 3 x 6 m room whose only light sits in a closet behind a 5 cm door gap.
// Seed 0 is the canonical layout (door centered at z = 3); other seeds move
// the door and its closet along the wall (center z in [1.5, 4.5]) and place
// the camera at random inside the room, looking towards the far half.
void build_door_room(HostScene &s, const pt_gen_params &p) {
    s.width = p.width ? p.width : 1920;
    s.height = p.height ? p.height : 1080;
    s.max_depth = p.max_depth ? p.max_depth : 16;
    const float cell = 0.0215f;
    float zc = 3.f;
    float eye[3] = {1.0f, 1.6f, 0.5f}, target[3] = {3.0f, 1.2f, 4.5f}, up[3] = {0.f, 1.f, 0.f};
    if (p.seed != 0) {
        SplitMix rng{p.seed * 0x9E3779B97F4A7C15ull + 0x2545F4914F6CDD1Dull};
        zc = rng.range(1.5f, 4.5f);
        eye[0] = rng.range(0.4f, 1.6f);
        eye[1] = rng.range(1.2f, 1.9f);
        eye[2] = rng.range(0.3f, 1.2f);
        target[0] = rng.range(2.4f, 3.6f);
        target[1] = rng.range(0.8f, 1.6f);
        target[2] = rng.range(3.5f, 5.5f);
    }
    const float d0 = zc - 0.025f, d1 = zc + 0.025f, c0 = zc - 0.5f, c1 = zc + 0.5f;
    uint32_t wall = s.add_material(make_diffuse(0.6f, 0.6f, 0.6f, true));
    uint32_t black = s.add_material(make_diffuse(0.f, 0.f, 0.f, true));
    // Room x in [0,4], y in [0,3], z in [0,6].
    add_wall(s, wall, 1, 0.f, 0.f, 6.f, 0.f, 4.f, cell);   // floor (axis y: u = z, v = x)
    add_wall(s, wall, 1, 3.f, 0.f, 6.f, 0.f, 4.f, cell);   // ceiling
    add_wall(s, wall, 0, 0.f, 0.f, 3.f, 0.f, 6.f, cell);   // x = 0 (u = y, v = z)
    add_wall(s, wall, 2, 0.f, 0.f, 4.f, 0.f, 3.f, cell);   // z = 0 (u = x, v = y)
    add_wall(s, wall, 2, 6.f, 0.f, 4.f, 0.f, 3.f, cell);   // z = 6
    // x = 4 wall with a door slit z in [zc - 0.025, zc + 0.025], y in [0, 2].
    add_wall(s, wall, 0, 4.f, 0.f, 3.f, 0.f, d0, cell);
    add_wall(s, wall, 0, 4.f, 0.f, 3.f, d1, 6.f, cell);
    add_wall(s, wall, 0, 4.f, 2.f, 3.f, d0, d1, cell);
    // Closet x in [4,5], z in [zc - 0.5, zc + 0.5].
    add_wall(s, wall, 1, 0.f, c0, c1, 4.f, 5.f, cell);
    add_wall(s, wall, 1, 3.f, c0, c1, 4.f, 5.f, cell);
    add_wall(s, wall, 0, 5.f, 0.f, 3.f, c0, c1, cell);
    add_wall(s, wall, 2, c0, 4.f, 5.f, 0.f, 3.f, cell);
    add_wall(s, wall, 2, c1, 4.f, 5.f, 0.f, 3.f, cell);
    pt_texture_desc le = make_rgb_texture(200.f, 200.f, 200.f);
    add_wall(s, black, 0, 4.99f, 1.0f, 2.0f, zc - 0.25f, zc + 0.25f, 0.25f, true, &le);
    Transform cam;
    cam.look_at(eye, target, up);
    s.set_camera(60.f, cam.m);
}

// Rectangle [-1,1]^2 (z = 0, normal +z) split into n x n cells.
MeshData make_grid_rectangle_mesh(int n) {
    MeshData m;
    for (int j = 0; j <= n; ++j)
        for (int i = 0; i <= n; ++i) {
            const float u = (float)i / (float)n, v = (float)j / (float)n;
            m.positions.insert(m.positions.end(), {2.f * u - 1.f, 2.f * v - 1.f, 0.f});
            m.normals.insert(m.normals.end(), {0.f, 0.f, 1.f});
            m.texcoords.insert(m.texcoords.end(), {u, v});
        }
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i) {
            uint32_t a = (uint32_t)(j * (n + 1) + i), b = a + 1, c = a + (uint32_t)(n + 1), d = c + 1;
            m.indices.insert(m.indices.end(), {a, b, d, a, d, c});
        }
    return m;
}

// Cube [-1,1]^3 with every face split into n x n cells (outward normals; the
// face layout and winding of make_cube_mesh).
MeshData make_grid_cube_mesh(int n) {
    MeshData m;
    const int axes[6] = {0, 0, 1, 1, 2, 2};
    const float signs[6] = {-1.f, 1.f, -1.f, 1.f, -1.f, 1.f};
    for (int f = 0; f < 6; ++f) {
        const int a = axes[f];
        const float sgn = signs[f];
        const int u = (a + 1) % 3, v = (a + 2) % 3;
        const uint32_t base = (uint32_t)(m.positions.size() / 3);
        for (int j = 0; j <= n; ++j)
            for (int i = 0; i <= n; ++i) {
                float p[3], nn[3] = {0.f, 0.f, 0.f};
                p[a] = sgn;
                p[u] = (2.f * (float)i / (float)n - 1.f) * (sgn > 0.f ? 1.f : -1.f);
                p[v] = 2.f * (float)j / (float)n - 1.f;
                nn[a] = sgn;
                m.positions.insert(m.positions.end(), p, p + 3);
                m.normals.insert(m.normals.end(), nn, nn + 3);
                m.texcoords.insert(m.texcoords.end(), {(float)i / (float)n, (float)j / (float)n});
            }
        for (int j = 0; j < n; ++j)
            for (int i = 0; i < n; ++i) {
                const uint32_t a0 = base + (uint32_t)(j * (n + 1) + i), b0 = a0 + 1, c0 = a0 + (uint32_t)(n + 1),
                               d0 = c0 + 1;
                m.indices.insert(m.indices.end(), {a0, b0, d0, a0, d0, c0});
            }
    }
    return m;
}

// Seeded variant of the Cornell box: the walls and both boxes tessellated so
// the scene has about `triangles` triangles (log-uniform in [50K, 2M] when 0),
// a random camera in front of / inside the open box, and a film size drawn from
// {256^2, 512^2, 768x432, 1024^2, 1920x1080}.  Max depth 8.
void build_cornell_variant(HostScene &s, const pt_gen_params &p) {
    SplitMix rng{p.seed * 0x9E3779B97F4A7C15ull + 0x5851F42D4C957F2Dull};
    static const uint32_t sizes[5][2] = {{256, 256}, {512, 512}, {768, 432}, {1024, 1024}, {1920, 1080}};
    const uint32_t pick = (uint32_t)(rng.next() % 5);
    s.width = p.width ? p.width : sizes[pick][0];
    s.height = p.height ? p.height : sizes[pick][1];
    s.max_depth = p.max_depth ? p.max_depth : 8;
    double target = p.triangles;
    if (target == 0.0) target = std::exp(std::log(5.0e4) + (std::log(2.0e6) - std::log(5.0e4)) * rng.uniform());
    // 5 walls + 12 box faces of 2 n^2 triangles each, plus the 2-triangle light.
    const int n = std::max(1, (int)std::lround(std::sqrt(std::max(0.0, target - 2.0) / 34.0)));
    uint32_t white = s.add_material(make_diffuse(0.73f, 0.73f, 0.73f, true));
    uint32_t red = s.add_material(make_diffuse(0.65f, 0.05f, 0.05f, true));
    uint32_t green = s.add_material(make_diffuse(0.12f, 0.45f, 0.15f, true));
    uint32_t black = s.add_material(make_diffuse(0.f, 0.f, 0.f, true));
    MeshData rect = make_grid_rectangle_mesh(n), cube = make_grid_cube_mesh(n), light = make_rectangle_mesh();
    s.add_instance(rect, trs(1, 1, 1, 1, 0, 0, -90.f, 0, 0, 0), white, 0, false, nullptr);
    s.add_instance(rect, trs(1, 1, 1, 1, 0, 0, 90.f, 0, 2, 0), white, 0, false, nullptr);
    s.add_instance(rect, trs(1, 1, 1, 0, 0, 0, 0.f, 0, 1, -1), white, 0, false, nullptr);
    s.add_instance(rect, trs(1, 1, 1, 0, 1, 0, 90.f, -1, 1, 0), red, 0, false, nullptr);
    s.add_instance(rect, trs(1, 1, 1, 0, 1, 0, -90.f, 1, 1, 0), green, 0, false, nullptr);
    s.add_instance(cube, trs(0.3f, 0.3f, 0.3f, 0, 1, 0, -17.f, 0.33f, 0.3f, 0.37f), white, 0, false, nullptr);
    s.add_instance(cube, trs(0.3f, 0.6f, 0.3f, 0, 1, 0, 17.f, -0.335f, 0.6f, -0.29f), white, 0, false, nullptr);
    pt_texture_desc le = make_rgb_texture(17.f, 12.f, 4.f);
    s.add_instance(light, trs(0.235f, 0.19f, 1, 1, 0, 0, 90.f, -0.005f, 1.98f, -0.03f), black, 0, true, &le);
    float eye[3] = {rng.range(-0.6f, 0.6f), rng.range(0.4f, 1.6f), rng.range(1.2f, 3.9f)};
    float tgt[3] = {rng.range(-0.4f, 0.4f), rng.range(0.4f, 1.4f), rng.range(-0.9f, 0.f)};
    float up[3] = {0.f, 1.f, 0.f};
    Transform cam;
    cam.look_at(eye, tgt, up);
    s.set_camera(rng.range(30.f, 60.f), cam.m);
}

// Closed cube [-1,1]^3 of twosided diffuse walls that all emit, seen from the
// center: every path stays inside, so the expected pixel value is
// Le * sum_{k < max_depth} albedo^k.
void build_furnace(HostScene &s, const pt_gen_params &p) {
    s.width = p.width ? p.width : 32;
    s.height = p.height ? p.height : 32;
    s.max_depth = p.max_depth ? p.max_depth : 6;
    const float albedo = p.albedo > 0.f ? p.albedo : 0.5f;
    const float le = p.radiance > 0.f ? p.radiance : 1.f;
    uint32_t wall = s.add_material(make_diffuse(albedo, albedo, albedo, true));
    pt_texture_desc rad = make_rgb_texture(le, le, le);
    s.add_instance(make_cube_mesh(), Mat4::identity(), wall, 0, true, &rad);
    float eye[3] = {0.1f, -0.05f, 0.2f}, target[3] = {0.3f, 0.1f, -1.f}, up[3] = {0.f, 1.f, 0.f};
    Transform cam;
    cam.look_at(eye, target, up);
    s.set_camera(70.f, cam.m);
}

}  // namespace

bool generate_scene(int kind, const pt_gen_params &params, HostScene &out, std::string &err) {
    out = HostScene();
    switch (kind) {
        case PT_SCENE_CORNELL: build_cornell(out, params); return true;
        case PT_SCENE_TERRAIN: build_terrain(out, params); return true;
        case PT_SCENE_DOOR_ROOM: build_door_room(out, params); return true;
        case PT_SCENE_FURNACE: build_furnace(out, params); return true;
        case PT_SCENE_CORNELL_VARIANT: build_cornell_variant(out, params); return true;
        default: err = "unknown scene kind"; return false;
    }
}

// ---------------------------------------------------------------------------
// IOR tables
namespace {
struct DielectricIor { const char *name; float value; };
const DielectricIor kDielectricIor[] = {
    {"vacuum", 1.0f}, {"helium", 1.000036f}, {"hydrogen", 1.000132f}, {"air", 1.000277f},
    {"carbon dioxide", 1.00045f}, {"water", 1.3330f}, {"acetone", 1.36f}, {"ethanol", 1.361f},
    {"carbon tetrachloride", 1.461f}, {"glycerol", 1.4729f}, {"benzene", 1.501f}, {"silicone oil", 1.52045f},
    {"bromine", 1.661f}, {"water ice", 1.31f}, {"fused quartz", 1.458f}, {"pyrex", 1.470f},
    {"acrylic glass", 1.49f}, {"polypropylene", 1.49f}, {"bk7", 1.5046f}, {"sodium chloride", 1.544f},
    {"amber", 1.55f}, {"pet", 1.5750f}, {"diamond", 2.419f},
};
struct ConductorIor { const char *name; float eta[3]; float k[3]; };
// Conductor eta / k at the RGB primaries, 73 named metals and compounds
// (render/material/ior.h:59-174; physical data, so the values are the same).
const ConductorIor kConductorIor[] = {
    {"a-C", {2.93785f, 2.22242f, 1.96400f}, {0.88555f, 0.79763f, 0.81356f}},
    {"Ag", {0.15494f, 0.11648f, 0.13809f}, {4.81810f, 3.11562f, 2.14240f}},
    {"Al", {1.65394f, 0.87850f, 0.52012f}, {9.20430f, 6.25621f, 4.82675f}},
    {"Al2O3", {2.12296f, 1.67556f, 1.61569f}, {2.11325f, 1.66785f, 1.60811f}},
    {"AlAs", {3.59737f, 3.22615f, 2.21235f}, {0.00067f, -0.00050f, 0.00741f}},
    {"AlAs_palik", {3.59737f, 3.22615f, 2.21235f}, {0.00067f, -0.00050f, 0.00741f}},
    {"AlSb", {4.70830f, 4.05196f, 4.57419f}, {-0.02918f, 0.09345f, 1.29784f}},
    {"AlSb_palik", {4.70830f, 4.05196f, 4.57419f}, {-0.02918f, 0.09345f, 1.29784f}},
    {"Au", {0.14282f, 0.37414f, 1.43944f}, {3.97472f, 2.38066f, 1.59981f}},
    {"Be", {4.17618f, 3.17830f, 2.77819f}, {3.82730f, 3.00374f, 2.86293f}},
    {"Be_palik", {4.17618f, 3.17830f, 2.77819f}, {3.82730f, 3.00374f, 2.86293f}},
    {"Cr", {4.36041f, 2.91052f, 1.65119f}, {5.19538f, 4.22239f, 3.74700f}},
    {"CsI", {2.14035f, 1.69870f, 1.65890f}, {0.00000f, 0.00000f, 0.00000f}},
    {"CsI_palik", {2.14035f, 1.69870f, 1.65890f}, {0.00000f, 0.00000f, 0.00000f}},
    {"Cu", {0.19999f, 0.92209f, 1.09988f}, {3.90464f, 2.44763f, 2.13765f}},
    {"Cu_palik", {0.21221f, 0.91805f, 1.10007f}, {3.91324f, 2.45193f, 2.13213f}},
    {"Cu2O", {3.54426f, 2.94365f, 2.74619f}, {0.11395f, 0.19341f, 0.60475f}},
    {"Cu2O_palik", {3.54426f, 2.94365f, 2.74619f}, {0.11395f, 0.19341f, 0.60475f}},
    {"CuO", {3.24062f, 2.44261f, 2.20556f}, {0.51999f, 0.56882f, 0.72064f}},
    {"CuO_palik", {3.24062f, 2.44261f, 2.20556f}, {0.51999f, 0.56882f, 0.72064f}},
    {"d-C", {2.90008f, 2.29322f, 2.22275f}, {0.00000f, 0.00000f, 0.00000f}},
    {"d-C_palik", {2.90008f, 2.29322f, 2.22275f}, {0.00000f, 0.00000f, 0.00000f}},
    {"Fe", {2.76404f, 1.95417f, 1.62766f}, {3.83077f, 2.73841f, 2.31812f}},
    {"GaN", {2.86583f, 2.29184f, 2.29848f}, {0.00020f, -0.00017f, 0.00128f}},
    {"Ge", {7.05832f, 4.78868f, 3.57814f}, {0.50903f, 2.19196f, 2.05716f}},
    {"Hg", {2.39384f, 1.43697f, 0.90762f}, {6.31420f, 4.36266f, 3.41454f}},
    {"Hg_palik", {2.39384f, 1.43697f, 0.90762f}, {6.31420f, 4.36266f, 3.41454f}},
    {"HgTe", {4.76940f, 3.22413f, 2.65439f}, {1.62853f, 1.57746f, 1.72592f}},
    {"HgTe_palik", {4.76940f, 3.22413f, 2.65439f}, {1.62853f, 1.57746f, 1.72592f}},
    {"Ir", {3.07986f, 2.07777f, 1.61446f}, {5.58028f, 4.05855f, 3.26033f}},
    {"Ir_palik", {3.07986f, 2.07777f, 1.61446f}, {5.58028f, 4.05855f, 3.26033f}},
    {"K", {0.06391f, 0.04631f, 0.03810f}, {2.09975f, 1.34607f, 0.91128f}},
    {"K_palik", {0.06391f, 0.04631f, 0.03810f}, {2.09975f, 1.34607f, 0.91128f}},
    {"Li", {0.26525f, 0.19519f, 0.22045f}, {3.53305f, 2.30618f, 1.66505f}},
    {"Li_palik", {0.26525f, 0.19519f, 0.22045f}, {3.53305f, 2.30618f, 1.66505f}},
    {"MgO", {2.08521f, 1.64721f, 1.59150f}, {0.00000f, -0.00000f, 0.00000f}},
    {"MgO_palik", {2.08521f, 1.64721f, 1.59150f}, {0.00000f, -0.00000f, 0.00000f}},
    {"Mo", {4.47417f, 3.51799f, 2.77018f}, {4.10240f, 3.41361f, 3.14393f}},
    {"Mo_palik", {4.47417f, 3.51799f, 2.77018f}, {4.10240f, 3.41361f, 3.14393f}},
    {"Na", {0.06014f, 0.05602f, 0.06186f}, {3.17254f, 2.10800f, 1.57575f}},
    {"Na_palik", {0.06014f, 0.05602f, 0.06186f}, {3.17254f, 2.10800f, 1.57575f}},
    {"Nb", {3.41288f, 2.78427f, 2.39051f}, {3.43408f, 2.73183f, 2.57445f}},
    {"Nb_palik", {3.41288f, 2.78427f, 2.39051f}, {3.43408f, 2.73183f, 2.57445f}},
    {"Ni", {2.36225f, 1.65983f, 1.46395f}, {4.48929f, 3.04369f, 2.34046f}},
    {"Ni_palik", {2.36225f, 1.65983f, 1.46395f}, {4.48929f, 3.04369f, 2.34046f}},
    {"Rh", {2.58031f, 1.85624f, 1.55114f}, {6.76790f, 4.69297f, 3.96766f}},
    {"Rh_palik", {2.58031f, 1.85624f, 1.55114f}, {6.76790f, 4.69297f, 3.96766f}},
    {"Se", {4.08037f, 2.83343f, 2.81458f}, {0.75378f, 0.63705f, 0.52047f}},
    {"Se_palik", {4.08037f, 2.83343f, 2.81458f}, {0.75378f, 0.63705f, 0.52047f}},
    {"SiC", {3.16562f, 2.52061f, 2.47411f}, {0.00000f, -0.00000f, 0.00001f}},
    {"SiC_palik", {3.16562f, 2.52061f, 2.47411f}, {0.00000f, -0.00000f, 0.00001f}},
    {"SnTe", {4.51562f, 1.97692f, 1.27897f}, {0.00000f, 0.00000f, 0.00000f}},
    {"SnTe_palik", {4.51562f, 1.97692f, 1.27897f}, {0.00000f, 0.00000f, 0.00000f}},
    {"Ta", {2.05820f, 2.38802f, 2.62250f}, {2.40293f, 1.73767f, 1.94291f}},
    {"Ta_palik", {2.05820f, 2.38802f, 2.62250f}, {2.40293f, 1.73767f, 1.94291f}},
    {"Te", {7.37519f, 4.47257f, 2.63149f}, {3.24919f, 3.51992f, 3.28521f}},
    {"Te_palik", {7.37519f, 4.47257f, 2.63149f}, {3.24919f, 3.51992f, 3.28521f}},
    {"ThF4", {1.82684f, 1.43917f, 1.38471f}, {0.00000f, 0.00000f, 0.00000f}},
    {"ThF4_palik", {1.82684f, 1.43917f, 1.38471f}, {0.00000f, 0.00000f, 0.00000f}},
    {"TiC", {3.69261f, 2.83141f, 2.57683f}, {3.25876f, 2.34656f, 2.16818f}},
    {"TiC_palik", {3.69261f, 2.83141f, 2.57683f}, {3.25876f, 2.34656f, 2.16818f}},
    {"TiN", {1.64497f, 1.14800f, 1.37685f}, {3.36132f, 1.93936f, 1.09967f}},
    {"TiN_palik", {1.64497f, 1.14800f, 1.37685f}, {3.36132f, 1.93936f, 1.09967f}},
    {"TiO2", {3.44929f, 2.79576f, 2.89899f}, {0.00010f, -0.00009f, 0.00063f}},
    {"TiO2_palik", {3.44929f, 2.79576f, 2.89899f}, {0.00010f, -0.00009f, 0.00063f}},
    {"VC", {3.64981f, 2.74689f, 2.52731f}, {3.06184f, 2.19400f, 1.95902f}},
    {"VC_palik", {3.64981f, 2.74689f, 2.52731f}, {3.06184f, 2.19400f, 1.95902f}},
    {"V", {4.26843f, 3.50571f, 2.75528f}, {3.48377f, 2.88322f, 3.10511f}},
    {"V_palik", {4.26843f, 3.50571f, 2.75528f}, {3.48377f, 2.88322f, 3.10511f}},
    {"VN", {2.85952f, 2.11468f, 1.93597f}, {3.02590f, 2.05174f, 1.61287f}},
    {"VN_palik", {2.85952f, 2.11468f, 1.93597f}, {3.02590f, 2.05174f, 1.61287f}},
    {"W", {4.36142f, 3.29330f, 2.99191f}, {3.49325f, 2.59934f, 2.26838f}},
    {"none", {0.0f, 0.0f, 0.0f}, {1.0f, 1.0f, 1.0f}},
};
}  // namespace

bool lookup_dielectric_ior(const std::string &name, float &value) {
    for (const auto &e : kDielectricIor)
        if (name == e.name) {
            value = e.value;
            return true;
        }
    return false;
}
bool lookup_conductor_ior(const std::string &name, float eta[3], float k[3]) {
    for (const auto &e : kConductorIor)
        if (name == e.name) {
            for (int i = 0; i < 3; ++i) {
                eta[i] = e.eta[i];
                k[i] = e.k[i];
            }
            return true;
        }
    return false;
}

}  // namespace pupil
