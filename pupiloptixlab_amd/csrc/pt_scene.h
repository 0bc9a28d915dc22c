// Host scene model: shape instances flattened to world space, materials,
// environment and camera, plus the procedural scene generators and the
// mitsuba-subset XML loader.  Mirrors resource::Scene + world::World of the
// reference (resource/scene.cpp, resource/shape.cpp, world/world.cpp) for a
// headless build.
#pragma once

#include <stdint.h>

#include <memory>
#include <string>
#include <vector>

#include "../../include/pupil_hip.h"

namespace pupil {

// Row-major 4x4 applied to column vectors (util/type.h Mat4 semantics).
struct Mat4 {
    float e[16];
    static Mat4 identity();
    Mat4 operator*(const Mat4 &b) const;
    Mat4 inverse() const;    // computed in double, rounded to float
    Mat4 transpose() const;
};

// util/transform.cpp:7-123 builders (pre-multiplying, so calls compose
// scale -> rotate -> translate as in the XML loader).
struct Transform {
    Mat4 m = Mat4::identity();
    void scale(float x, float y, float z);
    void rotate(float ux, float uy, float uz, float angle_deg);
    void translate(float x, float y, float z);
    void look_at(const float origin[3], const float target[3], const float up[3]);
};
void transform_point(const Mat4 &m, const float p[3], float out[3]);
void transform_normal_unnormalized(const Mat4 &inv_t, const float n[3], float out[3]);

// Camera matrices of util/camera.cpp:7-20: inverse of the right-handed
// perspective projection composed with the [-1,1] -> [0,1] film mapping.
Mat4 sample_to_camera_matrix(float fov_y_deg, float aspect, float near_clip, float far_clip);

struct MeshData {
    std::vector<float> positions;  // 3 per vertex, object space
    std::vector<float> normals;    // 3 per vertex or empty
    std::vector<float> texcoords;  // 2 per vertex or empty
    std::vector<uint32_t> indices; // 3 per triangle
};

pt_texture_desc make_rgb_texture(float r, float g, float b);
pt_texture_desc make_bitmap_texture(uint32_t bitmap);

// Decoded image shared by the bitmaps that sample it in different modes.
struct BitmapData {
    uint32_t width = 0, height = 0;
    std::vector<float> rgba;
};
pt_texture_desc make_checker_texture(const float c0[3], const float c1[3]);
pt_material_desc make_diffuse(float r, float g, float b, bool twosided);

class HostScene {
public:
    std::vector<float> positions, normals, texcoords;
    std::vector<uint32_t> indices;
    std::vector<pt_shape_desc> shapes;
    std::vector<pt_material_desc> materials;
    std::vector<pt_sphere_desc> spheres;
    std::vector<std::shared_ptr<BitmapData>> bitmap_data;  // per bitmap
    std::vector<pt_bitmap_desc> bitmaps;                   // rgba points into bitmap_data
    uint32_t env_type = PT_ENV_NONE;
    float env_radiance[3] = {0.f, 0.f, 0.f};
    uint32_t env_bitmap = 0;
    float env_scale = 1.f;
    float env_to_world[9] = {1.f, 0.f, 0.f, 0.f, 1.f, 0.f, 0.f, 0.f, 1.f};
    pt_camera_desc camera{};
    uint32_t width = 768, height = 576, max_depth = 1;
    bool any_normals = false, any_texcoords = false;
    // Loader notes for input the reference accepts with a fallback (unknown
    // conductor names, BSDF types and shape types).
    std::vector<std::string> warnings;

    uint32_t add_material(const pt_material_desc &m);
    // Flattens one instance: positions by `to_world`, normals by its inverse
    // transpose (not normalized).  Returns the shape index.
    uint32_t add_instance(const MeshData &mesh, const Mat4 &to_world, uint32_t material, uint32_t flags,
                          bool is_emitter, const pt_texture_desc *radiance);
    // Appends an analytic sphere shape: the unit sphere mapped by `to_world`
    // (only PT_SHAPE_FLIP_NORMALS of `flags` applies).  Returns the shape index.
    uint32_t add_sphere(const Mat4 &to_world, uint32_t material, uint32_t flags, bool is_emitter,
                        const pt_texture_desc *radiance);
    uint32_t add_bitmap(std::shared_ptr<BitmapData> data, uint32_t filter, uint32_t wrap);
    void set_camera(float fov_y_deg, const Mat4 &camera_to_world);
    pt_scene_desc desc() const;
};

MeshData make_rectangle_mesh();
MeshData make_cube_mesh();
MeshData make_icosphere_mesh(int subdivisions);

bool generate_scene(int kind, const pt_gen_params &params, HostScene &out, std::string &err);
bool load_xml_scene(const std::string &path, HostScene &out, std::string &err);
bool load_obj_mesh(const std::string &path, MeshData &mesh, std::string &err);
// util::BitmapTexture::Load: EXR / Radiance HDR / PNG as RGBA float, row 0 =
// top row; `err` starts with "io:" for file errors.
bool load_image(const std::string &path, uint32_t &w, uint32_t &h, std::vector<float> &rgba, std::string &err);
// Size limit shared by the image readers, checked before any pixel buffer is
// allocated; false (with `err`) above 2^24 pixels per side or 2^28 in all.
constexpr uint64_t kMaxImageSide = 1ull << 24, kMaxImagePixels = 1ull << 28;
bool image_size_ok(uint64_t width, uint64_t height, std::string &err);
bool write_exr_rgb(const std::string &path, const float *rgba, uint32_t width, uint32_t height, std::string &err);

// Dielectric IOR names (Hecht, Optics, 4th ed.) and conductor eta/k tables
// used by the XML material loader.
bool lookup_dielectric_ior(const std::string &name, float &value);
bool lookup_conductor_ior(const std::string &name, float eta[3], float k[3]);

}  // namespace pupil
