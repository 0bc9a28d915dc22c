/*
 * CPU oracle of the MIS path tracer -- TEST INFRASTRUCTURE ONLY.
 *
 * A scalar restatement of PupilOptixLab's path tracer
 * (example/path_tracer/main.cu:38-246 and the files it includes), used by the
 * tests, by smoke() and by the benchmark's CPU-baseline leg as the checker.
 * The product (pupiloptixlab_amd) never includes, links or calls it.
 *
 * Scene input: a struct with the byte layout of pt_scene_desc (declared here
 * independently as or_scene).  Shapes whose flags hold 32 are analytic
 * spheres: first_tri/num_tris then index spheres[], and sphere k has
 * primitive id (1 << 30) | k.  All functions are thread-safe.
 */
#ifndef PUPIL_ORACLE_H
#define PUPIL_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct or_texture {
    uint32_t type; /* 0 rgb, 1 checkerboard, 2 bitmap */
    uint32_t bitmap;
    float color0[3];
    float color1[3];
    float uv_row0[4];
    float uv_row1[4];
} or_texture;

typedef struct or_bitmap {
    uint32_t width, height;
    uint32_t filter; /* 0 point, 1 bilinear */
    uint32_t wrap;   /* 0 repeat, 1 clamp, 2 mirror, 3 border */
    const float *rgba; /* row 0 sampled at v = 0 */
} or_bitmap;

typedef struct or_material {
    uint32_t type;
    uint32_t twosided;
    uint32_t nonlinear;
    float int_ior;
    float ext_ior;
    or_texture tex[4];
} or_material;

typedef struct or_shape {
    uint32_t material;
    uint32_t flags;
    uint32_t first_tri;
    uint32_t num_tris;
    int32_t is_emitter;
    or_texture radiance;
} or_shape;

typedef struct or_scene {
    uint32_t num_vertices, num_triangles, num_shapes, num_materials;
    const float *positions;
    const float *normals;
    const float *texcoords;
    const uint32_t *indices;
    const or_shape *shapes;
    const or_material *materials;
    uint32_t env_type;
    float env_radiance[3];
    float sample_to_camera[16];
    float camera_to_world[16];
    uint32_t width, height, max_depth;
    uint32_t num_spheres;
    const float *spheres; /* 12 floats each: object -> world 3x4 rows of the unit sphere */
    uint32_t num_bitmaps;
    const or_bitmap *bitmaps;
    uint32_t env_bitmap;  /* env_type 2: environment map */
    float env_scale;
    float env_to_world[9];
} or_scene;

typedef struct or_params {
    uint32_t spp;
    uint32_t max_depth;
    uint32_t rr_depth;
    uint32_t mis_mode;  /* 0 reference, 1 consistent */
    uint32_t accumulate;
    uint32_t row_begin; /* render rows [row_begin, row_end); row_end 0 = height */
    uint32_t row_end;
    uint32_t threads;   /* 0 = all cores */
} or_params;

typedef struct or_counts {
    uint64_t rays_extension;
    uint64_t rays_shadow;
    uint64_t node_visits;
    uint64_t tri_tests;
    double build_ms;   /* acceleration structure + emitter table */
    double render_ms;  /* the per-pixel loop only */
    uint32_t threads;  /* worker threads used */
} or_counts;

/* Renders params->spp frames from a reset accumulator.  Outputs (any may be
 * NULL): accum float4/pixel, albedo float3, normal float3, test float,
 * prim uint32 (primary hit, 0xFFFFFFFF miss), sum64 double3 (sum of frames). */
int oracle_render(const or_scene *scene, const or_params *params, float *accum, float *albedo, float *normal,
                  float *test, uint32_t *prim, double *sum64, or_counts *counts);

/* Rays: 8 floats each (origin, tmin, direction, tmax).  Hits: 4 words
 * (t, prim, b1, b2).  brute != 0 tests every triangle. */
int oracle_trace(const or_scene *scene, const float *rays, uint32_t count, int any_hit, int brute, uint32_t *hits);

/* Unit entry points. */
void oracle_rng(uint32_t pixel, uint32_t sample, uint32_t n, float *out, uint32_t *seeds);
void oracle_sincos(float x, float *s, float *c);
/* Camera ray for pixel (x, y) with jitter (jx, jy): out = origin[3], dir[3]. */
void oracle_camera_ray(const or_scene *scene, uint32_t x, uint32_t y, float jx, float jy, float *out6);
/* BSDF of one material at uv (0,0).  wo/wi local.  eval: f[3], pdf.
 * sample: consumes draws from *rng_seed; out = wi[3], f[3], pdf, lobe. */
void oracle_bsdf_eval(const or_material *m, const float *wo, const float *wi, float *f3_pdf);
void oracle_bsdf_sample(const or_material *m, const float *wo, uint32_t *rng_seed, float *out8);
float oracle_fresnel_dielectric(float eta, float cos_i, float *cos_t);
float oracle_fresnel_conductor(float eta, float k, float cos_i);
float oracle_ggx_d(const float *wh, float alpha);
float oracle_ggx_pdf(const float *wo, const float *wh, float alpha);
void oracle_ggx_sample(const float *wo, float alpha, float u1, float u2, float *wh);
/* Sphere unit functions: atan2 and acos as the oracle evaluates them, the
 * sphere texcoord (phi / 2pi, theta / pi) of a unit direction, and the ray /
 * sphere test (nearest root in [tmin, tmax]; returns 1 on hit). */
/* Bitmap lookup (point / bilinear, repeat / clamp / mirror / border) at
 * texture coordinates (u, v). */
void oracle_bitmap_sample(const or_bitmap *b, float u, float v, float *out3);
float oracle_atan2(float y, float x);
float oracle_acos(float z);
void oracle_sphere_texcoord(const float *dir, float *uv);
int oracle_intersect_sphere(const float *o, const float *d, const float *center, float radius, float tmin,
                            float tmax, float *t);
/* Intersection of one ray with one triangle (watertight); returns 1 on hit. */
int oracle_intersect(const float *o, const float *d, const float *v0, const float *v1, const float *v2,
                     float *t_b1_b2);
/* Emitter table of the scene: per area emitter select prob, area; env prob. */
int oracle_emitter_table(const or_scene *scene, float *select_prob, float *area, uint32_t capacity,
                         uint32_t *count, float *env_prob);

#ifdef __cplusplus
}
#endif

#endif
