/*
 * pupil_hip.h -- C ABI of the MI355X-native wavefront MIS path tracer.
 *
 * This is the boundary that replaces the OptiX path-tracing pass of
 * PupilOptixLab.  Every entry point names the reference interface it stands
 * in for.  Plain C types only: pointers, sizes and POD structs.  All calls
 * return 0 on success or a negative pt_status; no call aborts on bad input.
 * A context is not thread-safe: callers serialize, as the reference's single
 * render thread does (system/system.cpp:93-106).
 *
 * Buffer conventions follow the reference pass: float4 RGBA with w = 1,
 * row 0 = bottom of the image (util/camera.cpp:9-16).
 */
#ifndef PUPIL_HIP_H
#define PUPIL_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum pt_status {
    PT_OK = 0,
    PT_ERR_INVALID_ARG = -1,
    PT_ERR_HIP = -2,
    PT_ERR_NO_SCENE = -3,
    PT_ERR_CAPACITY = -4,
    PT_ERR_IO = -5,
    PT_ERR_PARSE = -6,
    PT_ERR_OOM = -7,
    PT_ERR_UNSUPPORTED = -8,
    PT_ERR_NO_BVH = -9,
    PT_ERR_NO_GUIDE = -10 /* guide API called before pt_guide_enable */
};

/* Material types: EMatType order of decl/material_decl.inl:3-9. */
enum pt_material_type {
    PT_MAT_DIFFUSE = 1,
    PT_MAT_DIELECTRIC = 2,
    PT_MAT_ROUGH_DIELECTRIC = 3,
    PT_MAT_CONDUCTOR = 4,
    PT_MAT_ROUGH_CONDUCTOR = 5,
    PT_MAT_PLASTIC = 6,
    PT_MAT_ROUGH_PLASTIC = 7
};

/* Texture kinds of util/texture.h (RGB, checkerboard, bitmap). */
enum pt_texture_type { PT_TEX_RGB = 0, PT_TEX_CHECKER = 1, PT_TEX_BITMAP = 2 };

/* Shape flags (render/geometry.h:14-27 TriMesh and Sphere flags, flattened). */
enum pt_shape_flags {
    PT_SHAPE_FLIP_NORMALS = 1,
    PT_SHAPE_FLIP_TEXCOORDS = 2,
    PT_SHAPE_HAS_NORMALS = 4,
    PT_SHAPE_HAS_TEXCOORDS = 8,
    PT_SHAPE_NEG_FACE_NORMAL = 16,
    PT_SHAPE_SPHERE = 32 /* analytic spheres: first_tri/num_tris index spheres[] */
};

enum pt_env_type { PT_ENV_NONE = 0, PT_ENV_CONSTANT = 1, PT_ENV_MAP = 2 };

/* Bitmap sampling modes (util/texture.h:10-20, applied through the CUDA
 * texture object of cuda/texture.cpp:60-100). */
enum pt_bitmap_filter { PT_FILTER_POINT = 0, PT_FILTER_LINEAR = 1 };
enum pt_bitmap_wrap { PT_WRAP_REPEAT = 0, PT_WRAP_CLAMP = 1, PT_WRAP_MIRROR = 2, PT_WRAP_BORDER = 3 };

/* util::BitmapTexture (util/texture.h:38-50): RGBA float texels, row 0 is
 * the first (top) row of the image file and is sampled at v = 0.  Textures
 * reference a bitmap by index (pt_texture_desc.bitmap with PT_TEX_BITMAP). */
typedef struct pt_bitmap_desc {
    uint32_t width;
    uint32_t height;
    uint32_t filter;   /* pt_bitmap_filter */
    uint32_t wrap;     /* pt_bitmap_wrap */
    const float *rgba; /* width * height * 4 floats */
} pt_bitmap_desc;

/* cuda::Texture (cuda/texture.h:10-27): a color, a checkerboard of two
 * colors, or a bitmap; uv_row0/uv_row1 transform (u, v, 0, 1). */
typedef struct pt_texture_desc {
    uint32_t type;
    uint32_t bitmap;
    float color0[3];
    float color1[3];
    float uv_row0[4];
    float uv_row1[4];
} pt_texture_desc;

/* resource::Material (resource/material.h:59-74).  tex[] slots by type:
 *   diffuse:          0 reflectance
 *   dielectric:       0 specular_reflectance, 1 specular_transmittance
 *   rough dielectric: 0, 1 as dielectric, 2 alpha
 *   conductor:        0 specular_reflectance, 1 eta, 2 k
 *   rough conductor:  0, 1, 2 as conductor, 3 alpha
 *   plastic:          0 diffuse_reflectance, 1 specular_reflectance
 *   rough plastic:    0, 1 as plastic, 2 alpha                              */
typedef struct pt_material_desc {
    uint32_t type;
    uint32_t twosided;
    uint32_t nonlinear;
    float int_ior;
    float ext_ior;
    pt_texture_desc tex[4];
} pt_material_desc;

/* One shape instance flattened to world space (resource::ShapeInstance,
 * resource/shape.h:87-96 + world::RenderObject): a contiguous triangle
 * range (or, with PT_SHAPE_SPHERE, a contiguous range of spheres[]), its
 * material, and an area emitter when is_emitter != 0 (one emitter per
 * triangle or sphere, in shape order). */
typedef struct pt_shape_desc {
    uint32_t material;
    uint32_t flags;
    uint32_t first_tri;
    uint32_t num_tris;
    int32_t is_emitter;
    pt_texture_desc radiance;
} pt_shape_desc;

/* Analytic sphere (ShapeLoader<_sphere>, resource/shape.cpp:113-132 and
 * :245-247; render/geometry.h:23-27): the unit sphere at the origin mapped
 * by to_world (3x4 rows, object -> world; for the XML loader
 * to_world = transform * T(center) * S(radius)).  to_world must be a
 * similarity (rotation, uniform scale, translation), as the reference's
 * sphere emitter assumes (world/emitter.cpp:224-241). */
typedef struct pt_sphere_desc {
    float to_world[12];
} pt_sphere_desc;

/* optix::Camera (render/camera.h:7-10): row-major 4x4 matrices applied to
 * column vectors. */
typedef struct pt_camera_desc {
    float sample_to_camera[16];
    float camera_to_world[16];
} pt_camera_desc;

/* The world handed to the pass on SceneLoad (pt_pass.cpp:127-240):
 * world-space vertices (normals are M^-T n, not normalized), triangles,
 * shapes, materials, the environment and the camera. */
typedef struct pt_scene_desc {
    uint32_t num_vertices;
    uint32_t num_triangles;
    uint32_t num_shapes;
    uint32_t num_materials;
    const float *positions;  /* 3 per vertex */
    const float *normals;    /* 3 per vertex, or NULL */
    const float *texcoords;  /* 2 per vertex, or NULL */
    const uint32_t *indices; /* 3 per triangle */
    const pt_shape_desc *shapes;
    const pt_material_desc *materials;
    uint32_t env_type;
    float env_radiance[3];
    pt_camera_desc camera;
    uint32_t width;
    uint32_t height;
    uint32_t max_depth;
    uint32_t num_spheres;
    const pt_sphere_desc *spheres; /* num_spheres entries, or NULL */
    uint32_t num_bitmaps;
    const pt_bitmap_desc *bitmaps; /* num_bitmaps entries, or NULL */
    /* PT_ENV_MAP (resource::EnvMap, resource/scene.cpp:209-222, and
     * EnvMapEmitter, render/emitter/env.h:6-65): radiance = bitmaps[env_bitmap]
     * (bilinear, repeat) x env_scale; env_to_world = the 3x3 part of the
     * emitter's to_world (rows). */
    uint32_t env_bitmap;
    float env_scale;
    float env_to_world[9];
} pt_scene_desc;

typedef struct pt_bvh_opts {
    uint32_t num_bins; /* SAH bins, 0 = default (32) */
    uint32_t threads;  /* host build threads, 0 = all */
} pt_bvh_opts;

enum pt_mis_mode { PT_MIS_REFERENCE = 0, PT_MIS_CONSISTENT = 1 };

enum pt_render_flags {
    PT_RENDER_AOV = 1,        /* write albedo / normal / test / primitive id */
    PT_RENDER_FP64_ACCUM = 2, /* also keep a double-precision sum */
    PT_RENDER_TIMING = 4,     /* record per-kernel HIP event times */
    PT_RENDER_GUIDE = 8,      /* sample with the learned guiding proposal (needs
                                 pt_guide_enable and PT_MIS_CONSISTENT) */
    PT_RENDER_GUIDE_TRAIN = 16, /* train the guide online on this render's paths */
    PT_RENDER_ONE_CHAIN = 32    /* run every launch as one chain of kernels (by default a launch's
                                   paths are split into up to four chains of at least 10^6 paths,
                                   run concurrently on streams of their own; the image is the same
                                   either way, and with one chain each kernel's event time is that
                                   of the kernel alone) */
};

/* Launch parameters (example/path_tracer/type.h:9-33 config + pass knobs). */
typedef struct pt_render_params {
    uint32_t spp;            /* frames (samples per pixel) to add */
    uint32_t max_depth;      /* 0 = scene value */
    uint32_t rr_depth;       /* Russian roulette when depth > rr_depth (reference: 2) */
    uint32_t mis_mode;       /* pt_mis_mode */
    uint32_t accumulate;     /* running mean over frames (accumulated_flag) */
    uint32_t flags;          /* pt_render_flags */
    uint32_t spp_per_launch; /* frames traced concurrently, 0 = automatic */
    uint32_t block_size;     /* kernel block size, 0 = automatic */
} pt_render_params;

enum pt_buffer {
    PT_BUF_FRAME = 0,   /* "final result" float4 */
    PT_BUF_ACCUM = 1,   /* "pt accum buffer" float4 */
    PT_BUF_ALBEDO = 2,  /* "albedo" float3 */
    PT_BUF_NORMAL = 3,  /* "normal" float3 */
    PT_BUF_TEST = 4,    /* "test" float */
    PT_BUF_PRIM_ID = 5, /* primary-hit triangle id, uint32 (0xFFFFFFFF = miss) */
    PT_BUF_ACCUM64 = 6  /* double3 sum of samples (PT_RENDER_FP64_ACCUM) */
};

typedef struct pt_stats {
    uint64_t samples;        /* camera samples rendered since the last reset */
    uint64_t rays_extension; /* camera + bounce rays traced */
    uint64_t rays_shadow;    /* shadow rays traced */
    uint64_t node_visits;    /* closest-hit traversal; only with counting enabled */
    uint64_t tri_tests;
    uint64_t shadow_node_visits; /* any-hit traversal; only with counting enabled */
    uint64_t shadow_tri_tests;
    double ms_total;         /* wall time of pt_render calls */
    /* Per kernel kind, for launches made with PT_RENDER_TIMING: the time during
     * which at least one of its dispatches ran (HIP events; the kernels of a
     * launch's concurrent chains overlap) and the number of launches timed (a
     * bounce of all chains counts once). */
    double ms_raygen, ms_trace, ms_shade, ms_shadow, ms_accumulate;
    uint64_t launches_trace, launches_shade, launches_shadow;
    uint32_t bvh_nodes, bvh_leaf_tris, bvh_depth;
    double bvh_build_ms;
    double ms_guide;          /* event time of guide training (collect + steps), with timing */
    uint64_t guide_steps;     /* optimizer steps launched by pt_render / pt_guide_train */
    uint64_t guide_samples;   /* training samples collected by pt_render */
    double ms_guide_infer;    /* event time of the network inference kernels, with timing */
    uint64_t guide_queries;   /* network evaluations made for sampling */
    double ms_refit;          /* wall time of BVH refits after shape updates */
    uint64_t refits;
} pt_stats;

typedef struct pt_ctx pt_ctx;

/* cuda::Context::Init + PTPass ctor (cuda/context.cpp:9-49,
 * pt_pass.cpp:29-37): bind a device, create the pass stream. */
int pt_create(int hip_device, pt_ctx **out);
/* System::Destroy / ~PTPass. */
void pt_destroy(pt_ctx *ctx);
/* PTPass::SetScene (pt_pass.cpp:127-240) + World::LoadScene packing
 * (world/world.cpp:108-146, world/emitter.cpp:169-337): copies the arrays,
 * builds material and emitter tables, allocates frame buffers. */
int pt_set_scene(pt_ctx *ctx, const pt_scene_desc *scene);
/* GASManager/IASManager builds (world/gas_manager.cpp:69-245,
 * world/ias_manager.cpp:29-114): SAH 8-wide compressed BVH. */
int pt_build_bvh(pt_ctx *ctx, const pt_bvh_opts *opts);
/* CameraHelper::GetCudaMemory (world/camera.cpp:72-92) after a camera
 * change; restarts accumulation like EWorldEvent::CameraChange. */
int pt_set_camera(pt_ctx *ctx, const float sample_to_camera[16], const float camera_to_world[16]);
/* m_dirty restart of PTPass::OnRun (pt_pass.cpp:40-49). */
int pt_reset_accumulation(pt_ctx *ctx);
/* Frame index of the next frame (launch_params.random_seed, pt_pass.cpp:45,56,
 * which the reference restarts at 0 with the accumulation).  Setting it keeps
 * the accumulation; a large offset draws samples independent of frames
 * 0, 1, ... (reference renders that must not share samples with a test). */
int pt_set_random_seed(pt_ctx *ctx, uint32_t frame);
/* RenderObject::UpdateTransform (world/render_object.cpp:72-75) with the
 * world's RenderInstanceTransform / RenderInstanceUpdate handling
 * (world/world.cpp:45-59): the shape's geometry becomes `transform` (3x4
 * rows, affine) applied to the geometry the shape had in pt_set_scene; its
 * area emitters and the emitter selection table are rebuilt
 * (EmitterHelper::ResetAreaEmitter + ComputeProbability) and accumulation
 * restarts (pt_pass.cpp:242-249).  The BVH keeps its topology and is refit
 * on the GPU before the next render or trace, as IASManager::GetIASHandle
 * updates a dirty IAS (world/ias_manager.cpp:187-205).  Triangle shapes must
 * not share vertices with other shapes; spheres must stay similarity
 * transforms (PT_ERR_UNSUPPORTED otherwise; nothing changes then). */
int pt_set_shape_transform(pt_ctx *ctx, uint32_t shape, const float transform[12]);
/* Refits the BVH now if shape updates are pending (otherwise no-op). */
int pt_refit_bvh(pt_ctx *ctx);
/* PTPass::OnRun x spp (pt_pass.cpp:39-57) + optix::Pass::Run
 * (optix/pass.h:71-89): renders `spp` frames; blocks until done. */
int pt_render(pt_ctx *ctx, const pt_render_params *params);
/* Allocation-free variant of pt_render.  No counterpart in the reference:
 * it is for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  It renders
 * like pt_render (same kernels, identical images for the same frames) but
 * takes its path queues, queue counters and traversal stack spill area from a
 * 640 MiB array that the library holds in device memory per GPU for the life
 * of the process (resolved once per device in pt_create), calls no hipMalloc,
 * hipFree or other allocating routine, and does not synchronize: it returns
 * once the work is enqueued on the context's stream (pt_read_buffer waits for
 * it).  Frames per launch = spp_per_launch, or as many as the array holds;
 * one frame must fit (paths of 224 bytes: about 2.4M pixels; 1920x1080 fits),
 * else PT_ERR_CAPACITY with the limit in pt_last_error.  Guided rendering and
 * PT_RENDER_TIMING give PT_ERR_UNSUPPORTED; pending pt_set_shape_transform
 * updates must be refit first (pt_refit_bvh).  Ray counts in pt_stats are not
 * updated.  Every call on one GPU shares that GPU's array: calls that overlap
 * on one GPU (several contexts or threads) race and must be serialized by the
 * caller; each call clears the counters it reads. */
int pt_render_noalloc(pt_ctx *ctx, const pt_render_params *params);
/* BufferManager::GetBuffer + device-to-host copy (system/buffer.cpp:34-99). */
int pt_read_buffer(pt_ctx *ctx, int which, void *host_dst, size_t bytes);
int pt_get_stats(pt_ctx *ctx, pt_stats *out);
/* Node-visit / triangle-test counting in the traversal kernels (off by default). */
int pt_set_counting(pt_ctx *ctx, int enable);
/* Last error text of the context (or of the last failed pt_create when ctx is NULL). */
const char *pt_last_error(const pt_ctx *ctx);

/* Device-to-device streaming copy of `bytes` (x iters) on the pass stream;
 * reports read+write GB/s.  No reference counterpart: used to measure the
 * achievable HBM bandwidth next to the kernels' roofline figures. */
int pt_bandwidth_probe(pt_ctx *ctx, uint64_t bytes, uint32_t iters, double *gbps);

/* Traces caller-supplied rays against the built BVH on the GPU (closest hit
 * unless any_hit != 0).  rays: 8 floats each (origin xyz, tmin, dir xyz,
 * tmax).  hits: 4 words each (t, primitive id, b1, b2). */
int pt_trace_rays(pt_ctx *ctx, const float *rays, uint32_t count, int any_hit, void *hits);

/* ---- denoising ---------------------------------------------------------
 * DenoisePass (system/denoise_pass.cpp:31-93) runs optix::Denoiser
 * (optix/denoiser.h:9-38) on the noisy buffer with optional albedo / normal
 * guides (EMode UseAlbedo = 1, UseNormal = 2) and writes "final result".  The
 * OptiX model is closed; pt_denoise filters with an edge-avoiding a-trous
 * wavelet transform (Dammertz et al. 2010) with variance-guided luminance
 * stopping (Schied et al. 2017) on albedo-demodulated color. */
enum pt_denoise_flags { PT_DENOISE_ALBEDO = 1, PT_DENOISE_NORMAL = 2 };

typedef struct pt_denoise_params {
    uint32_t flags;        /* pt_denoise_flags */
    uint32_t iterations;   /* a-trous levels, 0 = 5 (strides 1..16) */
    float sigma_luminance; /* luminance stopping in local standard deviations, 0 = 4 */
    float sigma_normal;    /* normal weight exponent, 0 = 64 */
    float sigma_albedo;    /* albedo distance scale, 0 = 0.1 */
} pt_denoise_params;

/* Denoises the context's "pt accum buffer" (PT_BUF_ACCUM) with the AOVs of
 * the last rendered frame as guides into "final result" (PT_BUF_FRAME). */
int pt_denoise(pt_ctx *ctx, const pt_denoise_params *params);
/* Same filter on caller images (host memory): rgba 4 floats per pixel,
 * albedo / normal 3 floats per pixel or NULL (flags then ignored for them),
 * out 4 floats per pixel. */
int pt_denoise_image(pt_ctx *ctx, uint32_t width, uint32_t height, const float *rgba, const float *albedo,
                     const float *normal, float *out, const pt_denoise_params *params);

/* ---- learned path guiding ---------------------------------------------
 * No reference counterpart: PupilOptixLab samples directions from the BSDF
 * only (example/path_tracer/main.cu:146-160).  The guide is a hash-grid
 * encoding of the hit position followed by an MLP whose 41 outputs define an
 * 8-lobe von Mises-Fisher mixture and a selection probability c; directions
 * are drawn from c * guide + (1 - c) * BSDF (one-sample MIS), so renders stay
 * unbiased whatever the network has learned. */
typedef struct pt_guide_config {
    uint32_t log2_table;      /* hash-table entries per level = 2^log2_table, 10..24 (default 19) */
    uint32_t n_min, n_max;    /* grid resolution of the coarsest / finest level (16, 2048) */
    uint32_t seed;            /* parameter initialization */
    float learning_rate;      /* Adam step size (1e-2) */
    uint32_t batch_size;      /* samples per optimizer step, multiple of 64 (16384) */
    uint32_t steps_per_frame; /* optimizer steps per frame rendered with PT_RENDER_GUIDE_TRAIN (8) */
    uint32_t train_stride;    /* one camera path in train_stride records training data (1) */
    uint32_t loss_flags;      /* pt_guide_loss_flags: terms left out of the loss (0 = all) */
    float fixed_selection;    /* > 0: sample the guide with this probability instead of the learned c */
    float l2_weight;          /* weight of the relative-L2 radiance term (0 = default 1e-3) */
    uint32_t guided_depth;    /* path vertices (from the camera) that may sample the guide; deeper
                                 vertices sample the BSDF only (0 = all) */
} pt_guide_config;

enum pt_guide_loss_flags {
    PT_GUIDE_NO_RADIANCE_L2 = 1, /* relative L2 of the radiance model */
    PT_GUIDE_NO_GUIDE_CE = 2,    /* cross-entropy of the guide density */
    PT_GUIDE_NO_MIXTURE_CE = 4   /* cross-entropy of the one-sample mixture (trains c) */
};

typedef struct pt_guide_info {
    uint32_t enabled;
    uint32_t log2_table;
    uint32_t steps;           /* applied optimizer steps */
    uint32_t skipped;         /* steps dropped for non-finite gradients */
    float loss_scale;         /* current dynamic loss scale */
    float last_loss;          /* mean loss of the last step */
    uint32_t last_samples;    /* samples in the last step */
    uint32_t resolution[16];  /* grid resolution per level */
    uint32_t dense[16];       /* 1 when the level is indexed densely */
    float bounds_lo[3];       /* position normalization of grid 0: u = (p - lo) * inv */
    float bounds_inv[3];
    uint32_t grids;           /* hash grids (1, or one per batch view) */
} pt_guide_info;

/* Allocates and initializes the guide over the bounds of the current scene
 * (config NULL = defaults).  Re-enabling re-initializes the parameters. */
int pt_guide_enable(pt_ctx *ctx, const pt_guide_config *config);
int pt_guide_disable(pt_ctx *ctx);
int pt_guide_get_info(pt_ctx *ctx, pt_guide_info *out);
/* Raw network outputs (41 floats per position) for n world-space positions
 * (3 floats each). */
int pt_guide_query(pt_ctx *ctx, const float *positions, uint32_t n, float *outputs);
/* Guide for the batch (pt_batch_*): one hash grid per batch view, over that
 * view's scene bounds, all feeding one shared MLP; pt_batch_render then takes
 * PT_RENDER_GUIDE / PT_RENDER_GUIDE_TRAIN.  Adding or clearing views
 * disables it.  Table layouts of the calls below become
 * [views][16][2^log2_table][2]. */
int pt_guide_enable_batch(pt_ctx *ctx, const pt_guide_config *config);
/* pt_guide_query on hash grid `grid` (view index of a batch guide). */
int pt_guide_query_grid(pt_ctx *ctx, uint32_t grid, const float *positions, uint32_t n, float *outputs);
/* Position normalization of hash grid `grid`: u = (p - lo) * inv. */
int pt_guide_grid_bounds(pt_ctx *ctx, uint32_t grid, float lo[3], float inv[3]);
/* Training sample layout, 12 floats: position xyz, hash grid (0 with one
 * grid; with several, a sample naming no grid counts as unused), direction xyz (unit),
 * 0, radiance target (luminance of the incident radiance times |cos|), F/p
 * (luminance of f L |cos| / pdf), BSDF pdf of the direction, weight (0 =
 * unused slot; the loss is the weighted mean).  fp_norm scales F/p in the
 * cross-entropy terms. */
/* One optimizer step on n samples (n a multiple of 64); *loss = mean loss. */
int pt_guide_train(pt_ctx *ctx, const float *samples, uint32_t n, float fp_norm, float *loss);
/* Gradients of the mean loss over n samples, without a step:
 * table_grad [16][2^log2_table][2], weight_grad [4][64][64] ([in][out]). */
int pt_guide_gradients(pt_ctx *ctx, const float *samples, uint32_t n, float fp_norm, float *table_grad,
                       float *weight_grad, float *loss);
/* FP32 master parameters, same layouts as pt_guide_gradients. */
int pt_guide_get_params(pt_ctx *ctx, float *table, float *weights);
int pt_guide_set_params(pt_ctx *ctx, const float *table, const float *weights);
/* Copies up to max_samples of the training samples collected by the last
 * training launch of pt_render (12 floats each, layout of pt_guide_train);
 * *count receives the number copied. */
int pt_guide_read_samples(pt_ctx *ctx, float *samples, uint32_t max_samples, uint32_t *count);

/* ---- batched multi-view rendering ----------------------------------------
 * No counterpart in the reference (one scene and one camera per pass).  Views
 * -- each a scene with its own camera, film, depth and frame count -- are
 * added to the context's batch; pt_batch_render traces all of them in one
 * wavefront loop: every launch packs frames of several views into the same
 * queues, so paths of different scenes and lengths share the per-bounce
 * kernels.  Each view renders exactly what pt_render would for its scene
 * (same seeds, same bits). */
enum pt_batch_flags {
    PT_BATCH_GRAPH = 1,    /* capture one launch's kernels in a hipGraph and replay it */
    PT_BATCH_AUTOTUNE = 2, /* pick the launch shape by timing candidates first (pt_batch_autotune) */
    PT_BATCH_GUIDE_GATE = 4 /* with PT_RENDER_GUIDE_TRAIN: each view training on T >= 16 frames measures,
                               on its training frames [T/4, T/2), every other one sampling the BSDF
                               only, each frame's relMSE against the view's leave-one-out mean.  A view
                               whose guided frames reach neither in mean nor in median 1/1.1 of its
                               plain frames' error stops training there and renders its remaining
                               frames without the guide (no network queries); a view that keeps the
                               guide trains on to T and, with frames after T, restarts its
                               accumulation there, so the image holds only frames of the trained
                               guide */
};

typedef struct pt_batch_params {
    uint32_t rr_depth;           /* Russian roulette when depth > rr_depth */
    uint32_t mis_mode;           /* pt_mis_mode */
    uint32_t render_flags;       /* PT_RENDER_FP64_ACCUM, PT_RENDER_TIMING; PT_RENDER_GUIDE and
                                    PT_RENDER_GUIDE_TRAIN after pt_guide_enable_batch */
    uint32_t batch_flags;        /* pt_batch_flags */
    uint32_t paths_per_launch;   /* queue capacity, 0 = tuned or default (32M) */
    uint32_t shade_block;        /* shading block size (64, 128, 256), 0 = tuned or 256 */
    uint32_t shade_blocks_per_cu;/* 0 = tuned or 8 */
    uint32_t trace_blocks_per_cu;/* traversal blocks per CU, 0 = tuned or resident maximum */
    uint32_t guide_train_spp;    /* with PT_RENDER_GUIDE_TRAIN: each view trains the guide on its first
                                    guide_train_spp frames of the call (0 = all frames) */
    uint32_t train_paths_per_launch; /* with PT_RENDER_GUIDE_TRAIN: paths per training launch (at least
                                    one frame of the largest view), 0 = 8M; the network is updated
                                    after every training launch */
} pt_batch_params;

typedef struct pt_batch_tuning {
    uint32_t paths_per_launch, shade_block, shade_blocks_per_cu, trace_blocks_per_cu;
    uint32_t candidates;         /* launch shapes timed */
    uint32_t probe_paths;        /* paths per timed probe */
    double best_ns_per_path;
    double default_ns_per_path;  /* the default shape on the same probe */
} pt_batch_tuning;

typedef struct pt_batch_stats {
    uint32_t views, launches;
    uint64_t paths, rays_extension, rays_shadow, triangles;
    double ms_total, build_ms;
    uint32_t guide_gated_off; /* views whose guide the last gated render switched off */
    uint32_t pad;
} pt_batch_stats;

/* Removes all views and frees their device data. */
int pt_batch_clear(pt_ctx *ctx);
/* Adds a view: converts the scene, builds its BVH (one tree per view in a
 * shared node array), allocates its film.  spp = frames rendered per call
 * of pt_batch_render. */
int pt_batch_add_view(pt_ctx *ctx, const pt_scene_desc *scene, uint32_t spp, uint32_t *view_index);
/* Renders every view's spp frames; blocks until done. */
int pt_batch_render(pt_ctx *ctx, const pt_batch_params *params);
/* PT_BUF_FRAME, PT_BUF_ACCUM (float4) or PT_BUF_ACCUM64 (double3) of a view. */
int pt_batch_read(pt_ctx *ctx, uint32_t view, int which, void *host_dst, size_t bytes);
/* Restarts accumulation of every view. */
int pt_batch_reset(pt_ctx *ctx);
/* Times launch shapes (queue capacity, shading and traversal grids) on the
 * batch's own frames and keeps the fastest for later renders.  Accumulation
 * is restarted afterwards. */
int pt_batch_autotune(pt_ctx *ctx, const pt_batch_params *params, pt_batch_tuning *out);
int pt_batch_get_stats(pt_ctx *ctx, pt_batch_stats *out);

/* What the gate of the last gated guided render measured for one view
 * (PT_BATCH_GUIDE_GATE): mean and median per-frame relMSE of its guided and
 * of its BSDF-only evaluation frames, their counts, and whether the view's
 * guide was switched off.  No reference counterpart: the gate belongs to the
 * learned guide; this reports why a view kept or lost it. */
typedef struct pt_batch_gate_info {
    double guided_mean, guided_median, plain_mean, plain_median;
    uint32_t guided_frames, plain_frames, guide_off, pad;
} pt_batch_gate_info;
int pt_batch_get_gate(pt_ctx *ctx, uint32_t view, pt_batch_gate_info *out);

/* ---- host scene helpers (resource::Scene / World, headless) ----------- */
typedef struct pt_host_scene pt_host_scene;

enum pt_scene_kind {
    PT_SCENE_CORNELL = 1,   /* closed box, two rotated boxes, one area light */
    PT_SCENE_TERRAIN = 2,   /* fractal heightfield + icospheres, 7 materials, 4 lights + env */
    PT_SCENE_DOOR_ROOM = 3, /* room lit through a narrow door gap (seed > 0: door and camera moved) */
    PT_SCENE_FURNACE = 4,   /* closed emissive diffuse cube seen from its center */
    PT_SCENE_CORNELL_VARIANT = 5 /* Cornell box tessellated to 50K-2M tris, random camera and film size */
};

typedef struct pt_gen_params {
    uint64_t seed;
    uint32_t width, height; /* 0 = scene default */
    uint32_t grid;          /* heightfield vertices per side (terrain), 0 = default 1024 */
    uint32_t spheres;       /* icospheres (terrain), 0 = default 256 */
    uint32_t max_depth;     /* 0 = scene default */
    float albedo;           /* furnace wall albedo, 0 = 0.5 */
    float radiance;         /* furnace wall emission, 0 = 1 */
    uint32_t triangles;     /* Cornell variant: target triangle count, 0 = drawn from the seed */
} pt_gen_params;

/* Scene::LoadFromXML (resource/scene.cpp:27-56) for the mitsuba subset. */
int pt_scene_load_xml(const char *path, pt_host_scene **out, char *err, size_t err_len);
int pt_scene_generate(int kind, const pt_gen_params *params, pt_host_scene **out);
/* Loader warnings, one per line, for input the reference accepts with a
 * fallback and a log line (unknown bsdf and shape types, resource/material.cpp
 * and resource/shape.cpp; an unknown conductor name, which material.cpp:72-80
 * replaces by eta 0, k 1 silently).  Returns the number of warnings (< 0 on a
 * bad argument); buf (len bytes, may be NULL) gets them newline-separated,
 * truncated to fit. */
int pt_scene_get_warnings(const pt_host_scene *scene, char *buf, size_t len);
/* Pointers in the returned desc stay valid until pt_scene_destroy. */
int pt_scene_get_desc(const pt_host_scene *scene, pt_scene_desc *out);
void pt_scene_destroy(pt_host_scene *scene);
/* util::BitmapTexture::Load (util/texture.cpp:87-174): reads an EXR
 * (NONE/RLE/ZIPS/ZIP), Radiance HDR or PNG file as RGBA float, row 0 = top
 * row, 8-bit channels linearized with (v/255)^2.2.  Call with rgba = NULL to
 * get the size; capacity is in floats. */
int pt_load_image(const char *path, uint32_t *width, uint32_t *height, float *rgba, size_t capacity,
                  char *err, size_t err_len);
/* Writes RGB float EXR (uncompressed scanlines), bottom row first in the
 * buffer as in util/texture.cpp:23-85. */
int pt_write_exr(const char *path, const float *rgba, uint32_t width, uint32_t height);
/* Host build + traversal of the same wide BVH the GPU uses (for checks on
 * machines without a GPU). */
int pt_host_trace_rays(const pt_scene_desc *scene, const float *rays, uint32_t count, int any_hit, void *hits,
                       uint64_t *node_visits, uint64_t *tri_tests);

#ifdef __cplusplus
}
#endif

#endif /* PUPIL_HIP_H */
