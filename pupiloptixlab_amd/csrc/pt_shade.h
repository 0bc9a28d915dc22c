// Per-path shading step of the wavefront integrator: what
// example/path_tracer/main.cu:87-188 does between two optixTrace calls,
// written as one step of a state machine over (ray, hit, path state).
//
// A path arrives with the result of its last extension ray.  The step adds
// the emission found there (with the MIS weight of main.cu:174-185 or the
// environment term of main.cu:168-172), then runs one loop iteration of
// main.cu:106-166: depth test, Russian roulette, next-event estimation and
// BSDF sampling.  It emits at most one continuation ray and at most one
// shadow ray carrying the pending NEE contribution; the shadow stage adds
// that contribution when the ray is unoccluded, so radiance terms are summed
// in the reference order.
#pragma once

#include "pt_bvh.h"
#include "pt_records.h"

namespace pupil {

enum PathFlags : uint32_t {
    kPathDepthMask = 0xFFu,
    kPathPrimary = 1u << 8,     // the last ray was the camera ray
    kPathPrevDelta = 1u << 9,   // the last BSDF sample came from a delta lobe
    kPathEnvUpper = 1u << 10,   // last direction lies in the upper hemisphere of the last normal
};

constexpr uint32_t kMissPrim = 0xFFFFFFFFu;
constexpr float kExtTmin = 0.001f;   // main.cu:82, 163
constexpr float kShadowEps = 0.0001f; // main.cu:126

struct SceneView {
    const uint32_t *tri;        // 4 per triangle: i0, i1, i2, shape
    const f4 *vpos;
    const f4 *vnrm;             // may be null
    const f2 *vuv;              // may be null
    const ShapeRec *shapes;
    const SphereRec *spheres;   // may be null without spheres
    const MaterialRec *materials;
    const EmitterRec *emitters;
    const float *emitter_cdf;
    uint32_t num_emitters;
    uint32_t pad;
    TexArrays tex;              // bitmaps (may be null without bitmap textures)
    EnvRec env;
};

struct ShadeParams {
    uint32_t max_depth;
    uint32_t rr_depth;
    uint32_t mis_mode;  // 0 reference, 1 consistent
    uint32_t aov;
};

struct PathState {
    f3 o, d;          // last ray
    float prev_pdf;   // pdf of the BSDF sample that produced d
    f3 T, L;          // throughput, radiance
    Rng rng;
    uint32_t flags;   // depth | PathFlags
};

struct ShadeOut {
    bool cont;        // continuation ray in (o, d)
    bool done;        // path finished: L is final
    bool shadow;      // shadow ray pending
    f3 so, sd;        // shadow ray
    float stmax;
    f3 contrib;       // NEE contribution added when unoccluded
    // AOVs of the primary hit
    bool wrote_aov;
    f3 albedo, normal;
    float test;
};

// Sphere hit (render/geometry.h:94-100): position o + t d, texcoord and
// normal from the object-space point (the normal through the transposed
// world -> object matrix, i.e. the inverse transpose of object -> world).
PT_HD void sphere_hit_geometry(const SceneView &s, uint32_t k, float t, f3 ray_o, f3 ray_dir, LocalGeometry &g,
                               uint32_t &material, int32_t &emitter_index, uint32_t &flags) {
    const SphereRec &sp = s.spheres[k];
    const ShapeRec sh = s.shapes[sp.shape];
    g.position = ray_o + t * ray_dir;
    const f4 p = f4{g.position.x, g.position.y, g.position.z, 1.f};
    const f3 local = mk3(dot(sp.w2o[0], p), dot(sp.w2o[1], p), dot(sp.w2o[2], p));
    g.texcoord = sphere_texcoord(normalize(local));
    const f3 n = mk3(sp.w2o[0].x * local.x + sp.w2o[1].x * local.y + sp.w2o[2].x * local.z,
                     sp.w2o[0].y * local.x + sp.w2o[1].y * local.y + sp.w2o[2].y * local.z,
                     sp.w2o[0].z * local.x + sp.w2o[1].z * local.y + sp.w2o[2].z * local.z);
    g.normal = normalize(n);
    material = sh.material;
    flags = sh.flags;
    emitter_index = sh.emitter_offset >= 0 ? sh.emitter_offset + (int32_t)(k - sh.first_tri) : -1;
}

// Reconstructs the hit (render/geometry.h:60-100, 176-180) in world space.
// Triangles use the barycentrics (b1, b2), spheres the ray and its t.
PT_HD void hit_geometry(const SceneView &s, uint32_t prim, float t_hit, float b1, float b2, f3 ray_o, f3 ray_dir,
                        LocalGeometry &g, uint32_t &material, int32_t &emitter_index, uint32_t feat = kFeatAll) {
    if ((feat & kFeatSphere) && (prim & kPrimSphere)) {
        uint32_t flags;
        sphere_hit_geometry(s, prim & kPrimIndexMask, t_hit, ray_o, ray_dir, g, material, emitter_index, flags);
        if (flags & kShapeFlipNormals) g.normal = g.normal * -1.f;
        if (dot(-ray_dir, g.normal) < 0.f && (s.materials[material].flags & kMatTwosided)) g.normal = -g.normal;
        return;
    }
    const uint32_t *t = s.tri + 4 * (size_t)prim;
    const uint32_t i0 = t[0], i1 = t[1], i2 = t[2];
    const ShapeRec sh = s.shapes[t[3]];
    const f4 q0 = s.vpos[i0], q1 = s.vpos[i1], q2 = s.vpos[i2];
    const f3 p0 = mk3(q0.x, q0.y, q0.z), p1 = mk3(q1.x, q1.y, q1.z), p2 = mk3(q2.x, q2.y, q2.z);
    const float w = 1.f - b1 - b2;
    g.position = w * p0 + b1 * p1 + b2 * p2;
    f3 n;
    if (sh.flags & kShapeHasNormals) {
        const f4 m0 = s.vnrm[i0], m1 = s.vnrm[i1], m2 = s.vnrm[i2];
        n = w * mk3(m0.x, m0.y, m0.z) + b1 * mk3(m1.x, m1.y, m1.z) + b2 * mk3(m2.x, m2.y, m2.z);
    } else {
        n = cross(p1 - p0, p2 - p0);
        if (sh.flags & kShapeNegFaceNormal) n = -n;
    }
    n = normalize(n);
    if (sh.flags & kShapeFlipNormals) n = n * -1.f;
    f2 uv = mk2(0.f, 0.f);
    if (sh.flags & kShapeHasTexcoords) {
        const f2 t0 = s.vuv[i0], t1 = s.vuv[i1], t2 = s.vuv[i2];
        uv = w * t0 + b1 * t1 + b2 * t2;
        if (sh.flags & kShapeFlipTexcoords) uv.y = 1.f - uv.y;
    }
    material = sh.material;
    const MaterialRec &m = s.materials[sh.material];
    if (dot(-ray_dir, n) < 0.f && (m.flags & kMatTwosided)) n = -n;
    g.normal = n;
    g.texcoord = uv;
    emitter_index = sh.emitter_offset >= 0 ? sh.emitter_offset + (int32_t)(prim - sh.first_tri) : -1;
}

// EmitterGroup::SelectOneEmiiter (render/emitter.h:110-136) as a binary search
// over prefix sums formed in the same order as the reference's running sum.
// Returns an area emitter index, -1 for the environment, -2 for none.
PT_HD int32_t select_emitter(const SceneView &s, float p) {
    uint32_t lo = 0, hi = s.num_emitters;
    while (lo < hi) {
        uint32_t mid = (lo + hi) >> 1;
        if (p <= s.emitter_cdf[mid]) hi = mid;
        else lo = mid + 1;
    }
    if (lo < s.num_emitters) return (int32_t)lo;
    if (s.env.type != kEmitterNone) return -1;
    if (s.num_emitters > 0) return (int32_t)s.num_emitters - 1;
    return -2;
}

// Receives the AOVs of a camera path's first vertex (albedo, shading normal,
// the `test` draw) when ShadeParams::aov is set, at the point they are
// known, so a kernel can store them there instead of carrying them through
// the rest of the step; the default keeps them in ShadeOut only.
struct NoAovWrite {
    PT_HD void operator()(f3, f3, float) const {}
};

// `feat`: the scene features compiled in (kFeat*; a constant of the calling
// kernel, so the code of absent features is dropped).
template <typename AovWrite = NoAovWrite>
PT_HD void shade_path(const SceneView &s, const ShadeParams &prm, PathState &ps, uint32_t prim, float t_hit,
                      float b1, float b2, ShadeOut &out, uint32_t feat = kFeatAll,
                      const AovWrite &aov_write = AovWrite()) {
    out.cont = false;
    out.done = false;
    out.shadow = false;
    out.wrote_aov = false;
    const bool primary = (ps.flags & kPathPrimary) != 0;
    const bool consistent = prm.mis_mode != 0;
    uint32_t depth = ps.flags & kPathDepthMask;

    if (prim == kMissPrim) {
        // __miss__default (main.cu:199-215) and the escape branch (main.cu:168-172).
        f3 env_rad = mk3(0.f);
        float env_pdf = 0.f;    // the environment's Eval pdf (reference weighting)
        float light_pdf = 0.f;  // selection x next-event density (consistent weighting)
        if (s.env.type == kEmitterConstEnv) {
            env_rad = mk3(s.env.color.x, s.env.color.y, s.env.color.z);
            env_pdf = 0.25f * kInvPi;
            light_pdf = (ps.flags & kPathEnvUpper) ? s.env.select_prob * (kInvPi * 0.5f) : 0.f;
        } else if ((feat & kFeatEnvMap) && s.env.type == kEmitterEnvMap) {
            if (consistent) {
                float p;
                env_map_eval(s.env, ps.d, true, s.tex, env_rad, p);
                light_pdf = s.env.select_prob * p;
            } else {
                // The miss program evaluates the direction (o + d) - o.
                env_map_eval(s.env, normalize((ps.o + ps.d) - ps.o), false, s.tex, env_rad, env_pdf);
            }
        }
        if (primary) {
            out.test = rng_next(ps.rng);
            if (prm.aov) {
                out.wrote_aov = true;
                out.albedo = mk3(0.f);
                out.normal = mk3(0.f);
                aov_write(out.albedo, out.normal, out.test);
            }
        } else {
            float mis;
            if (consistent) {
                mis = (ps.flags & kPathPrevDelta) ? 1.f : mis_weight(ps.prev_pdf, light_pdf);
            } else {
                mis = mis_weight(ps.prev_pdf, env_pdf);
            }
            env_rad = env_rad * (ps.T * mis);
        }
        ps.L = ps.L + env_rad;
        out.done = true;
        return;
    }

    LocalGeometry g;
    uint32_t mat_index;
    int32_t emitter_index;
    hit_geometry(s, prim, t_hit, b1, b2, ps.o, ps.d, g, mat_index, emitter_index, feat);
    const MaterialRec &mat = s.materials[mat_index];
    const LocalBsdf bsdf = material_local_bsdf(mat, g.texcoord, s.tex, feat);

    if (primary) {
        if (emitter_index >= 0)
            ps.L = ps.L + texture_sample(s.emitters[emitter_index].radiance, g.texcoord, s.tex, feat);
        out.test = rng_next(ps.rng);
        if (prm.aov) {
            out.wrote_aov = true;
            out.albedo = bsdf.c0;
            out.normal = g.normal;
            aov_write(out.albedo, out.normal, out.test);
        }
    } else if (emitter_index >= 0) {
        const EmitterRec &e = s.emitters[emitter_index];
        f3 rad;
        float epdf;
        area_eval(e, g, ps.o, rad, epdf, s.tex, feat);
        if (!is_zero(epdf)) {
            // Consistent mode weighs against the pdf next-event estimation
            // would assign to this point, which uses the emitter's own normals.
            const float lpdf = consistent ? emitter_nee_pdf(e, g.position, b1, b2, ps.o, feat) : epdf;
            float mis = (ps.flags & kPathPrevDelta) ? 1.f : mis_weight(ps.prev_pdf, lpdf * e.select_prob);
            ps.L = ps.L + ps.T * rad * mis;
        }
    }

    // One iteration of the bounce loop (main.cu:106-166).
    ++depth;
    if (depth >= prm.max_depth) {
        out.done = true;
        return;
    }
    const float rr = depth > prm.rr_depth ? 0.95f : 1.0f;
    if (rng_next(ps.rng) > rr) {
        out.done = true;
        return;
    }
    ps.T = ps.T / rr;

    // Next-event estimation (main.cu:116-144).
    {
        const float psel = rng_next(ps.rng);
        const float x = rng_next(ps.rng);
        const float y = rng_next(ps.rng);
        const int32_t sel = select_emitter(s, psel);
        EmitterSample es;
        es.pdf = 0.f;
        float sel_prob = 0.f;
        if (sel >= 0) {
            emitter_sample_direct(s.emitters[sel], g, mk2(x, y), es, s.tex, feat);
            sel_prob = s.emitters[sel].select_prob;
        } else if (sel == -1) {
            env_sample_direct(s.env, g, mk2(x, y), consistent, es, s.tex, feat);
            sel_prob = s.env.select_prob;
        }
        if (es.pdf != 0.f) {
            const f3 wi_l = to_local(es.wi, g.normal);
            const f3 wo_l = to_local(-ps.d, g.normal);
            f3 f;
            float pdf;
            bsdf_eval(bsdf, wo_l, wi_l, f, pdf);
            if (!is_zero(f * es.pdf)) {
                const float nol = dot(g.normal, es.wi);
                if (nol > 0.f) {
                    float mis;
                    float lpdf;
                    if (consistent) {
                        lpdf = es.pdf * sel_prob;
                        mis = mis_weight(lpdf, pdf);
                    } else {
                        mis = mis_weight(es.pdf, pdf);
                        lpdf = es.pdf * sel_prob;
                    }
                    out.contrib = ps.T * es.radiance * f * nol * mis / lpdf;
                    out.shadow = true;
                    out.so = g.position;
                    out.sd = es.wi;
                    out.stmax = es.distance - kShadowEps;
                }
            }
        }
    }

    // BSDF sampling (main.cu:146-160).
    {
        const f3 wo_l = to_local(-ps.d, g.normal);
        BsdfSample bs;
        bsdf_sample(bsdf, wo_l, ps.rng, bs);
        if (is_zero(bs.f * fabsf(bs.wi.z)) || is_zero(bs.pdf) || !finite_sample(bs.f, bs.pdf)) {
            out.done = true;
            return;
        }
        ps.T = ps.T * (bs.f * fabsf(bs.wi.z) / bs.pdf);
        ps.o = g.position;
        ps.d = to_world(bs.wi, g.normal);
        ps.prev_pdf = bs.pdf;
        uint32_t fl = depth;
        if (bs.lobe & kLobeDelta) fl |= kPathPrevDelta;
        if (dot(g.normal, ps.d) > 0.f) fl |= kPathEnvUpper;
        ps.flags = fl;
        out.cont = true;
    }
}

}  // namespace pupil
