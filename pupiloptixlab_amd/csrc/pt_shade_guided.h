// Shading step of the guided integrator.
//
// Same state machine as shade_path (pt_shade.h) with balance-heuristic MIS
// that applies the emitter selection probability on both sides (the
// consistent mode), and one change in how directions are drawn: at a vertex
// with a non-delta BSDF the direction comes from the learned vMF mixture with
// probability c and from the BSDF otherwise, and every pdf the estimator uses
// for that vertex (throughput division, the MIS weight of next-event
// estimation, the MIS weight of an emitter hit by the sampled ray) is the
// effective density p = c q + (1 - c) p_bsdf.  With c = 0 (no network, delta
// or smooth-plastic BSDFs) the step reduces to plain BSDF sampling.
//
// Split in two so that the wave can run the network between them: guided_begin
// (emission, depth test, Russian roulette) and guided_scatter (next-event
// estimation and direction sampling).
#pragma once

#include "pt_guide.h"
#include "pt_shade.h"

namespace pupil {

constexpr uint32_t kPathRecorded = 1u << 11;  // the last vertex wrote a training record
constexpr uint32_t kPathNoTrain = 1u << 12;   // the path's frame does not train the guide (batched renders)
constexpr uint32_t kPathNoGuide = 1u << 13;   // the path samples the BSDF only (batched renders: a plain
                                              // evaluation frame, or a view whose guide was gated off)

struct GuidedVertex {
    LocalGeometry g;
    LocalBsdf bsdf;
    uint32_t depth;     // depth after this vertex
    bool guidable;      // non-delta BSDF
};

// Training-record sink of the guided step (null `rec` disables recording).
// Record component 2's w holds the hash grid of the vertex (its batch view).
struct RecordSink {
    f4 *rec;             // component-major: 5 x capacity float4
    uint32_t capacity;
    uint32_t stride, phase, max_depth, epoch;
};

PT_HD bool guidable_material(uint32_t type) {
    return type == kMatDiffuse || type == kMatRoughConductor || type == kMatRoughDielectric ||
           type == kMatRoughPlastic;
}

PT_HD bool training_path(const RecordSink &r, uint32_t path, uint32_t flags) {
    return r.rec != nullptr && (flags & kPathNoTrain) == 0u && path % r.stride == r.phase;
}

PT_HD uint32_t record_index(const RecordSink &r, uint32_t path, uint32_t depth) {
    return (path / r.stride) * r.max_depth + depth;
}

// Completes the record of the previous vertex with L_after - T Le (Le = the
// emission found here, unweighted).
PT_HD void complete_record(const RecordSink &r, uint32_t path, const PathState &ps, f3 le) {
    if (!(ps.flags & kPathRecorded)) return;
    const uint32_t depth = ps.flags & kPathDepthMask;
    if (depth == 0u) return;
    const uint32_t idx = record_index(r, path, depth - 1u);
    if (idx >= r.capacity) return;
    const f3 adj = ps.L - ps.T * le;
    r.rec[4 * (size_t)r.capacity + idx] = f4{adj.x, adj.y, adj.z, u2f(r.epoch)};
}

// Returns true when the path scatters at this vertex (guided_scatter follows).
PT_HD bool guided_begin(const SceneView &s, const ShadeParams &prm, PathState &ps, uint32_t prim, float t_hit,
                        float b1, float b2, const RecordSink &rs, uint32_t path, GuidedVertex &v, ShadeOut &out) {
    out.cont = false;
    out.done = false;
    out.shadow = false;
    out.wrote_aov = false;
    const bool primary = (ps.flags & kPathPrimary) != 0;
    const bool train = training_path(rs, path, ps.flags);
    uint32_t depth = ps.flags & kPathDepthMask;

    if (prim == kMissPrim) {
        f3 env_rad = mk3(0.f);
        float light_pdf = 0.f;
        if (s.env.type == kEmitterConstEnv) {
            env_rad = mk3(s.env.color.x, s.env.color.y, s.env.color.z);
            light_pdf = (ps.flags & kPathEnvUpper) ? s.env.select_prob * (kInvPi * 0.5f) : 0.f;
        } else if (s.env.type == kEmitterEnvMap) {
            float p;
            env_map_eval(s.env, ps.d, true, s.tex, env_rad, p);
            light_pdf = s.env.select_prob * p;
        }
        if (primary) {
            if (prm.aov) {
                out.wrote_aov = true;
                out.albedo = mk3(0.f);
                out.normal = mk3(0.f);
            }
            out.test = rng_next(ps.rng);
            ps.L = ps.L + env_rad;
        } else {
            const float mis = (ps.flags & kPathPrevDelta) ? 1.f : mis_weight(ps.prev_pdf, light_pdf);
            ps.L = ps.L + env_rad * (ps.T * mis);
            if (train) complete_record(rs, path, ps, env_rad);
        }
        out.done = true;
        return false;
    }

    uint32_t mat_index;
    int32_t emitter_index;
    hit_geometry(s, prim, t_hit, b1, b2, ps.o, ps.d, v.g, mat_index, emitter_index);
    const MaterialRec &mat = s.materials[mat_index];
    v.bsdf = material_local_bsdf(mat, v.g.texcoord, s.tex);
    v.guidable = guidable_material(v.bsdf.type);

    f3 le = mk3(0.f);
    if (primary) {
        if (emitter_index >= 0)
            ps.L = ps.L + texture_sample(s.emitters[emitter_index].radiance, v.g.texcoord, s.tex);
        if (prm.aov) {
            out.wrote_aov = true;
            out.albedo = v.bsdf.c0;
            out.normal = v.g.normal;
        }
        out.test = rng_next(ps.rng);
    } else if (emitter_index >= 0) {
        const EmitterRec &e = s.emitters[emitter_index];
        f3 rad;
        float epdf;
        area_eval(e, v.g, ps.o, rad, epdf, s.tex);
        if (!is_zero(epdf)) {
            const float lpdf = emitter_nee_pdf(e, v.g.position, b1, b2, ps.o);
            const float mis = (ps.flags & kPathPrevDelta) ? 1.f : mis_weight(ps.prev_pdf, lpdf * e.select_prob);
            ps.L = ps.L + ps.T * rad * mis;
            le = rad;
        }
    }
    if (train && !primary) complete_record(rs, path, ps, le);

    ++depth;
    v.depth = depth;
    if (depth >= prm.max_depth) {
        out.done = true;
        return false;
    }
    const float rr = depth > prm.rr_depth ? 0.95f : 1.0f;
    if (rng_next(ps.rng) > rr) {
        out.done = true;
        return false;
    }
    ps.T = ps.T / rr;
    return true;
}

// Next-event estimation and direction sampling.  `use_guide` selects the
// mixture path; `m` is read only then.
PT_HD void guided_scatter(const SceneView &s, PathState &ps, const GuidedVertex &v, bool use_guide,
                          const GuideMixture &m, const RecordSink &rs, uint32_t path, ShadeOut &out,
                          uint32_t grid = 0u) {
    const LocalGeometry &g = v.g;
    const float c = use_guide ? m.c : 0.f;
    const f3 wo_l = to_local(-ps.d, g.normal);

    // Next-event estimation, weighted against the effective sampling density.
    {
        const float psel = rng_next(ps.rng);
        const float x = rng_next(ps.rng);
        const float y = rng_next(ps.rng);
        const int32_t sel = select_emitter(s, psel);
        EmitterSample es;
        es.pdf = 0.f;
        float sel_prob = 0.f;
        if (sel >= 0) {
            emitter_sample_direct(s.emitters[sel], g, mk2(x, y), es, s.tex);
            sel_prob = s.emitters[sel].select_prob;
        } else if (sel == -1) {
            env_sample_direct(s.env, g, mk2(x, y), true, es, s.tex);
            sel_prob = s.env.select_prob;
        }
        if (es.pdf != 0.f) {
            const f3 wi_l = to_local(es.wi, g.normal);
            f3 f;
            float pdf;
            bsdf_eval(v.bsdf, wo_l, wi_l, f, pdf);
            if (!is_zero(f * es.pdf)) {
                const float nol = dot(g.normal, es.wi);
                if (nol > 0.f) {
                    const float lpdf = es.pdf * sel_prob;
                    const float pdir = use_guide ? c * guide_pdf(m, es.wi) + (1.f - c) * pdf : pdf;
                    const float mis = mis_weight(lpdf, pdir);
                    out.contrib = ps.T * es.radiance * f * nol * mis / lpdf;
                    out.shadow = true;
                    out.so = g.position;
                    out.sd = es.wi;
                    out.stmax = es.distance - kShadowEps;
                }
            }
        }
    }

    // One-sample MIS between the guide and the BSDF.
    f3 wi_w, wi_l, f;
    float pb;
    uint32_t lobe;
    if (use_guide && rng_next(ps.rng) < c) {
        const float u0 = rng_next(ps.rng);
        const float u1 = rng_next(ps.rng);
        const float u2 = rng_next(ps.rng);
        wi_w = guide_sample(m, u0, u1, u2);
        wi_l = to_local(wi_w, g.normal);
        bsdf_eval(v.bsdf, wo_l, wi_l, f, pb);
        lobe = kLobeUnknown;
    } else {
        BsdfSample bs;
        bsdf_sample(v.bsdf, wo_l, ps.rng, bs);
        wi_l = bs.wi;
        f = bs.f;
        pb = bs.pdf;
        lobe = bs.lobe;
        wi_w = to_world(bs.wi, g.normal);
    }
    const bool delta = (lobe & kLobeDelta) != 0u;
    float p = pb;
    if (use_guide && !delta) p = c * guide_pdf(m, wi_w) + (1.f - c) * pb;
    const bool record = v.guidable && !delta && training_path(rs, path, ps.flags);
    const uint32_t ridx = record ? record_index(rs, path, v.depth - 1u) : 0u;
    const size_t rn = rs.capacity;
    if (is_zero(f * fabsf(wi_l.z)) || is_zero(p) || !finite_sample(f, p)) {
        // A direction that carries nothing: recorded complete with a zero
        // target, so the guide learns to avoid it.
        if (record && ridx < rs.capacity) {
            rs.rec[ridx] = f4{g.position.x, g.position.y, g.position.z, pb};
            rs.rec[rn + ridx] = f4{wi_w.x, wi_w.y, wi_w.z, fabsf(wi_l.z)};
            rs.rec[2 * rn + ridx] = f4{0.f, 0.f, 0.f, (float)grid};
            rs.rec[3 * rn + ridx] = f4{0.f, 0.f, 0.f, 0.f};
            rs.rec[4 * rn + ridx] = f4{0.f, 0.f, 0.f, u2f(rs.epoch)};
        }
        out.done = true;
        return;
    }
    const f3 weight = f * fabsf(wi_l.z) / p;
    ps.T = ps.T * weight;
    ps.o = g.position;
    ps.d = wi_w;
    ps.prev_pdf = p;
    uint32_t fl = v.depth | (ps.flags & (kPathNoTrain | kPathNoGuide));
    if (delta) fl |= kPathPrevDelta;
    if (dot(g.normal, ps.d) > 0.f) fl |= kPathEnvUpper;
    if (record && ridx < rs.capacity) {
        rs.rec[ridx] = f4{g.position.x, g.position.y, g.position.z, pb};
        rs.rec[rn + ridx] = f4{wi_w.x, wi_w.y, wi_w.z, fabsf(wi_l.z)};
        rs.rec[2 * rn + ridx] = f4{ps.T.x, ps.T.y, ps.T.z, (float)grid};
        rs.rec[3 * rn + ridx] = f4{weight.x, weight.y, weight.z, 0.f};
        fl |= kPathRecorded;
    }
    ps.flags = fl;
    out.cont = true;
}

}  // namespace pupil
