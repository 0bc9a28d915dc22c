// gfx950 kernels of the wavefront path tracer and the C-ABI context that
// drives them.
//
// One frame (= one sample per pixel, as one optixLaunch of the reference,
// example/path_tracer/pt_pass.cpp:39-57) runs as:
//   raygen -> { trace -> shade -> shadow } x max_depth -> accumulate
// over structure-of-arrays path queues.  `spp_per_launch` frames run in the
// same queues concurrently; accumulation then folds them in frame order, so
// the result equals frame-by-frame rendering bit for bit.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <map>
#include <set>
#include <mutex>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "../../include/pupil_hip.h"
#include "pt_bvh.h"
#include "pt_batch_plan.h"
#include "pt_chains.h"
#include "pt_timing.h"
#include "pt_bvh_build.h"
#include "pt_denoise.h"
#include "pt_guide_state.h"
#include "pt_shade.h"
#include "pt_shade_guided.h"

using namespace pupil;

namespace {

enum KernelKind { kKindRaygen = 0, kKindTrace, kKindShade, kKindShadow, kKindAccumulate, kKindGuide, kKindGuideInfer };

constexpr int kTraceBlock = 128;
constexpr int kLdsStack = 8;   // stack entries per lane kept in LDS
constexpr uint32_t kShadeRounds = 64;  // shading grid = this many rounds of resident blocks
constexpr uint32_t kSpillEntries = kMaxBvhDepth + 2 - kLdsStack;

// --- queue layout -------------------------------------------------------------
struct PathQueue {
    float4 *ray0;  // origin.xyz, path id bits
    float4 *ray1;  // direction.xyz, previous BSDF pdf
    float4 *tl;    // throughput.xyz, radiance.x
    float4 *lr;    // radiance.y, radiance.z, rng bits, flags bits
    float4 *hit;   // t, primitive bits, b1, b2
};
struct ShadowQueue {
    float4 *s0;  // origin.xyz, tmax
    float4 *s1;  // direction.xyz, target bits
    float4 *s2;  // contribution.xyz
};

struct DeviceBvh {
    const WideNode *nodes;
    const LeafTri *tris;
    uint2 *spill;
};

__device__ __forceinline__ void load_node_words(const WideNode *nodes, uint32_t i, NodeWords &n) {
    const uint4 *p = reinterpret_cast<const uint4 *>(nodes + i);
    for (int k = 0; k < 5; ++k) {
        const uint4 v = p[k];
        n.w[4 * k + 0] = v.x;
        n.w[4 * k + 1] = v.y;
        n.w[4 * k + 2] = v.z;
        n.w[4 * k + 3] = v.w;
    }
}

__device__ __forceinline__ void load_node_dev(const WideNode *nodes, uint32_t i, WideNode &n) {
    const uint4 *p = reinterpret_cast<const uint4 *>(nodes + i);
    uint4 a = p[0], b = p[1], c = p[2], d = p[3], e = p[4];
    uint4 *q = reinterpret_cast<uint4 *>(&n);
    q[0] = a;
    q[1] = b;
    q[2] = c;
    q[3] = d;
    q[4] = e;
}

struct LdsStack {
    uint2 *lds;    // this lane's column: lds[k * stride]
    uint2 *spill;  // this lane's global overflow area
    uint32_t stride;
    uint32_t sp;
    __device__ void push(uint32_t a, uint32_t b) {
        if (sp < (uint32_t)kLdsStack) lds[sp * stride] = make_uint2(a, b);
        else spill[sp - kLdsStack] = make_uint2(a, b);
        ++sp;
    }
    __device__ void pop(uint32_t &a, uint32_t &b) {
        --sp;
        uint2 v = sp < (uint32_t)kLdsStack ? lds[sp * stride] : spill[sp - kLdsStack];
        a = v.x;
        b = v.y;
    }
    __device__ bool empty() const { return sp == 0; }
};

template <bool kAnyHit, bool kCount, bool kSph>
__device__ __forceinline__ bool trace_one(const DeviceBvh &bvh, uint2 *lds_col, uint32_t gtid, f3 o, f3 d,
                                          float tmin, float tmax, Hit &h, TraverseStats &st, uint32_t root = 0u) {
    LdsStack stack{lds_col, bvh.spill + (size_t)gtid * kSpillEntries, (uint32_t)kTraceBlock, 0u};
    auto ln = [&](uint32_t i, NodeWords &n) { load_node_words(bvh.nodes, i, n); };
    auto lt = [&](uint32_t i, LeafTri &t) {
        const float4 *p = reinterpret_cast<const float4 *>(bvh.tris + i);
        float4 a = p[0], b = p[1], c = p[2];
        t.v0 = f4{a.x, a.y, a.z, a.w};
        t.v1 = f4{b.x, b.y, b.z, b.w};
        t.v2 = f4{c.x, c.y, c.z, c.w};
    };
    return traverse_bvh<kAnyHit, kSph>(ln, lt, o, d, tmin, tmax, stack, h, kCount ? &st : nullptr, root);
}

__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }

// Wave-level stream compaction: one atomic per wave reserves slots for the
// lanes whose predicate holds; each lane gets base + its rank in the ballot.
__device__ __forceinline__ uint32_t wave_alloc(bool pred, uint32_t *counter) {
    const unsigned long long mask = __ballot(pred);
    if (mask == 0ull) return 0u;
    const uint32_t lane = lane_id();
    const int leader = __ffsll((long long)mask) - 1;
    uint32_t base = 0;
    if ((int)lane == leader) base = atomicAdd(counter, (uint32_t)__popcll(mask));
    base = __shfl(base, leader);
    const unsigned long long lt = lane == 0 ? 0ull : (mask & ((~0ull) >> (64 - lane)));
    return base + (uint32_t)__popcll(lt);
}

__device__ __forceinline__ f3 f3of(float4 v) { return mk3(v.x, v.y, v.z); }

// --- raygen ---------------------------------------------------------------------
// main.cu:39-78 for every (frame, pixel) of the launch.
// Queue order of a launch's camera paths: slot i traces frame i % frames of
// the pixel at place i / frames of the film's tile order.  The lanes of a
// wave thus carry a few neighbouring pixels in all the launch's frames: their
// camera rays nearly coincide, their later rays leave from nearly the same
// points, and compaction keeps them together at later bounces (terrain
// 1080p, against one frame after another: traversal 2.6% faster, launches
// 1.7% faster at 20 frames and 2.1% at 64).  Tile order (tiled_pixel): full
// 8x8 tiles first, then the pixels of a partial right strip and bottom strip
// in row order.  The path id keeps frame and pixel (frame * pixels + pixel),
// so films, AOVs and random streams do not depend on the order.
__device__ __forceinline__ uint32_t tiled_pixel(uint32_t j, uint32_t w, uint32_t h) {
    const uint32_t w8 = w & ~7u, h8 = h & ~7u, tiled = w8 * h8;
    if (j < tiled) {
        const uint32_t t = j >> 6, k = j & 63u, tiles_per_row = w8 >> 3;
        return ((t / tiles_per_row) * 8u + (k >> 3)) * w + (t % tiles_per_row) * 8u + (k & 7u);
    }
    j -= tiled;
    const uint32_t rw = w - w8;
    if (j < rw * h8) return (j / rw) * w + w8 + j % rw;
    return h8 * w + (j - rw * h8);
}

// One chain of a launch: frames [frame0, frame0 + frames) of the pixels at
// places [pix0, pix0 + pixels) of the tile order; frame f of the launch
// draws from random-stream key frame_seed0 + f.
__global__ void k_raygen(PathQueue q, CameraRec cam, uint32_t width, uint32_t height, uint32_t frames,
                         uint32_t frame_seed0, uint32_t *counters, uint32_t num_counters, uint32_t frame0,
                         uint32_t pix0, uint32_t pixels) {
    const uint32_t npix = width * height;
    const uint32_t total = pixels * frames;
    const uint32_t tid = blockIdx.x * blockDim.x + threadIdx.x;
    if (tid == 0) {
        counters[0] = total;
        for (uint32_t i = 1; i < num_counters; ++i) counters[i] = 0u;
    }
    for (uint32_t i = tid; i < total; i += gridDim.x * blockDim.x) {
        uint32_t f, place;
        chain_slot(frame0, frames, pix0, i, f, place);
        const uint32_t pixel = tiled_pixel(place, width, height);
        const uint32_t x = pixel % width, y = pixel / width;
        Rng rng = rng_init(pixel, frame_seed0 + f);
        const float jx = rng_next(rng);
        const float jy = rng_next(rng);
        f3 o, d;
        camera_ray(cam, x, y, width, height, jx, jy, o, d);
        q.ray0[i] = make_float4(o.x, o.y, o.z, u2f(f * npix + pixel));
        q.ray1[i] = make_float4(d.x, d.y, d.z, 0.f);
        q.tl[i] = make_float4(1.f, 1.f, 1.f, 0.f);
        q.lr[i] = make_float4(0.f, 0.f, u2f(rng.seed), u2f(kPathPrimary));
    }
}

// Wave-level work fetch: lane 0 reserves `n` consecutive items for the wave.
__device__ __forceinline__ uint32_t wave_fetch(uint32_t *work, uint32_t n) {
    uint32_t base = 0;
    if (lane_id() == 0) base = atomicAdd(work, n);
    return __shfl(base, 0);
}

// Work distribution of the persistent traversal kernels, 64 items per wave
// at a time: a wave's first chunk is fixed by its index, later chunks come
// from the shared cursor (which starts at zero) past the statically assigned
// range.  When the queue is shorter than the grid, idle waves exit without
// touching the cursor: one cursor atomic per wave would serialize the whole
// grid on one address even when there is nothing to trace.
struct WaveWork {
    uint32_t base, dyn0;
    __device__ __forceinline__ WaveWork() {
        const uint32_t wpb = blockDim.x >> 6;
        base = (blockIdx.x * wpb + (threadIdx.x >> 6)) * 64u;
        dyn0 = gridDim.x * wpb * 64u;
    }
    __device__ __forceinline__ void next(uint32_t *work) { base = dyn0 + wave_fetch(work, 64u); }
};

// --- closest-hit traversal ------------------------------------------------------
// Persistent grid sized to the resident occupancy; each wave pulls 64 rays at
// a time, so waves that drew short traversals keep working.
// Several scenes can share the queues (batched views): the view of a path
// sits above bit `path_bits` of its id and selects its tree's root in `roots`
// (null for a single scene, where path_bits is 31).
// Five waves per SIMD (at most 96 VGPRs): the closest-hit traversal keeps the
// next node's load in flight while it tests a leaf's triangles.
template <bool kCount, bool kSph>
__global__ __launch_bounds__(kTraceBlock, 5) void k_trace(PathQueue q, DeviceBvh bvh, const uint32_t *count_ptr,
                                                       uint32_t *work, unsigned long long *stat,
                                                       const uint32_t *roots, uint32_t path_bits) {
    __shared__ uint2 lds[kLdsStack * kTraceBlock];
    const uint32_t count = *count_ptr;
    const uint32_t gtid = blockIdx.x * blockDim.x + threadIdx.x;
    TraverseStats st{0, 0};
    for (WaveWork ww;; ww.next(work)) {
        const uint32_t base = ww.base;
        if (base >= count) break;
        const uint32_t i = base + lane_id();
        if (i < count) {
            const float4 r0 = q.ray0[i], r1 = q.ray1[i];
            const uint32_t root = roots ? roots[f2u(r0.w) >> path_bits] : 0u;
            Hit h;
            trace_one<false, kCount, kSph>(bvh, lds + threadIdx.x, gtid, f3of(r0), f3of(r1), kExtTmin,
                                           kMaxDistance, h, st, root);
            q.hit[i] = make_float4(h.t, u2f(h.prim), h.b1, h.b2);
        }
    }
    if (kCount) {
        atomicAdd(stat + 0, (unsigned long long)st.nodes);
        atomicAdd(stat + 1, (unsigned long long)st.tris);
    }
}

// --- shadow (any-hit) traversal ----------------------------------------------------
// Shadow-ray targets: bit 31 set = output slot, else next-queue slot; the
// view sits in bits [path_bits, 31), the slot below.
template <bool kCount, bool kSph>
__global__ __launch_bounds__(kTraceBlock) void k_shadow(ShadowQueue sq, DeviceBvh bvh, const uint32_t *count_ptr,
                                                        uint32_t *work, PathQueue next, float4 *out_l,
                                                        unsigned long long *stat, const uint32_t *roots,
                                                        uint32_t path_bits) {
    __shared__ uint2 lds[kLdsStack * kTraceBlock];
    const uint32_t count = *count_ptr;
    const uint32_t gtid = blockIdx.x * blockDim.x + threadIdx.x;
    TraverseStats st{0, 0};
    for (WaveWork ww;; ww.next(work)) {
        const uint32_t base = ww.base;
        if (base >= count) break;
        const uint32_t i = base + lane_id();
        bool occluded = true;
        float4 s1 = make_float4(0.f, 0.f, 0.f, 0.f);
        if (i < count) {
            const float4 s0 = sq.s0[i];
            s1 = sq.s1[i];
            const uint32_t root = roots ? roots[(f2u(s1.w) & 0x7FFFFFFFu) >> path_bits] : 0u;
            Hit h;
            occluded = trace_one<true, kCount, kSph>(bvh, lds + threadIdx.x, gtid, f3of(s0), f3of(s1), kShadowEps,
                                                     s0.w, h, st, root);
        }
        if (!occluded) {
            const float4 c = sq.s2[i];
            const uint32_t target = f2u(s1.w);
            const uint32_t slot = target & ((1u << path_bits) - 1u);
            if (target & 0x80000000u) {
                const uint32_t p = slot;
                float4 l = out_l[p];
                l.x = l.x + c.x;
                l.y = l.y + c.y;
                l.z = l.z + c.z;
                out_l[p] = l;
            } else {
                float4 tl = next.tl[slot], lr = next.lr[slot];
                tl.w = tl.w + c.x;
                lr.x = lr.x + c.y;
                lr.y = lr.y + c.z;
                next.tl[slot] = tl;
                next.lr[slot] = lr;
            }
        }
    }
    if (kCount) {
        atomicAdd(stat + 0, (unsigned long long)st.nodes);
        atomicAdd(stat + 1, (unsigned long long)st.tris);
    }
}

// --- shading -----------------------------------------------------------------------
struct AovBuffers {
    float *albedo;     // 3 per pixel
    float *normal;     // 3 per pixel
    float *test;       // 1 per pixel
    uint32_t *prim;    // 1 per pixel
    uint32_t last_frame;  // AOVs are kept from this frame of the launch
    uint32_t npix;
};

// Per-view records of a batched render (k_raygen_batch, k_shade,
// k_accumulate_batch).
struct BatchViewDev {
    CameraRec camera;
    EnvRec env;
    uint32_t width, height, max_depth, root;
    uint32_t emitter_begin, emitter_count;
    uint32_t guide_off;  // guided batches: the view's guide was gated off (its paths sample the BSDF)
    uint32_t pad1;
    float4 *accum, *frame;
    double *accum64;
    void *pad2;
};

// With `views` set (batched render) the path id carries the view above bit
// `path_bits`; the view selects its emitters, environment and depth.
// kFeat: the scene features compiled in (kFeat*): 0 for triangle scenes with
// RGB textures and no environment map, kFeatAll otherwise.
template <bool kBatch, uint32_t kFeat>
__global__ __launch_bounds__(256, kBatch ? 1 : 5) void k_shade(SceneView scene, ShadeParams prm, PathQueue in,
                                               const uint32_t *count_in, PathQueue out, uint32_t *count_out,
                                               ShadowQueue sq, uint32_t *count_shadow, float4 *out_l,
                                               AovBuffers aov, const BatchViewDev *views, uint32_t path_bits) {
    const uint32_t count = *count_in;
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t base = blockIdx.x * blockDim.x; base < count; base += stride) {
        const uint32_t i = base + threadIdx.x;
        const bool valid = i < count;
        PathState ps;
        ShadeOut so;
        so.cont = so.done = so.shadow = so.wrote_aov = false;
        uint32_t path = 0, packed = 0, vbits = 0, prim = kMissPrim;
        if (valid) {
            const float4 r0 = in.ray0[i], r1 = in.ray1[i], tl = in.tl[i], lr = in.lr[i], h = in.hit[i];
            ps.o = f3of(r0);
            packed = path = f2u(r0.w);
            ps.d = f3of(r1);
            ps.prev_pdf = r1.w;
            ps.T = f3of(tl);
            ps.L = mk3(tl.w, lr.x, lr.y);
            ps.rng.seed = f2u(lr.z);
            ps.flags = f2u(lr.w);
            prim = f2u(h.y);
            // AOVs of the launch's last frame, stored where shading finds them.
            const auto aov_write = [&](f3 albedo, f3 normal, float test) {
                if (!aov.albedo) return;
                const uint32_t f = path / aov.npix;
                if (f != aov.last_frame) return;
                const uint32_t px = path - f * aov.npix;
                aov.albedo[3 * px + 0] = albedo.x;
                aov.albedo[3 * px + 1] = albedo.y;
                aov.albedo[3 * px + 2] = albedo.z;
                aov.normal[3 * px + 0] = normal.x;
                aov.normal[3 * px + 1] = normal.y;
                aov.normal[3 * px + 2] = normal.z;
                aov.test[px] = test;
                aov.prim[px] = prim;
            };
            if constexpr (kBatch) {
                const uint32_t view = packed >> path_bits;
                path = packed & ((1u << path_bits) - 1u);
                vbits = view << path_bits;
                const BatchViewDev &bv = views[view];
                SceneView sv = scene;
                sv.emitters += bv.emitter_begin;
                sv.emitter_cdf += bv.emitter_begin;
                sv.num_emitters = bv.emitter_count;
                sv.env = bv.env;
                ShadeParams pv = prm;
                pv.max_depth = bv.max_depth;
                shade_path(sv, pv, ps, prim, h.x, h.z, h.w, so, kFeat, aov_write);
            } else {
                shade_path(scene, prm, ps, prim, h.x, h.z, h.w, so, kFeat, aov_write);
            }
        }
        const uint32_t j = wave_alloc(valid && so.cont, count_out);
        const uint32_t k = wave_alloc(valid && so.shadow, count_shadow);
        if (!valid) continue;
        if (so.cont) {
            out.ray0[j] = make_float4(ps.o.x, ps.o.y, ps.o.z, u2f(packed));
            out.ray1[j] = make_float4(ps.d.x, ps.d.y, ps.d.z, ps.prev_pdf);
            out.tl[j] = make_float4(ps.T.x, ps.T.y, ps.T.z, ps.L.x);
            out.lr[j] = make_float4(ps.L.y, ps.L.z, u2f(ps.rng.seed), u2f(ps.flags));
        } else {
            out_l[path] = make_float4(ps.L.x, ps.L.y, ps.L.z, 1.f);
        }
        if (so.shadow) {
            const uint32_t target = so.cont ? (vbits | j) : (0x80000000u | vbits | path);
            sq.s0[k] = make_float4(so.so.x, so.so.y, so.so.z, so.stmax);
            sq.s1[k] = make_float4(so.sd.x, so.sd.y, so.sd.z, u2f(target));
            sq.s2[k] = make_float4(so.contrib.x, so.contrib.y, so.contrib.z, 0.f);
        }
    }
}

// --- guided shading ----------------------------------------------------------------
struct GuideShadeArgs {
    const uint4 *gout;  // network outputs per queue slot (48 halves, 41 used); null: no guide sampling
    RecordSink rs;      // rs.rec null: no training records
    float fixed_c;      // > 0: selection probability used instead of the learned one
};

__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ uint32_t pack_half2(float a, float b) {
    union {
        _Float16 h[2];
        uint32_t u;
    } c;
    c.h[0] = (_Float16)a;
    c.h[1] = (_Float16)b;
    return c.u;
}

__device__ __forceinline__ void unpack_half2(uint32_t v, float &a, float &b) {
    union {
        _Float16 h[2];
        uint32_t u;
    } c;
    c.u = v;
    a = (float)c.h[0];
    b = (float)c.h[1];
}

// Network inference for every hit in the extension queue, ahead of shading.
// The block keeps the pre-permuted FP16 weight operands in LDS (32 KB); each
// wave takes 64 queue slots at a time as four 16-query MFMA tiles, queries the
// hit point o + t d, and stores the 41 outputs (FP16, 48 halves = 96 bytes
// per slot) straight from the MFMA result registers: lane (g, i) of a tile
// holds outputs 16 rt + 4 g .. +3 of query i.  Waves stride over the queue;
// a wave with no hit skips the network.  In a batched render the path id's
// view bits (above `path_bits`) select the query's hash grid.  Only hits on
// materials the guided step samples are queried (misses and delta / smooth
// plastic surfaces gather no table entries; their slots are left unwritten).
constexpr int kInferBlock = 512;

// Scene records k_guide_infer needs to tell a guidable hit.
struct InferScene {
    const uint32_t *tri;  // 4 per triangle: i0, i1, i2, shape
    const ShapeRec *shapes;
    const SphereRec *spheres;
    const MaterialRec *materials;
};

__global__ __launch_bounds__(kInferBlock) void k_guide_infer(GuideView G, PathQueue q, const uint32_t *count_ptr,
                                                             uint2 *gout, unsigned long long *queries,
                                                             uint32_t path_bits, InferScene sc) {
    __shared__ half8 wl[4 * 4 * 2 * 64];
    const uint32_t waves = kInferBlock / 64;
    if (blockIdx.x * waves * 64u >= *count_ptr) return;  // no slot for this block: skip the weight load
    for (uint32_t t = threadIdx.x; t < 4 * 4 * 2 * 64; t += kInferBlock) wl[t] = G.fwd_a[t];
    __syncthreads();
    const uint32_t wave = threadIdx.x >> 6;
    const int lane = (int)lane_id();
    const int g = lane >> 4, il = lane & 15;
    const uint32_t count = *count_ptr;
    const uint32_t wstride = gridDim.x * waves * 64u;
    uint32_t nq = 0;  // queries of this wave (counted by lane 0)
    for (uint32_t base = (blockIdx.x * waves + wave) * 64u; base < count; base += wstride) {
        const uint32_t i = base + lane;
        bool valid = i < count;
        f3 p = mk3(0.f);
        uint32_t grid = 0u;
        if (valid) {
            const float4 r0 = q.ray0[i], r1 = q.ray1[i], h = q.hit[i];
            const uint32_t pflags = f2u(reinterpret_cast<const float *>(q.lr + i)[3]);
            const uint32_t prim = f2u(h.y);
            valid = prim != kMissPrim && !(pflags & kPathNoGuide);
            if (valid) {
                const uint32_t shape =
                    (prim & kPrimSphere) ? sc.spheres[prim & kPrimIndexMask].shape : sc.tri[4 * (size_t)prim + 3];
                valid = guidable_material(sc.materials[sc.shapes[shape].material].type);
                p = f3of(r0) + f3of(r1) * h.x;
            }
            grid = min(f2u(r0.w) >> path_bits, G.grids - 1u);
        }
        const unsigned long long vmask = __ballot(valid);
        if (vmask == 0ull) continue;
        nq += (uint32_t)__popcll(vmask);
        const f3 u = guide_grid_pos(G, grid, p);
        const int vbit = valid ? 1 : 0;
#pragma unroll 1
        for (int qt = 0; qt < 4; ++qt) {
            // Keep the weight operands in LDS: re-read them per tile instead
            // of letting them be hoisted into (and occupy) 128 VGPRs.
            asm volatile("" ::: "memory");
            const int qi = qt * 16 + il;
            const f3 uq = mk3(__shfl(u.x, qi), __shfl(u.y, qi), __shfl(u.z, qi));
            const uint32_t gq = (uint32_t)__shfl((int)grid, qi);
            const bool vq = __shfl(vbit, qi) != 0;
            half8 in0 = {};
            if (vq) in0 = guide_encode(G, uq, g, gq);  // no table gathers for an unused query
            floatx4 out[4];
            guide_mlp_tile_shared(wl, lane, in0, out);
            if (vq) {
                uint2 *dst = gout + (size_t)(base + qi) * 12u + (uint32_t)g;
#pragma unroll
                for (int rt = 0; rt < 3; ++rt)
                    dst[rt * 4] = make_uint2(pack_half2(out[rt][0], out[rt][1]), pack_half2(out[rt][2], out[rt][3]));
            }
        }
    }
    if (lane == 0 && nq) atomicAdd(queries, (unsigned long long)nq);
}

// k_shade with the guided step; the mixture of each guidable vertex comes
// from k_guide_infer's outputs for the same queue slot.  With `views` set
// (batched render) the path id carries the view above bit `path_bits`, as in
// k_shade, the view is also the hash grid its records name, and the launch's
// record epoch and phase come from `ld` (so a captured launch replays with
// the values of each launch).
template <bool kBatch>
__global__ __launch_bounds__(256) void k_shade_guided(SceneView scene, ShadeParams prm, PathQueue in,
                                                      const uint32_t *count_in, PathQueue out, uint32_t *count_out,
                                                      ShadowQueue sq, uint32_t *count_shadow, float4 *out_l,
                                                      AovBuffers aov, GuideShadeArgs ga, const BatchViewDev *views,
                                                      uint32_t path_bits, const LaunchDesc *ld) {
    const uint32_t count = *count_in;
    const uint32_t stride = gridDim.x * blockDim.x;
    RecordSink rs = ga.rs;
    if constexpr (kBatch) {
        rs.epoch = ld->epoch;
        rs.phase = ld->phase;
    }
    for (uint32_t base = blockIdx.x * blockDim.x; base < count; base += stride) {
        const uint32_t i = base + threadIdx.x;
        const bool valid = i < count;
        PathState ps;
        ShadeOut so;
        so.cont = so.done = so.shadow = so.wrote_aov = false;
        GuidedVertex v;
        v.guidable = false;
        uint32_t path = 0, packed = 0, vbits = 0, view = 0, prim = kMissPrim;
        bool scatter = false;
        SceneView sv = scene;
        if (valid) {
            const float4 r0 = in.ray0[i], r1 = in.ray1[i], tl = in.tl[i], lr = in.lr[i], h = in.hit[i];
            ps.o = f3of(r0);
            packed = path = f2u(r0.w);
            ps.d = f3of(r1);
            ps.prev_pdf = r1.w;
            ps.T = f3of(tl);
            ps.L = mk3(tl.w, lr.x, lr.y);
            ps.rng.seed = f2u(lr.z);
            ps.flags = f2u(lr.w);
            prim = f2u(h.y);
            ShadeParams pv = prm;
            if constexpr (kBatch) {
                view = packed >> path_bits;
                path = packed & ((1u << path_bits) - 1u);
                vbits = view << path_bits;
                const BatchViewDev &bv = views[view];
                sv.emitters += bv.emitter_begin;
                sv.emitter_cdf += bv.emitter_begin;
                sv.num_emitters = bv.emitter_count;
                sv.env = bv.env;
                pv.max_depth = bv.max_depth;
            }
            scatter = guided_begin(sv, pv, ps, prim, h.x, h.z, h.w, rs, path, v, so);
        }
        bool use_guide = ga.gout != nullptr && scatter && v.guidable && !(ps.flags & kPathNoGuide);
        GuideMixture m;
        if (use_guide) {
            float o[48];
#pragma unroll
            for (int k = 0; k < 6; ++k) {
                const uint4 u = ga.gout[6 * (size_t)i + k];
                unpack_half2(u.x, o[8 * k], o[8 * k + 1]);
                unpack_half2(u.y, o[8 * k + 2], o[8 * k + 3]);
                unpack_half2(u.z, o[8 * k + 4], o[8 * k + 5]);
                unpack_half2(u.w, o[8 * k + 6], o[8 * k + 7]);
            }
            guide_decode(o, m);
            if (ga.fixed_c > 0.f) m.c = ga.fixed_c;
            if (!(m.c >= 0.f && m.c <= 1.f) || !(m.wsum > 0.f) || !isfinite(m.wsum)) use_guide = false;
        }
        if (scatter) guided_scatter(sv, ps, v, use_guide, m, rs, path, so, view);
        const uint32_t j = wave_alloc(valid && so.cont, count_out);
        const uint32_t k = wave_alloc(valid && so.shadow, count_shadow);
        if (!valid) continue;
        if (so.wrote_aov && aov.albedo) {
            const uint32_t f = path / aov.npix;
            if (f == aov.last_frame) {
                const uint32_t px = path - f * aov.npix;
                aov.albedo[3 * px + 0] = so.albedo.x;
                aov.albedo[3 * px + 1] = so.albedo.y;
                aov.albedo[3 * px + 2] = so.albedo.z;
                aov.normal[3 * px + 0] = so.normal.x;
                aov.normal[3 * px + 1] = so.normal.y;
                aov.normal[3 * px + 2] = so.normal.z;
                aov.test[px] = so.test;
                aov.prim[px] = prim;
            }
        }
        if (so.cont) {
            out.ray0[j] = make_float4(ps.o.x, ps.o.y, ps.o.z, u2f(packed));
            out.ray1[j] = make_float4(ps.d.x, ps.d.y, ps.d.z, ps.prev_pdf);
            out.tl[j] = make_float4(ps.T.x, ps.T.y, ps.T.z, ps.L.x);
            out.lr[j] = make_float4(ps.L.y, ps.L.z, u2f(ps.rng.seed), u2f(ps.flags));
        } else {
            out_l[path] = make_float4(ps.L.x, ps.L.y, ps.L.z, 1.f);
        }
        if (so.shadow) {
            const uint32_t target = so.cont ? (vbits | j) : (0x80000000u | vbits | path);
            sq.s0[k] = make_float4(so.so.x, so.so.y, so.so.z, so.stmax);
            sq.s1[k] = make_float4(so.sd.x, so.sd.y, so.sd.z, u2f(target));
            sq.s2[k] = make_float4(so.contrib.x, so.contrib.y, so.contrib.z, 0.f);
        }
    }
}

// --- accumulation ---------------------------------------------------------------
// main.cu:190-196: running mean acc = lerp(acc, L, 1 / (n + 1)) per frame.
__global__ void k_accumulate(const float4 *out_l, float4 *accum, float4 *frame, double *accum64, uint32_t npix,
                             uint32_t frames, uint32_t sample_cnt0, uint32_t accumulate) {
    for (uint32_t p = blockIdx.x * blockDim.x + threadIdx.x; p < npix; p += gridDim.x * blockDim.x) {
        float4 acc = accum[p];
        f3 a = mk3(acc.x, acc.y, acc.z);
        uint32_t n = sample_cnt0;
        double sx = 0, sy = 0, sz = 0;
        if (accum64) {
            sx = accum64[3 * (size_t)p];
            sy = accum64[3 * (size_t)p + 1];
            sz = accum64[3 * (size_t)p + 2];
        }
        for (uint32_t f = 0; f < frames; ++f) {
            const float4 lv = out_l[(size_t)f * npix + p];
            f3 L = mk3(lv.x, lv.y, lv.z);
            if (accumulate && n > 0) {
                const float t = 1.f / ((float)n + 1.f);
                L = lerp(a, L, t);
            }
            a = L;
            if (accumulate) n += 1;
            sx += (double)lv.x;
            sy += (double)lv.y;
            sz += (double)lv.z;
        }
        const float4 r = make_float4(a.x, a.y, a.z, 1.f);
        accum[p] = r;
        frame[p] = r;
        if (accum64) {
            accum64[3 * (size_t)p] = sx;
            accum64[3 * (size_t)p + 1] = sy;
            accum64[3 * (size_t)p + 2] = sz;
        }
    }
}

// --- batched views ----------------------------------------------------------------
// main.cu:39-78 for every (view, frame, pixel) of a batched launch.
// Chain `chain` of `chains` takes the launch's slots [lo, hi) with
// lo = total * chain / chains, into its own queue part.
__global__ void k_raygen_batch(PathQueue q, const LaunchDesc *ld, const BatchViewDev *views, uint32_t path_bits,
                               uint32_t *counters, uint32_t num_counters, uint32_t chain, uint32_t chains) {
    const uint32_t all = ld->total_paths;
    const uint32_t lo = batch_chain_begin(all, chain, chains);
    const uint32_t total = batch_chain_begin(all, chain + 1, chains) - lo;
    const uint32_t ns = ld->n_segs;
    const uint32_t tid = blockIdx.x * blockDim.x + threadIdx.x;
    if (tid == 0) {
        counters[0] = total;
        for (uint32_t i = 1; i < num_counters; ++i) counters[i] = 0u;
    }
    for (uint32_t i = tid; i < total; i += gridDim.x * blockDim.x) {
        // The segment holding launch slot g; within a segment (consecutive
        // frames of one view) frames are interleaved as in k_raygen.
        const uint32_t g = lo + i;
        uint32_t a = 0, z = ns - 1;
        while (a < z) {
            const uint32_t mid = (a + z + 1u) >> 1;
            if (ld->e[ld->g[mid].first_entry].path_begin <= g) a = mid;
            else z = mid - 1u;
        }
        const LaunchSeg &sg = ld->g[a];
        const uint32_t local = g - ld->e[sg.first_entry].path_begin;
        const LaunchEntry en = ld->e[sg.first_entry + local % sg.frames];
        const uint32_t view = en.view & kEntryViewMask;
        const BatchViewDev &v = views[view];
        const uint32_t pixel = tiled_pixel(local / sg.frames, v.width, v.height);
        const uint32_t x = pixel % v.width, y = pixel / v.width;
        Rng rng = rng_init(pixel, en.frame_seed);
        const float jx = rng_next(rng);
        const float jy = rng_next(rng);
        f3 o, d;
        camera_ray(v.camera, x, y, v.width, v.height, jx, jy, o, d);
        uint32_t flags = kPathPrimary | ((en.view & kEntryTrain) ? 0u : kPathNoTrain);
        if ((en.view & kEntryPlain) || v.guide_off) flags |= kPathNoGuide;
        q.ray0[i] = make_float4(o.x, o.y, o.z, u2f((view << path_bits) | (en.path_begin + pixel)));
        q.ray1[i] = make_float4(d.x, d.y, d.z, 0.f);
        q.tl[i] = make_float4(1.f, 1.f, 1.f, 0.f);
        q.lr[i] = make_float4(0.f, 0.f, u2f(rng.seed), u2f(flags));
    }
}

// k_accumulate per segment: the running mean of main.cu:190-196 over the
// segment's frames in order, into its view's film.
__global__ void k_accumulate_batch(const float4 *out_l, const LaunchDesc *ld, const BatchViewDev *views,
                                   uint32_t fp64) {
    const uint32_t total = ld->seg_pixels;
    const uint32_t ns = ld->n_segs;
    for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < total; t += gridDim.x * blockDim.x) {
        uint32_t lo = 0, hi = ns - 1;
        while (lo < hi) {
            const uint32_t mid = (lo + hi + 1u) >> 1;
            if (ld->g[mid].pix_begin <= t) lo = mid;
            else hi = mid - 1u;
        }
        const LaunchSeg sg = ld->g[lo];
        const BatchViewDev &v = views[sg.view];
        const uint32_t p = t - sg.pix_begin;
        const float4 acc = v.accum[p];
        f3 a = mk3(acc.x, acc.y, acc.z);
        const bool restart = sg.restart && !v.guide_off;
        uint32_t n = restart ? sg.restart_cnt0 : sg.sample_cnt0;
        double sx = 0, sy = 0, sz = 0;
        if (fp64 && !(restart && n == 0)) {
            sx = v.accum64[3 * (size_t)p];
            sy = v.accum64[3 * (size_t)p + 1];
            sz = v.accum64[3 * (size_t)p + 2];
        }
        for (uint32_t f = 0; f < sg.frames; ++f) {
            const float4 lv = out_l[ld->e[sg.first_entry + f].path_begin + p];
            f3 L = mk3(lv.x, lv.y, lv.z);
            if (sg.accumulate && n > 0) {
                const float w = 1.f / ((float)n + 1.f);
                L = lerp(a, L, w);
            }
            a = L;
            if (sg.accumulate) n += 1;
            sx += (double)lv.x;
            sy += (double)lv.y;
            sz += (double)lv.z;
        }
        const float4 r = make_float4(a.x, a.y, a.z, 1.f);
        v.accum[p] = r;
        v.frame[p] = r;
        if (fp64) {
            v.accum64[3 * (size_t)p] = sx;
            v.accum64[3 * (size_t)p + 1] = sy;
            v.accum64[3 * (size_t)p + 2] = sz;
        }
    }
}

// Adds one launch's queue sizes to running ray counts (graph-capturable; no
// host read-back per launch).
// Guide gate accumulators: per view, per kind (guided, plain), per frame slot
// (frame index mod kGateSlots), the frame's error and a frame count.
constexpr uint32_t kGateSlots = 128;
constexpr size_t kGateDoublesPerView = 2 * kGateSlots * 2;
__host__ __device__ __forceinline__ size_t gate_slot(uint32_t view, uint32_t plain, uint32_t slot) {
    return (((size_t)view * 2 + plain) * kGateSlots + slot) * 2;
}

// Per-frame error of the evaluation frames of a guided batch's launch: for
// each frame flagged kEntryEval, the mean over pixels and channels of
// (L - mu)^2 / (mu^2 + 0.01) -- the relMSE one frame would have against the
// view's mean -- goes to the frame's slot of its view and kind.
// mu is the leave-one-out mean of the view's N frames so far,
// (N mean - L) / (N - 1): the mean after this launch includes the frame
// itself, which would shrink a firefly's deviation to at most (N - 1)^2 of
// its own share and hide exactly the variance that guiding removes.
// blockIdx.y = launch entry.
__global__ __launch_bounds__(256) void k_eval_error(const float4 *out_l, const LaunchDesc *ld,
                                                    const BatchViewDev *views, double *acc) {
    __shared__ float part[4];
    __shared__ uint32_t frames_so_far;
    const uint32_t en = blockIdx.y;
    if (en >= ld->n_entries) return;
    const LaunchEntry e = ld->e[en];
    if (!(e.view & kEntryEval)) return;
    if (threadIdx.x == 0) {
        uint32_t n = 0;
        for (uint32_t k = 0; k < ld->n_segs; ++k) {
            const LaunchSeg &sg = ld->g[k];
            if (en >= sg.first_entry && en < sg.first_entry + sg.frames) n = sg.sample_cnt0 + sg.frames;
        }
        frames_so_far = n;
    }
    __syncthreads();
    const uint32_t n = frames_so_far;
    if (n < 2u) return;
    const float fn = (float)n, inv = 1.f / (float)(n - 1u);
    const uint32_t v = e.view & kEntryViewMask;
    const float4 *mu = views[v].accum;
    float sum = 0.f;
    for (uint32_t p = blockIdx.x * blockDim.x + threadIdx.x; p < e.pixels; p += gridDim.x * blockDim.x) {
        const float4 l = out_l[e.path_begin + p], ma = mu[p];
        const float4 m = make_float4((fn * ma.x - l.x) * inv, (fn * ma.y - l.y) * inv, (fn * ma.z - l.z) * inv, 0.f);
        const float dx = l.x - m.x, dy = l.y - m.y, dz = l.z - m.z;
        sum += (dx * dx / (m.x * m.x + 0.01f) + dy * dy / (m.y * m.y + 0.01f) + dz * dz / (m.z * m.z + 0.01f)) *
               (1.f / 3.f);
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) sum += __shfl_xor(sum, off);
    if (lane_id() == 0) part[threadIdx.x >> 6] = sum;
    __syncthreads();
    if (threadIdx.x == 0) {
        const double t = (double)part[0] + (double)part[1] + (double)part[2] + (double)part[3];
        const size_t k = gate_slot(v, (e.view & kEntryPlain) ? 1u : 0u, e.frame_seed % kGateSlots);
        atomicAdd(acc + k, t / (double)e.pixels);
        if (blockIdx.x == 0) atomicAdd(acc + k + 1, 1.0);
    }
}

// Mean and median of the per-frame errors of one view and kind; false
// without frames.
__device__ bool gate_stats(const double *acc, uint32_t view, uint32_t plain, double &mean, double &median) {
    double x[kGateSlots];
    uint32_t n = 0;
    for (uint32_t s = 0; s < kGateSlots; ++s) {
        const size_t k = gate_slot(view, plain, s);
        if (acc[k + 1] > 0.0) {
            double val = acc[k] / acc[k + 1];
            uint32_t j = n++;  // insertion sort
            while (j > 0 && x[j - 1] > val) {
                x[j] = x[j - 1];
                --j;
            }
            x[j] = val;
        }
    }
    if (n == 0) return false;
    double sum = 0.0;
    for (uint32_t i = 0; i < n; ++i) sum += x[i];
    mean = sum / n;
    median = (n & 1u) ? x[n / 2] : 0.5 * (x[n / 2 - 1] + x[n / 2]);
    return true;
}

// Gate of a guided batch, for each view whose last evaluation frame is in
// this launch (kEntryDecide): the view's remaining frames sample the BSDF
// only when neither the mean nor the median of its guided evaluation frames'
// errors is below 1/kGateGain of its plain ones.  The median keeps one
// heavy-tailed frame from deciding; the mean keeps the tail in the picture.
// The bar is low because the decision comes at half the training, while the
// guide is still learning: at 1080p, in batches of two or three door-room
// views, their guided frames 8..15 were 1.2-1.5x better in mean (and later
// frames 5-10x), while terrain views, lit directly, read 0.88-1.08x.  Views
// without evaluation frames keep the guide.  One thread per launch entry.
constexpr double kGateGain = 1.1;
__global__ void k_guide_gate(BatchViewDev *views, const LaunchDesc *ld, double *acc, uint32_t *gated_off) {
    const uint32_t en = blockIdx.x * blockDim.x + threadIdx.x;
    if (en >= ld->n_entries || !(ld->e[en].view & kEntryDecide)) return;
    const uint32_t v = ld->e[en].view & kEntryViewMask;
    double gm, gmed, pm, pmed;
    const bool off = gate_stats(acc, v, 0u, gm, gmed) && gate_stats(acc, v, 1u, pm, pmed) &&
                     gm * kGateGain >= pm && gmed * kGateGain >= pmed;
    views[v].guide_off = off ? 1u : 0u;
    if (off) atomicAdd(gated_off, 1u);
}

__global__ void k_guide_gate_reset(BatchViewDev *views, uint32_t n, double *acc, uint32_t *gated_off) {
    for (size_t i = blockIdx.x * blockDim.x + threadIdx.x; i < (size_t)n * kGateDoublesPerView;
         i += gridDim.x * blockDim.x) {
        acc[i] = 0.0;
        if (i % kGateDoublesPerView == 0) views[i / kGateDoublesPerView].guide_off = 0u;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) *gated_off = 0u;
}

__global__ void k_batch_count(const uint32_t *counters, uint32_t max_depth, unsigned long long *rays) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    unsigned long long e = 0, sh = 0;
    for (uint32_t b = 0; b < max_depth; ++b) {
        e += counters[b];
        sh += counters[max_depth + 2 + b];
    }
    rays[0] += e;
    rays[1] += sh;
}

// Streaming copy used to measure achievable HBM bandwidth and to calibrate
// hardware byte counters (pt_bandwidth_probe).
__global__ void k_copy(const float4 *__restrict__ src, float4 *__restrict__ dst, size_t n) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

// Closest-hit / any-hit over caller rays (pt_trace_rays).
template <bool kAnyHit, bool kSph>
__global__ __launch_bounds__(kTraceBlock) void k_trace_rays(const float *rays, uint32_t count, DeviceBvh bvh,
                                                            uint4 *hits) {
    __shared__ uint2 lds[kLdsStack * kTraceBlock];
    const uint32_t gtid = blockIdx.x * blockDim.x + threadIdx.x;
    TraverseStats st{0, 0};
    for (uint32_t i = gtid; i < count; i += gridDim.x * blockDim.x) {
        const float *r = rays + 8 * (size_t)i;
        Hit h;
        trace_one<kAnyHit, false, kSph>(bvh, lds + threadIdx.x, gtid, mk3(r[0], r[1], r[2]), mk3(r[4], r[5], r[6]),
                                        r[3], r[7], h, st);
        hits[i] = make_uint4(f2u(h.t), h.prim, f2u(h.b1), f2u(h.b2));
    }
}

// Traversal variants: counting on/off x sphere leaves present/absent (a
// triangle-only tree runs without the sphere test, like an OptiX pipeline
// built without the sphere primitive flag, optix/pipeline.cpp:14-28).
using TraceKernel = void (*)(PathQueue, DeviceBvh, const uint32_t *, uint32_t *, unsigned long long *,
                             const uint32_t *, uint32_t);
using ShadowKernel = void (*)(ShadowQueue, DeviceBvh, const uint32_t *, uint32_t *, PathQueue, float4 *,
                              unsigned long long *, const uint32_t *, uint32_t);
using TraceRaysKernel = void (*)(const float *, uint32_t, DeviceBvh, uint4 *);
TraceKernel trace_kernel(bool count, bool sph) {
    return count ? (sph ? k_trace<true, true> : k_trace<true, false>)
                 : (sph ? k_trace<false, true> : k_trace<false, false>);
}
ShadowKernel shadow_kernel(bool count, bool sph) {
    return count ? (sph ? k_shadow<true, true> : k_shadow<true, false>)
                 : (sph ? k_shadow<false, true> : k_shadow<false, false>);
}
using ShadeKernel = void (*)(SceneView, ShadeParams, PathQueue, const uint32_t *, PathQueue, uint32_t *, ShadowQueue,
                            uint32_t *, float4 *, AovBuffers, const BatchViewDev *, uint32_t);
// The shading kernel for a scene using the kFeat* features `feat`.
ShadeKernel shade_kernel(bool batch, uint32_t feat) {
    if (batch) return feat ? k_shade<true, kFeatAll> : k_shade<true, 0u>;
    return feat ? k_shade<false, kFeatAll> : k_shade<false, 0u>;
}
TraceRaysKernel trace_rays_kernel(bool any_hit, bool sph) {
    return any_hit ? (sph ? k_trace_rays<true, true> : k_trace_rays<true, false>)
                   : (sph ? k_trace_rays<false, true> : k_trace_rays<false, false>);
}

// --- BVH refit after shape updates ------------------------------------------------
// Leaf records copy their primitive's current geometry (the primitive id and
// the sphere tag stay).
__global__ void k_refit_leaves(LeafTri *leaf, uint32_t n, const uint32_t *tri, const float4 *pos,
                               const float4 *sph4) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        float4 *p = reinterpret_cast<float4 *>(leaf + i);
        const float4 a = p[0], b = p[1];
        const uint32_t prim = f2u(a.w);
        if (f2u(b.w) != 0u) {
            const float4 q = sph4[prim & kPrimIndexMask];
            p[0] = make_float4(q.x, q.y, q.z, a.w);
            p[1] = make_float4(q.w, 0.f, 0.f, b.w);
        } else {
            const uint4 t = *reinterpret_cast<const uint4 *>(tri + 4 * (size_t)prim);
            const float4 p0 = pos[t.x], p1 = pos[t.y], p2 = pos[t.z];
            p[0] = make_float4(p0.x, p0.y, p0.z, a.w);
            p[1] = make_float4(p1.x, p1.y, p1.z, 0.f);
            p[2] = make_float4(p2.x, p2.y, p2.z, 0.f);
        }
    }
}

// One tree level, bottom-up: each node's children boxes (child nodes from
// the level below, leaves from their primitives), the node box, and the
// quantized child boxes against the new box (the builder's rule).  Boxes
// are kept per node for the parents; `failed` is set when an axis cannot be
// quantized.
__global__ void k_refit_nodes(WideNode *nodes, float4 *boxes, const LeafTri *leaf, const uint32_t *list,
                              uint32_t count, uint32_t *failed) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= count) return;
    const uint32_t idx = list[t];
    WideNode n;
    load_node_dev(nodes, idx, n);
    float clo[3][8], chi[3][8];
    int slot[8];
    int ns = 0;
    float nlo[3] = {INFINITY, INFINITY, INFINITY}, nhi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int k = 0; k < 8; ++k) {
        const uint32_t meta = n.meta[k];
        if (meta == 0) continue;
        float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
        if (n.imask & (1u << k)) {
            const uint32_t c = n.child_base + __builtin_popcount((uint32_t)n.imask & ((1u << k) - 1u));
            const float4 a = boxes[2 * (size_t)c], b = boxes[2 * (size_t)c + 1];
            lo[0] = a.x; lo[1] = a.y; lo[2] = a.z;
            hi[0] = b.x; hi[1] = b.y; hi[2] = b.z;
        } else {
            const uint32_t cnt = __builtin_popcount(meta >> 5), off = meta & 31u;
            for (uint32_t j = 0; j < cnt; ++j) {
                const float4 *p = reinterpret_cast<const float4 *>(leaf + n.tri_base + off + j);
                const float4 a = p[0], b = p[1];
                if (f2u(b.w) != 0u) {
                    const float c3[3] = {a.x, a.y, a.z};
                    for (int ax = 0; ax < 3; ++ax) {
                        lo[ax] = fminf(lo[ax], c3[ax] - b.x);
                        hi[ax] = fmaxf(hi[ax], c3[ax] + b.x);
                    }
                } else {
                    const float4 c = p[2];
                    const float v[3][3] = {{a.x, a.y, a.z}, {b.x, b.y, b.z}, {c.x, c.y, c.z}};
                    for (int q = 0; q < 3; ++q)
                        for (int ax = 0; ax < 3; ++ax) {
                            lo[ax] = fminf(lo[ax], v[q][ax]);
                            hi[ax] = fmaxf(hi[ax], v[q][ax]);
                        }
                }
            }
        }
        for (int ax = 0; ax < 3; ++ax) {
            clo[ax][ns] = lo[ax];
            chi[ax][ns] = hi[ax];
            nlo[ax] = fminf(nlo[ax], lo[ax]);
            nhi[ax] = fmaxf(nhi[ax], hi[ax]);
        }
        slot[ns++] = k;
    }
    if (ns == 0) return;
    uint8_t e[3], ql[3][8], qh[3][8];
    const float np[3] = {nlo[0], nlo[1], nlo[2]};
    for (int ax = 0; ax < 3; ++ax)
        if (!quantize_node_axis(np[ax], nhi[ax], clo[ax], chi[ax], ns, e[ax], ql[ax], qh[ax])) {
            atomicOr(failed, 1u);
            return;
        }
    n.px = np[0];
    n.py = np[1];
    n.pz = np[2];
    n.ex = e[0];
    n.ey = e[1];
    n.ez = e[2];
    for (int j = 0; j < ns; ++j) {
        const int k = slot[j];
        n.qlo_x[k] = ql[0][j];
        n.qlo_y[k] = ql[1][j];
        n.qlo_z[k] = ql[2][j];
        n.qhi_x[k] = qh[0][j];
        n.qhi_y[k] = qh[1][j];
        n.qhi_z[k] = qh[2][j];
    }
    const uint4 *src = reinterpret_cast<const uint4 *>(&n);
    uint4 *dst = reinterpret_cast<uint4 *>(nodes + idx);
    for (int q = 0; q < 5; ++q) dst[q] = src[q];
    boxes[2 * (size_t)idx] = make_float4(nlo[0], nlo[1], nlo[2], 0.f);
    boxes[2 * (size_t)idx + 1] = make_float4(nhi[0], nhi[1], nhi[2], 0.f);
}

}  // namespace

// =====================================================================================
// Context
// =====================================================================================
// One view of a batched render (host side).
struct BatchView {
    uint32_t width = 0, height = 0, spp = 0, max_depth = 1;
    float lo[3]{}, hi[3]{};  // scene bounds (the view's guide hash grid covers them)
    CameraRec camera{};
    EnvRec env{};
    uint32_t root = 0, emitter_begin = 0, emitter_count = 0, tri_count = 0;
    size_t env_tables = 0;  // offset of this view's env-map tables in BatchState::env_tables
    float4 *d_accum = nullptr, *d_frame = nullptr;
    double *d_accum64 = nullptr;
    uint32_t sample_cnt = 0, random_seed = 0;  // progressive state, as pt_pass.cpp:55-56 per view
};

struct BatchState {
    std::vector<BatchView> views;
    // concatenated scene (global vertex / triangle / shape / material / emitter indices)
    std::vector<float> pos4, nrm4, uv2;
    std::vector<uint32_t> tri4;
    std::vector<ShapeRec> shapes;
    std::vector<SphereRec> spheres;
    std::vector<MaterialRec> materials;
    std::vector<EmitterRec> emitters;
    std::vector<float> cdf;
    std::vector<BitmapRec> bitmaps;
    std::vector<float> texels;
    std::vector<float> env_tables;
    std::vector<WideNode> nodes;
    std::vector<LeafTri> leaf;
    bool any_normals = false, any_texcoords = false;
    bool dirty = true;
    uint32_t feat = 0;  // kFeat* used by any view
    uint32_t *d_tri = nullptr;
    float4 *d_pos = nullptr, *d_nrm = nullptr;
    float2 *d_uv = nullptr;
    ShapeRec *d_shapes = nullptr;
    SphereRec *d_spheres = nullptr;
    MaterialRec *d_materials = nullptr;
    EmitterRec *d_emitters = nullptr;
    float *d_cdf = nullptr;
    BitmapRec *d_bitmaps = nullptr;
    float4 *d_texels = nullptr;
    float *d_env_tables = nullptr;
    WideNode *d_nodes = nullptr;
    LeafTri *d_leaf = nullptr;
    BatchViewDev *d_views = nullptr;
    uint32_t *d_roots = nullptr;
    LaunchDesc *d_launch = nullptr;
    LaunchDesc *h_launch = nullptr;  // pinned ring, one slot per launch of a render
    uint32_t h_launch_slots = 0;
    unsigned long long *d_rays = nullptr;  // extension, shadow
    double *d_gate_acc = nullptr;          // guide gate: 4 per view (guided / plain error sums and frames)
    uint32_t *d_gate_off = nullptr;        // guide gate: views switched off by the last gated render
    uint32_t gate_views = 0;
    bool tuned = false;
    pt_batch_tuning tuning{};
    pt_batch_stats stats{};
};

// Host facts of a shape for dynamic updates.
constexpr uint32_t kVertexUnused = 0xFFFFFFFFu, kVertexShared = 0xFFFFFFFEu;
struct ShapeHost {
    uint32_t flags0 = 0;        // ShapeRec flags at pt_set_scene
    uint32_t first = 0, count = 0;
    uint32_t vmin = 0, vmax = 0;  // vertex range used (triangle shapes)
    bool sphere = false, emitter = false, shared = false;
    TextureRec radiance{};
    float radiance_weight = 0.f;
};

struct pt_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    // Streams of the second to fourth chains of a launch (pt_render and
    // batched launches); dep_events fork them from and join them to the pass
    // stream.
    hipStream_t stream2 = nullptr, stream3 = nullptr, stream4 = nullptr;
    std::vector<hipEvent_t> dep_events;
    std::string error;
    int num_cu = 256;
    uint32_t trace_blocks_per_cu = 8;       // resident closest-hit traversal blocks per CU, sphere variant
    uint32_t trace_blocks_per_cu_tri = 8;   // triangle-only variant
    uint32_t shade_blocks_per_cu = 8;       // resident k_shade<false> blocks (256 threads) per CU
    uint32_t shade_guided_blocks_per_cu = 2;  // resident k_shade_guided<false> blocks per CU
    uint32_t shadow_blocks_per_cu = 8;      // resident any-hit (shadow) traversal blocks per CU, spheres
    uint32_t shadow_blocks_per_cu_tri = 8;  // triangle-only variant

    // scene (host copies kept for the BVH build)
    bool has_scene = false;
    bool has_bvh = false;
    std::vector<float> h_pos4, h_nrm4, h_uv2;
    std::vector<uint32_t> h_tri4;
    std::vector<ShapeRec> h_shapes;
    std::vector<SphereRec> h_spheres;
    std::vector<float> h_sph4;  // world center xyz + radius per sphere
    std::vector<MaterialRec> h_materials;
    std::vector<EmitterRec> h_emitters;
    std::vector<float> h_cdf;
    EnvRec env{};
    CameraRec camera{};
    uint32_t width = 0, height = 0, scene_max_depth = 1;
    uint32_t num_tris = 0, num_spheres = 0;
    uint32_t shade_feat = kFeatAll;  // kFeat* of the scene: picks the shading kernel
    // dynamic updates: geometry as given to pt_set_scene, per-shape facts
    std::vector<float> h_pos0, h_nrm0, h_sph_m0;  // 3 / 3 per vertex, 12 per sphere
    std::vector<ShapeHost> shape_host;
    std::vector<uint32_t> vert_owner;             // shape of each vertex, kVertexUnused / kVertexShared
    bool has_env = false;
    bool bvh_dirty = false;

    // device scene
    uint32_t *d_tri = nullptr;
    float4 *d_pos = nullptr, *d_nrm = nullptr;
    float2 *d_uv = nullptr;
    ShapeRec *d_shapes = nullptr;
    SphereRec *d_spheres = nullptr;
    MaterialRec *d_materials = nullptr;
    EmitterRec *d_emitters = nullptr;
    float *d_cdf = nullptr;
    BitmapRec *d_bitmaps = nullptr;
    float4 *d_texels = nullptr;
    float *d_env_tables = nullptr;
    WideNode *d_nodes = nullptr;
    LeafTri *d_leaf = nullptr;
    float4 *d_sph4 = nullptr;       // sphere center + radius (refit)
    uint32_t *d_levels = nullptr;   // node indices level by level (refit)
    float4 *d_node_box = nullptr;   // refit: per node box lo, hi
    uint32_t *d_refit_err = nullptr;
    std::vector<uint32_t> level_off;
    uint32_t num_nodes = 0, num_leaf = 0;
    BvhBuildStats bvh_stats{};

    // frame buffers
    float4 *d_accum = nullptr, *d_frame = nullptr;
    double *d_accum64 = nullptr;
    float *d_albedo = nullptr, *d_normal = nullptr, *d_test = nullptr;
    uint32_t *d_prim = nullptr;
    float4 *d_dn_scratch = nullptr;  // denoiser ping-pong, 2 w h
    size_t dn_scratch_px = 0;

    // queues
    uint32_t capacity = 0;  // paths
    PathQueue q[2]{};
    ShadowQueue sq{};
    float4 *d_out_l = nullptr;
    uint32_t *d_counters = nullptr;
    uint32_t num_counters = 0;
    uint2 *d_spill = nullptr;
    uint32_t spill_threads = 0;
    unsigned long long *d_stat = nullptr;  // closest nodes, closest tris, shadow nodes, shadow tris, guide queries
    bool counting = false;
    uint32_t *d_count_hist = nullptr;      // per-launch copies of the queue counters
    uint32_t count_hist_slots = 0;         // capacity in counters

    // kernel timing (PT_RENDER_TIMING)
    struct Pending {
        int kind;
        hipEvent_t a, b;
        uint64_t group;  // timers of one group time one logical launch
    };
    uint64_t timer_groups = 0;  // source of single-timer groups
    std::vector<hipEvent_t> event_pool;
    size_t event_next = 0;
    std::vector<Pending> pending;

    // progressive state (pt_pass.cpp:55-56)
    uint32_t sample_cnt = 0;
    uint32_t random_seed = 0;
    bool accum64_valid = false;

    pt_stats stats{};

    // learned guiding (pt_guide_*)
    float scene_lo[3]{}, scene_hi[3]{};
    GuideState guide;
    uint4 *d_gout = nullptr;  // network outputs per queue slot
    uint32_t gout_cap = 0;
    uint32_t infer_blocks_per_cu = 2;

    // batched views (pt_batch_*)
    BatchState batch;
};

namespace {

thread_local std::string g_create_error;

// Device memory of pt_render_noalloc: path queues, shadow queue, radiance
// slots, queue counters and the traversal stack spill area of one render call,
// carved out of a file-scope array instead of hipMalloc.  Each GPU that loads
// the library holds its own copy until the process ends.  In host code the
// array's name is not a device address: its per-device address comes from
// hipGetSymbolAddress (kept in g_scratch_ptr) and only that is handed to
// kernels.
constexpr size_t kNoAllocScratchBytes = 640ull << 20;
__device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];
constexpr int kMaxScratchDevices = 64;
std::mutex g_scratch_mutex;
unsigned char *g_scratch_ptr[kMaxScratchDevices] = {};

// Resolves this device's copy of `scratch` once (the device must be current).
hipError_t resolve_scratch(int device) {
    if (device < 0 || device >= kMaxScratchDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (g_scratch_ptr[device]) return hipSuccess;
    void *ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_scratch_ptr[device] = static_cast<unsigned char *>(ptr);
    return hipSuccess;
}

#define PT_HIP(call)                                                                         \
    do {                                                                                    \
        hipError_t e_ = (call);                                                             \
        if (e_ != hipSuccess) {                                                             \
            ctx->error = std::string(#call) + ": " + hipGetErrorString(e_);                 \
            return PT_ERR_HIP;                                                              \
        }                                                                                   \
    } while (0)

template <typename T>
void dfree(T *&p) {
    if (p) (void)hipFree((void *)p);
    p = nullptr;
}

void free_scene(pt_ctx *ctx) {
    dfree(ctx->d_tri);
    dfree(ctx->d_pos);
    dfree(ctx->d_nrm);
    dfree(ctx->d_uv);
    dfree(ctx->d_shapes);
    dfree(ctx->d_spheres);
    dfree(ctx->d_materials);
    dfree(ctx->d_emitters);
    dfree(ctx->d_cdf);
    dfree(ctx->d_bitmaps);
    dfree(ctx->d_texels);
    dfree(ctx->d_env_tables);
    dfree(ctx->d_nodes);
    dfree(ctx->d_leaf);
    dfree(ctx->d_sph4);
    dfree(ctx->d_levels);
    dfree(ctx->d_node_box);
    dfree(ctx->d_accum);
    dfree(ctx->d_frame);
    dfree(ctx->d_accum64);
    dfree(ctx->d_albedo);
    dfree(ctx->d_normal);
    dfree(ctx->d_test);
    dfree(ctx->d_prim);
    dfree(ctx->d_dn_scratch);
    ctx->dn_scratch_px = 0;
    ctx->has_scene = false;
    ctx->has_bvh = false;
}

void free_queues(pt_ctx *ctx) {
    for (auto &q : ctx->q) {
        dfree(q.ray0);
        dfree(q.ray1);
        dfree(q.tl);
        dfree(q.lr);
        dfree(q.hit);
    }
    dfree(ctx->sq.s0);
    dfree(ctx->sq.s1);
    dfree(ctx->sq.s2);
    dfree(ctx->d_out_l);
    dfree(ctx->d_counters);
    ctx->capacity = 0;
    ctx->num_counters = 0;
}

template <typename T>
int upload(pt_ctx *ctx, T *&dst, const T *src, size_t n) {
    dfree(dst);
    if (n == 0) return PT_OK;
    PT_HIP(hipMalloc((void **)&dst, n * sizeof(T)));
    PT_HIP(hipMemcpy(dst, src, n * sizeof(T), hipMemcpyHostToDevice));
    return PT_OK;
}

TextureRec to_texture_rec(const pt_texture_desc &t) {
    TextureRec r;
    std::memset(&r, 0, sizeof(r));
    r.type = t.type;
    r.bitmap = t.bitmap;
    r.c0 = f4{t.color0[0], t.color0[1], t.color0[2], 0.f};
    r.c1 = f4{t.color1[0], t.color1[1], t.color1[2], 0.f};
    r.r0 = f4{t.uv_row0[0], t.uv_row0[1], t.uv_row0[2], t.uv_row0[3]};
    r.r1 = f4{t.uv_row1[0], t.uv_row1[1], t.uv_row1[2], t.uv_row1[3]};
    return r;
}

// Average color of a texture (GetPixelAverage, optix_material.cpp:9-36); a
// bitmap averages its texels in row order.
f3 texture_average(const pt_texture_desc &t, const pt_bitmap_desc *bitmaps) {
    if (t.type == PT_TEX_CHECKER)
        return mk3(t.color0[0] + t.color1[0], t.color0[1] + t.color1[1], t.color0[2] + t.color1[2]) * 0.5f;
    if (t.type == PT_TEX_BITMAP) {
        const pt_bitmap_desc &b = bitmaps[t.bitmap];
        float r = 0.f, g = 0.f, bl = 0.f;
        const size_t n = (size_t)b.width * b.height;
        for (size_t i = 0; i < n; ++i) {
            r += b.rgba[4 * i];
            g += b.rgba[4 * i + 1];
            bl += b.rgba[4 * i + 2];
        }
        return mk3(r, g, bl) / (1.f * (float)b.height * (float)b.width);
    }
    return mk3(t.color0[0], t.color0[1], t.color0[2]);
}

// Emitter selection weight of a radiance texture (GetWeight, world/emitter.cpp:
// 73-101): the largest channel, averaged over a bitmap's texels (all texels;
// world/emitter.cpp's loop indexes a square image's texels).
float texture_weight(const pt_texture_desc &t, const pt_bitmap_desc *bitmaps) {
    auto mx = [](float r, float g, float b) { return r > g ? (r > b ? r : b) : (g > b ? g : b); };
    if (t.type == PT_TEX_CHECKER)
        return (mx(t.color0[0], t.color0[1], t.color0[2]) + mx(t.color1[0], t.color1[1], t.color1[2])) * 0.5f;
    if (t.type == PT_TEX_BITMAP) {
        const pt_bitmap_desc &b = bitmaps[t.bitmap];
        float w = 0.f;
        const size_t n = (size_t)b.width * b.height;
        for (size_t i = 0; i < n; ++i) w += mx(b.rgba[4 * i], b.rgba[4 * i + 1], b.rgba[4 * i + 2]);
        return w / (1.f * (float)b.width * (float)b.height);
    }
    return mx(t.color0[0], t.color0[1], t.color0[2]);
}

// optix_material.cpp:41-132 host packing.
MaterialRec to_material_rec(const pt_material_desc &m, const pt_bitmap_desc *bitmaps) {
    MaterialRec r;
    std::memset(&r, 0, sizeof(r));
    r.type = m.type;
    r.flags = (m.twosided ? kMatTwosided : 0u) | (m.nonlinear ? kMatNonlinear : 0u);
    for (int i = 0; i < 4; ++i) r.tex[i] = to_texture_rec(m.tex[i]);
    r.eta = m.int_ior / m.ext_ior;
    if (m.type == PT_MAT_PLASTIC || m.type == PT_MAT_ROUGH_PLASTIC) {
        float diffuse_lum = luminance(texture_average(m.tex[0], bitmaps));
        float specular_lum = luminance(texture_average(m.tex[1], bitmaps));
        r.spec_weight = specular_lum / (specular_lum + diffuse_lum);
        r.int_fdr = fresnel_diffuse_reflectance(1.f / r.eta);
    }
    return r;
}

int ensure_queues(pt_ctx *ctx, uint32_t paths, uint32_t max_depth) {
    uint32_t need_counters = 4 * (max_depth + 2);
    if (paths <= ctx->capacity && need_counters <= ctx->num_counters) return PT_OK;
    free_queues(ctx);
    const size_t n = paths;
    for (auto &q : ctx->q) {
        PT_HIP(hipMalloc((void **)&q.ray0, n * sizeof(float4)));
        PT_HIP(hipMalloc((void **)&q.ray1, n * sizeof(float4)));
        PT_HIP(hipMalloc((void **)&q.tl, n * sizeof(float4)));
        PT_HIP(hipMalloc((void **)&q.lr, n * sizeof(float4)));
        PT_HIP(hipMalloc((void **)&q.hit, n * sizeof(float4)));
    }
    PT_HIP(hipMalloc((void **)&ctx->sq.s0, n * sizeof(float4)));
    PT_HIP(hipMalloc((void **)&ctx->sq.s1, n * sizeof(float4)));
    PT_HIP(hipMalloc((void **)&ctx->sq.s2, n * sizeof(float4)));
    PT_HIP(hipMalloc((void **)&ctx->d_out_l, n * sizeof(float4)));
    // One set of counters per chain of a pt_render launch (four at most).
    PT_HIP(hipMalloc((void **)&ctx->d_counters, 4 * need_counters * sizeof(uint32_t)));
    ctx->capacity = paths;
    ctx->num_counters = need_counters;
    return PT_OK;
}

uint32_t trace_grid(pt_ctx *ctx) {
    return (uint32_t)ctx->num_cu * (ctx->num_spheres > 0 ? ctx->trace_blocks_per_cu : ctx->trace_blocks_per_cu_tri);
}
uint32_t shadow_grid(pt_ctx *ctx, bool spheres) {
    return (uint32_t)ctx->num_cu * (spheres ? ctx->shadow_blocks_per_cu : ctx->shadow_blocks_per_cu_tri);
}
// Shading grids.  k_shade<false>: kShadeRounds times the blocks that can be
// resident at once.  Every block strides over the queue; with many more
// blocks than slots the dispatcher balances uneven shading work across CUs,
// and blocks that find no work exit at once.  Terrain 1080p, 64 spp,
// k_shade<false> per render at 1, 4, 16 and 64 rounds: 49.2, 48.0, 46.9 and
// 46.3 ms.  k_shade_guided<false> uses the same rounds when it only samples;
// while training it keeps a grid of 8 blocks per CU (its training records are
// indexed by path and vertex, so the grid does not change them).
uint32_t shade_grid(pt_ctx *ctx, bool guided, bool training) {
    if (guided && training) return (uint32_t)ctx->num_cu * 8u;
    return (uint32_t)ctx->num_cu * kShadeRounds * (guided ? ctx->shade_guided_blocks_per_cu : ctx->shade_blocks_per_cu);
}

int ensure_spill(pt_ctx *ctx) {
    const uint32_t threads =
        (uint32_t)ctx->num_cu *
        std::max(std::max(ctx->trace_blocks_per_cu, ctx->trace_blocks_per_cu_tri),
                 std::max(ctx->shadow_blocks_per_cu, ctx->shadow_blocks_per_cu_tri)) *
        kTraceBlock;
    if (ctx->d_spill && ctx->spill_threads >= threads) return PT_OK;
    dfree(ctx->d_spill);
    // Four areas: one per chain of a launch.
    PT_HIP(hipMalloc((void **)&ctx->d_spill, 4 * (size_t)threads * kSpillEntries * sizeof(uint2)));
    ctx->spill_threads = threads;
    return PT_OK;
}

DeviceBvh device_bvh(pt_ctx *ctx) { return DeviceBvh{ctx->d_nodes, ctx->d_leaf, ctx->d_spill}; }
// The tree with the spill area of chain `area` of a launch.
DeviceBvh device_bvh_area(pt_ctx *ctx, uint32_t area) {
    return DeviceBvh{ctx->d_nodes, ctx->d_leaf, ctx->d_spill + (size_t)area * ctx->spill_threads * kSpillEntries};
}
PathQueue queue_at(const PathQueue &q, size_t off) {
    return PathQueue{q.ray0 + off, q.ray1 + off, q.tl + off, q.lr + off, q.hit + off};
}
ShadowQueue shadow_queue_at(const ShadowQueue &q, size_t off) { return ShadowQueue{q.s0 + off, q.s1 + off, q.s2 + off}; }

// Events ordering the streams of a render: the fork and the joins of a
// launch's chains.
int ensure_dep_events(pt_ctx *ctx, uint32_t n) {
    while (ctx->dep_events.size() < n) {
        hipEvent_t e = nullptr;
        PT_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        ctx->dep_events.push_back(e);
    }
    return PT_OK;
}

SceneView scene_view(pt_ctx *ctx) {
    SceneView v;
    v.tri = ctx->d_tri;
    v.vpos = reinterpret_cast<const f4 *>(ctx->d_pos);
    v.vnrm = reinterpret_cast<const f4 *>(ctx->d_nrm);
    v.vuv = reinterpret_cast<const f2 *>(ctx->d_uv);
    v.shapes = ctx->d_shapes;
    v.spheres = ctx->d_spheres;
    v.materials = ctx->d_materials;
    v.emitters = ctx->d_emitters;
    v.emitter_cdf = ctx->d_cdf;
    v.num_emitters = (uint32_t)ctx->h_emitters.size();
    v.pad = 0;
    v.tex = TexArrays{ctx->d_bitmaps, reinterpret_cast<const f4 *>(ctx->d_texels)};
    v.env = ctx->env;
    return v;
}

hipEvent_t next_event(pt_ctx *ctx) {
    if (ctx->event_next == ctx->event_pool.size()) {
        hipEvent_t e = nullptr;
        if (hipEventCreate(&e) != hipSuccess) return nullptr;
        ctx->event_pool.push_back(e);
    }
    return ctx->event_pool[ctx->event_next++];
}

// Brackets one kernel launch with events on the pass stream; durations are
// read after the render's final synchronization, so timing adds no host
// synchronization inside the timed region.
// A kernel kind's time is the union of its timers' intervals (the time
// during which at least one of its dispatches ran; kernels of concurrent
// chains overlap), and its launch count the number of distinct groups: timers
// sharing a nonzero `group` (the same bounce of the chains of one pt_render
// launch) count as one launch.
constexpr uint64_t kGroupSingle = 1ull << 63;  // groups of single timers
struct EventTimer {
    pt_ctx *ctx;
    bool on;
    hipStream_t stream;
    uint64_t group;
    hipEvent_t a = nullptr;
    EventTimer(pt_ctx *c, bool enabled, hipStream_t s = nullptr, uint64_t g = 0)
        : ctx(c), on(enabled), stream(s ? s : c->stream), group(g ? g : (kGroupSingle | ++c->timer_groups)) {
        if (on) {
            a = next_event(ctx);
            if (a) (void)hipEventRecord(a, stream);
        }
    }
    void stop(int kind) {
        if (!on || !a) return;
        hipEvent_t b = next_event(ctx);
        if (!b) return;
        (void)hipEventRecord(b, stream);
        ctx->pending.push_back(pt_ctx::Pending{kind, a, b, group});
    }
};

void collect_event_times(pt_ctx *ctx) {
    // Intervals relative to the first timer's start, per kind.
    std::map<int, std::vector<std::pair<float, float>>> spans;
    std::map<int, std::set<uint64_t>> groups;
    const hipEvent_t ref = ctx->pending.empty() ? nullptr : ctx->pending.front().a;
    for (const auto &p : ctx->pending) {
        float t0 = 0.f, t1 = 0.f;
        if (hipEventElapsedTime(&t0, ref, p.a) != hipSuccess || hipEventElapsedTime(&t1, ref, p.b) != hipSuccess)
            continue;
        spans[p.kind].emplace_back(t0, t1);
        groups[p.kind].insert(p.group);
    }
    for (auto &e : spans) {
        const double ms = interval_union(e.second);
        const uint64_t n = groups[e.first].size();
        switch (e.first) {
            case kKindRaygen: ctx->stats.ms_raygen += ms; break;
            case kKindTrace: ctx->stats.ms_trace += ms; ctx->stats.launches_trace += n; break;
            case kKindShade: ctx->stats.ms_shade += ms; ctx->stats.launches_shade += n; break;
            case kKindShadow: ctx->stats.ms_shadow += ms; ctx->stats.launches_shadow += n; break;
            case kKindGuide: ctx->stats.ms_guide += ms; break;
            case kKindGuideInfer: ctx->stats.ms_guide_infer += ms; break;
            default: ctx->stats.ms_accumulate += ms; break;
        }
    }
    ctx->pending.clear();
    ctx->event_next = 0;
}

// A scene description converted to the device record layout (host copies).
struct SceneHostData {
    std::vector<float> h_pos4, h_nrm4, h_uv2;
    std::vector<uint32_t> h_tri4;
    std::vector<ShapeRec> h_shapes;
    std::vector<SphereRec> h_spheres;
    std::vector<float> h_sph4;
    std::vector<MaterialRec> h_materials;
    std::vector<EmitterRec> h_emitters;
    std::vector<float> h_cdf;
    std::vector<BitmapRec> h_bitmaps;
    std::vector<float> h_texels;     // 4 floats per texel
    std::vector<float> h_env_tables; // env map: row cdf (h + 1) | column cdfs ((w + 1) h) | row weights (h)
    std::vector<ShapeHost> shape_host;
    std::vector<uint32_t> vert_owner;
    EnvRec env{};
    CameraRec camera{};
    uint32_t width = 0, height = 0, max_depth = 1, num_tris = 0, num_spheres = 0;
    float lo[3]{}, hi[3]{};
    bool has_normals = false, has_texcoords = false;
    uint32_t feat = 0;  // kFeat* the scene uses (shading kernel choice)
};

// The kFeat* features a scene's shading needs.
uint32_t scene_features(const SceneHostData &d) {
    uint32_t f = 0;
    auto tex = [&](const TextureRec &t) {
        if (t.type == kTexChecker) f |= kFeatChecker;
        if (t.type == kTexBitmap) f |= kFeatBitmap;
    };
    for (const MaterialRec &m : d.h_materials)
        for (const TextureRec &t : m.tex) tex(t);
    for (const EmitterRec &e : d.h_emitters) {
        tex(e.radiance);
        if (e.type == kEmitterSphere) f |= kFeatSphere;
    }
    if (!d.h_spheres.empty()) f |= kFeatSphere;
    if (d.env.type == kEmitterEnvMap) f |= kFeatEnvMap | kFeatBitmap;
    return f;
}

// EnvMapEmitter set-up (world/emitter.cpp:107-149 BuildEnvMapCdfTable and
// :293-312): per row the running luminance sums normalized to a column cdf,
// per row the weight sin((y + 1/2) pi / h), the row cdf over the weighted row
// sums, and the normalization 1 / (sum (2 pi / w) (pi / h)).  Sums run in
// float in the same order.  A black row keeps the 0/0 entries the reference
// table holds.  Frame: to_world = the 3x3 of the emitter transform, to_local
// its inverse.
void build_env_map(const pt_scene_desc *s, SceneHostData &out) {
    const pt_bitmap_desc &bm = s->bitmaps[s->env_bitmap];
    const size_t w = bm.width, h = bm.height;
    std::vector<float> &t = out.h_env_tables;
    t.assign((h + 1) + (w + 1) * h + h, 0.f);
    float *row_cdf = t.data(), *col_cdf = row_cdf + (h + 1), *row_weight = col_cdf + (w + 1) * h;
    size_t ci = 0, ri = 0;
    float row_sum = 0.f;
    row_cdf[ri++] = 0.f;
    for (size_t y = 0; y < h; ++y) {
        float col_sum = 0.f;
        col_cdf[ci++] = 0.f;
        for (size_t x = 0; x < w; ++x) {
            const float *p = bm.rgba + 4 * (y * w + x);
            col_sum += luminance(mk3(p[0], p[1], p[2]));
            col_cdf[ci++] = col_sum;
        }
        for (size_t x = 1; x < w; ++x) col_cdf[ci - x - 1] /= col_sum;
        col_cdf[ci - 1] = 1.f;
        const float weight = std::sin(((float)y + 0.5f) * kPi / (float)h);
        row_weight[y] = weight;
        row_sum += col_sum * weight;
        row_cdf[ri++] = row_sum;
    }
    for (size_t y = 1; y < h; ++y) row_cdf[ri - y - 1] /= row_sum;
    row_cdf[ri - 1] = 1.f;
    EnvRec &e = out.env;
    e.bitmap = s->env_bitmap;
    e.scale = s->env_scale;
    e.w = (uint32_t)w;
    e.h = (uint32_t)h;
    e.normalization = 1.f / (row_sum * (2.f * kPi / (float)w) * (kPi / (float)h));
    const float *m = s->env_to_world;
    double a[3][3], inv[3][3];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) a[i][j] = m[3 * i + j];
    const double det = a[0][0] * (a[1][1] * a[2][2] - a[1][2] * a[2][1]) -
                       a[0][1] * (a[1][0] * a[2][2] - a[1][2] * a[2][0]) +
                       a[0][2] * (a[1][0] * a[2][1] - a[1][1] * a[2][0]);
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            const int r0 = (j + 1) % 3, r1 = (j + 2) % 3, c0 = (i + 1) % 3, c1 = (i + 2) % 3;
            inv[i][j] = (a[r0][c0] * a[r1][c1] - a[r0][c1] * a[r1][c0]) / det;
        }
    for (int i = 0; i < 3; ++i) {
        e.to_world[i] = f4{m[3 * i], m[3 * i + 1], m[3 * i + 2], 0.f};
        e.to_local[i] = f4{(float)inv[i][0], (float)inv[i][1], (float)inv[i][2], 0.f};
    }
}

// Device pointers of an env map's tables inside one uploaded array.
void env_table_pointers(EnvRec &e, const float *base) {
    if (e.type != kEmitterEnvMap || !base) {
        e.row_cdf = e.col_cdf = e.row_weight = nullptr;
        return;
    }
    e.row_cdf = base;
    e.col_cdf = base + (e.h + 1u);
    e.row_weight = e.col_cdf + (size_t)(e.w + 1u) * e.h;
}

// One emitting triangle (SetMeshAreaEmitter, world/emitter.cpp:169-222):
// vertex normals (normalized component-wise as util::Transform::
// TransformNormal does) or the face normal.
EmitterRec tri_emitter(const f3 p[3], const f3 nrm[3], bool has_normals, bool neg_face, const f2 uv[3],
                       const TextureRec &radiance, float radiance_weight) {
    EmitterRec e;
    std::memset(&e, 0, sizeof(e));
    e.type = kEmitterTriArea;
    f3 n[3];
    for (int k = 0; k < 3; ++k) {
        f3 nn = has_normals ? nrm[k] : mk3(0.f);
        if (!has_normals) {
            nn = cross(p[1] - p[0], p[2] - p[0]);
            if (neg_face) nn = -nn;
        }
        float len = sqrtf(nn.x * nn.x + nn.y * nn.y + nn.z * nn.z);
        n[k] = mk3(nn.x / len, nn.y / len, nn.z / len);
    }
    e.p0 = f4{p[0].x, p[0].y, p[0].z, 0.f};
    e.p1 = f4{p[1].x, p[1].y, p[1].z, 0.f};
    e.p2 = f4{p[2].x, p[2].y, p[2].z, 0.f};
    e.n0 = f4{n[0].x, n[0].y, n[0].z, 0.f};
    e.n1 = f4{n[1].x, n[1].y, n[1].z, 0.f};
    e.n2 = f4{n[2].x, n[2].y, n[2].z, 0.f};
    e.uv01 = f4{uv[0].x, uv[0].y, uv[1].x, uv[1].y};
    e.uv2 = f4{uv[2].x, uv[2].y, 0.f, 0.f};
    e.area = length(cross(p[1] - p[0], p[2] - p[0])) * 0.5f;
    e.radiance = radiance;
    e.weight = radiance_weight * e.area;
    return e;
}

// SetSphereAreaEmitter (world/emitter.cpp:224-241); sph4 = center, radius.
EmitterRec sphere_emitter(const float *sph4, const TextureRec &radiance, float radiance_weight) {
    EmitterRec e;
    std::memset(&e, 0, sizeof(e));
    e.type = kEmitterSphere;
    const float rad = sph4[3];
    e.p0 = f4{sph4[0], sph4[1], sph4[2], rad};
    e.area = 4.f * 3.14159265358979323846f * rad * rad;
    e.radiance = radiance;
    e.weight = radiance_weight * e.area;
    return e;
}

// EmitterHelper::ComputeProbability (world/emitter.cpp:321-337) and the
// running-sum table select_emitter searches.
void emitter_probabilities(std::vector<EmitterRec> &em, bool has_env, std::vector<float> &cdf, float &env_prob) {
    float sum = 0.f;
    for (auto &e : em) sum += e.weight;
    const size_t na = em.size();
    const size_t total = na + (has_env ? 1 : 0);
    for (auto &e : em) {
        e.select_prob = e.weight / sum * (float)na;
        e.select_prob = e.select_prob / (float)total;
    }
    cdf.resize(na);
    float acc = 0.f;
    for (size_t i = 0; i < na; ++i) {
        acc = acc + em[i].select_prob;
        cdf[i] = acc;
    }
    env_prob = has_env ? 1.f / (float)total : 0.f;
}

// Rows of the world -> object transform of a sphere placed by `m` (3x4).
void sphere_inverse_rows(const float *m, SphereRec &rec) {
    const double a00 = m[0], a01 = m[1], a02 = m[2], a10 = m[4], a11 = m[5], a12 = m[6];
    const double a20 = m[8], a21 = m[9], a22 = m[10];
    const double c00 = a11 * a22 - a12 * a21, c01 = a02 * a21 - a01 * a22, c02 = a01 * a12 - a02 * a11;
    const double c10 = a12 * a20 - a10 * a22, c11 = a00 * a22 - a02 * a20, c12 = a02 * a10 - a00 * a12;
    const double c20 = a10 * a21 - a11 * a20, c21 = a01 * a20 - a00 * a21, c22 = a00 * a11 - a01 * a10;
    const double det = a00 * c00 + a01 * c10 + a02 * c20;
    const double inv[3][3] = {{c00 / det, c01 / det, c02 / det},
                              {c10 / det, c11 / det, c12 / det},
                              {c20 / det, c21 / det, c22 / det}};
    const double t[3] = {m[3], m[7], m[11]};
    for (int i = 0; i < 3; ++i) {
        const double ti = -(inv[i][0] * t[0] + inv[i][1] * t[1] + inv[i][2] * t[2]);
        rec.w2o[i] = f4{(float)inv[i][0], (float)inv[i][1], (float)inv[i][2], (float)ti};
    }
}

// Whether a 3x4 placement is a finite similarity (uniform scale).
bool is_similarity(const float *m) {
    for (int i = 0; i < 12; ++i)
        if (!std::isfinite(m[i])) return false;
    double c[3][3];
    for (int j = 0; j < 3; ++j)
        for (int i = 0; i < 3; ++i) c[j][i] = m[4 * i + j];
    auto dotc = [&](int a, int b) { return c[a][0] * c[b][0] + c[a][1] * c[b][1] + c[a][2] * c[b][2]; };
    const double n0 = dotc(0, 0), tol = 1e-4 * n0;
    return n0 > 0.0 && std::fabs(dotc(1, 1) - n0) <= tol && std::fabs(dotc(2, 2) - n0) <= tol &&
           std::fabs(dotc(0, 1)) <= tol && std::fabs(dotc(0, 2)) <= tol && std::fabs(dotc(1, 2)) <= tol;
}

// Validation and packing of PTPass::SetScene + World::LoadScene
// (pt_pass.cpp:127-240, world/world.cpp:108-146, world/emitter.cpp:169-337).
// Nothing is produced for a malformed description.
int convert_scene(const pt_scene_desc *s, const char *who, SceneHostData &out, std::string &err) {
    if (!s || (s->num_vertices && !s->positions) || (s->num_triangles && !s->indices) ||
        (s->num_shapes && !s->shapes) || (s->num_materials && !s->materials) || (s->num_spheres && !s->spheres) ||
        s->width == 0 || s->height == 0) {
        err = std::string(who) + ": missing arrays or zero film size";
        return PT_ERR_INVALID_ARG;
    }
    const uint32_t nv = s->num_vertices, nt = s->num_triangles, ns = s->num_shapes, nm = s->num_materials;
    const uint32_t nsph = s->num_spheres;
    if (nt >= kPrimSphere || nsph >= kPrimSphere) {
        err = std::string(who) + ": too many primitives";
        return PT_ERR_CAPACITY;
    }
    // Validate ranges first: nothing is uploaded for a malformed description.
    uint32_t covered = 0, covered_sph = 0;
    for (uint32_t i = 0; i < ns; ++i) {
        const pt_shape_desc &sh = s->shapes[i];
        uint32_t &cov = (sh.flags & PT_SHAPE_SPHERE) ? covered_sph : covered;
        if (sh.first_tri != cov || sh.material >= nm) {
            err = std::string(who) +
                  ": shapes must cover the triangles (and spheres) in order and reference valid materials";
            return PT_ERR_INVALID_ARG;
        }
        cov += sh.num_tris;
    }
    if (covered != nt || covered_sph != nsph) {
        err = std::string(who) + ": shape ranges do not cover all triangles and spheres";
        return PT_ERR_INVALID_ARG;
    }
    for (uint32_t k = 0; k < nsph; ++k) {
        if (!is_similarity(s->spheres[k].to_world)) {
            err = std::string(who) + ": sphere to_world must be a finite similarity transform (uniform scale)";
            return PT_ERR_UNSUPPORTED;
        }
    }
    // Bitmaps and the textures that name them.
    const uint32_t nb = s->num_bitmaps;
    if (nb && !s->bitmaps) {
        err = std::string(who) + ": missing bitmap array";
        return PT_ERR_INVALID_ARG;
    }
    for (uint32_t i = 0; i < nb; ++i) {
        const pt_bitmap_desc &b = s->bitmaps[i];
        if (!b.rgba || b.width == 0 || b.height == 0 || b.width > 65536 || b.height > 65536 ||
            b.filter > PT_FILTER_LINEAR || b.wrap > PT_WRAP_BORDER) {
            err = std::string(who) + ": invalid bitmap " + std::to_string(i);
            return PT_ERR_INVALID_ARG;
        }
    }
    auto tex_ok = [&](const pt_texture_desc &t) {
        return t.type == PT_TEX_RGB || t.type == PT_TEX_CHECKER || (t.type == PT_TEX_BITMAP && t.bitmap < nb);
    };
    for (uint32_t i = 0; i < nm; ++i)
        for (int k = 0; k < 4; ++k)
            if (!tex_ok(s->materials[i].tex[k])) {
                err = std::string(who) + ": material texture names an unknown type or bitmap";
                return PT_ERR_INVALID_ARG;
            }
    for (uint32_t i = 0; i < ns; ++i)
        if (s->shapes[i].is_emitter && !tex_ok(s->shapes[i].radiance)) {
            err = std::string(who) + ": emitter radiance names an unknown texture type or bitmap";
            return PT_ERR_INVALID_ARG;
        }
    if (s->env_type > PT_ENV_MAP ||
        (s->env_type == PT_ENV_MAP && (s->env_bitmap >= nb || s->bitmaps[s->env_bitmap].height < 2))) {
        err = std::string(who) + ": invalid environment (an env map needs a bitmap of height >= 2)";
        return PT_ERR_INVALID_ARG;
    }
    for (uint32_t i = 0; i < 3 * nt; ++i)
        if (s->indices[i] >= nv) {
            err = std::string(who) + ": vertex index out of range";
            return PT_ERR_INVALID_ARG;
        }
    for (uint32_t i = 0; i < nm; ++i)
        if (s->materials[i].type < PT_MAT_DIFFUSE || s->materials[i].type > PT_MAT_ROUGH_PLASTIC) {
            err = std::string(who) + ": unknown material type";
            return PT_ERR_UNSUPPORTED;
        }
    out.h_pos4.assign((size_t)nv * 4, 0.f);
    out.h_nrm4.assign(s->normals ? (size_t)nv * 4 : 0, 0.f);
    out.h_uv2.assign(s->texcoords ? (size_t)nv * 2 : 0, 0.f);
    for (uint32_t i = 0; i < nv; ++i) {
        for (int k = 0; k < 3; ++k) out.h_pos4[4 * i + k] = s->positions[3 * i + k];
        if (s->normals)
            for (int k = 0; k < 3; ++k) out.h_nrm4[4 * i + k] = s->normals[3 * i + k];
        if (s->texcoords)
            for (int k = 0; k < 2; ++k) out.h_uv2[2 * i + k] = s->texcoords[2 * i + k];
    }
    for (int k = 0; k < 3; ++k) {
        out.lo[k] = INFINITY;
        out.hi[k] = -INFINITY;
    }
    for (uint32_t i = 0; i < nv; ++i)
        for (int k = 0; k < 3; ++k) {
            out.lo[k] = std::min(out.lo[k], s->positions[3 * i + k]);
            out.hi[k] = std::max(out.hi[k], s->positions[3 * i + k]);
        }
    // Spheres: world center = to_world * (0, 0, 0), radius = |to_world * (1,
    // 0, 0) - center| (world/emitter.cpp:224-241); rows of the inverse
    // transform for shading.
    out.h_sph4.assign((size_t)nsph * 4, 0.f);
    out.h_spheres.assign(nsph, SphereRec{});
    for (uint32_t k = 0; k < nsph; ++k) {
        const float *m = s->spheres[k].to_world;
        float *q = &out.h_sph4[4 * (size_t)k];
        sphere_center_radius(m, q);
        for (int a = 0; a < 3; ++a) {
            out.lo[a] = std::min(out.lo[a], q[a] - q[3]);
            out.hi[a] = std::max(out.hi[a], q[a] + q[3]);
        }
        sphere_inverse_rows(m, out.h_spheres[k]);
    }
    for (int k = 0; k < 3; ++k)
        if (!(out.lo[k] <= out.hi[k])) out.lo[k] = out.hi[k] = 0.f;
    out.h_tri4.assign((size_t)nt * 4, 0u);
    out.h_shapes.resize(ns);
    out.shape_host.assign(ns, ShapeHost());
    out.vert_owner.assign(nv, kVertexUnused);
    for (uint32_t si = 0; si < ns; ++si) {
        const pt_shape_desc &sh = s->shapes[si];
        ShapeHost &h = out.shape_host[si];
        h.first = sh.first_tri;
        h.count = sh.num_tris;
        h.sphere = (sh.flags & PT_SHAPE_SPHERE) != 0;
        h.emitter = sh.is_emitter != 0;
        h.radiance = to_texture_rec(sh.radiance);
        h.radiance_weight = texture_weight(sh.radiance, s->bitmaps);
        if (h.sphere) continue;
        h.vmin = 0xFFFFFFFFu;
        h.vmax = 0;
        for (uint64_t i = 3ull * sh.first_tri; i < 3ull * ((uint64_t)sh.first_tri + sh.num_tris); ++i) {
            const uint32_t v = s->indices[i];
            h.vmin = std::min(h.vmin, v);
            h.vmax = std::max(h.vmax, v);
            uint32_t &o = out.vert_owner[v];
            if (o == kVertexUnused) {
                o = si;
            } else if (o != si) {
                if (o != kVertexShared) out.shape_host[o].shared = true;
                h.shared = true;
                o = kVertexShared;
            }
        }
        if (sh.num_tris == 0) h.vmin = h.vmax = 0;
    }
    // Area emitters: one per triangle of each emitting shape, in shape order
    // (world/emitter.cpp:169-222, pt_pass.cpp:219-235).
    out.h_emitters.clear();
    for (uint32_t si = 0; si < ns; ++si) {
        const pt_shape_desc &sh = s->shapes[si];
        ShapeRec r;
        r.material = sh.material;
        r.flags = sh.flags;
        if (!s->normals) r.flags &= ~(uint32_t)PT_SHAPE_HAS_NORMALS;
        if (!s->texcoords) r.flags &= ~(uint32_t)PT_SHAPE_HAS_TEXCOORDS;
        r.first_tri = sh.first_tri;
        r.emitter_offset = sh.is_emitter ? (int32_t)out.h_emitters.size() : -1;
        if (sh.flags & PT_SHAPE_SPHERE) {
            r.flags = sh.flags & (uint32_t)(PT_SHAPE_SPHERE | PT_SHAPE_FLIP_NORMALS);
            out.h_shapes[si] = r;
            out.shape_host[si].flags0 = r.flags;
            for (uint32_t k = sh.first_tri; k < sh.first_tri + sh.num_tris; ++k) {
                out.h_spheres[k].shape = si;
                if (!sh.is_emitter) continue;
                out.h_emitters.push_back(sphere_emitter(&out.h_sph4[4 * (size_t)k], to_texture_rec(sh.radiance),
                                                        texture_weight(sh.radiance, s->bitmaps)));
            }
            continue;
        }
        out.h_shapes[si] = r;
        out.shape_host[si].flags0 = r.flags;
        for (uint32_t t = sh.first_tri; t < sh.first_tri + sh.num_tris; ++t) {
            for (int k = 0; k < 3; ++k) out.h_tri4[4 * (size_t)t + k] = s->indices[3 * (size_t)t + k];
            out.h_tri4[4 * (size_t)t + 3] = si;
            if (!sh.is_emitter) continue;
            f3 p[3], n[3];
            f2 uv[3];
            const bool has_n = s->normals && (r.flags & kShapeHasNormals);
            for (int k = 0; k < 3; ++k) {
                uint32_t vi = s->indices[3 * (size_t)t + k];
                p[k] = mk3(s->positions[3 * vi], s->positions[3 * vi + 1], s->positions[3 * vi + 2]);
                n[k] = has_n ? mk3(s->normals[3 * vi], s->normals[3 * vi + 1], s->normals[3 * vi + 2]) : mk3(0.f);
                uv[k] = s->texcoords ? mk2(s->texcoords[2 * vi], s->texcoords[2 * vi + 1]) : mk2(0.f, 0.f);
            }
            out.h_emitters.push_back(tri_emitter(p, n, has_n, (r.flags & kShapeNegFaceNormal) != 0, uv,
                                                 to_texture_rec(sh.radiance), texture_weight(sh.radiance, s->bitmaps)));
        }
    }
    {
        float env_prob;
        emitter_probabilities(out.h_emitters, s->env_type != PT_ENV_NONE, out.h_cdf, env_prob);
        std::memset(&out.env, 0, sizeof(out.env));
        const float total_inv = env_prob;
        if (s->env_type == PT_ENV_CONSTANT) {
            out.env.type = kEmitterConstEnv;
            out.env.select_prob = total_inv;
            out.env.color = f4{s->env_radiance[0], s->env_radiance[1], s->env_radiance[2], 0.f};
        } else if (s->env_type == PT_ENV_MAP) {
            out.env.type = kEmitterEnvMap;
            out.env.select_prob = total_inv;
        }
    }
    out.h_materials.resize(nm);
    for (uint32_t i = 0; i < nm; ++i) out.h_materials[i] = to_material_rec(s->materials[i], s->bitmaps);
    {
        // Texels are stored once per image: bitmaps that sample the same
        // texel array in other modes share it.
        std::map<const float *, size_t> stored;
        out.h_bitmaps.resize(nb);
        out.h_texels.clear();
        for (uint32_t i = 0; i < nb; ++i) {
            const pt_bitmap_desc &b = s->bitmaps[i];
            const size_t n = (size_t)b.width * b.height;
            auto it = stored.find(b.rgba);
            size_t off;
            if (it != stored.end()) {
                off = it->second;
            } else {
                off = out.h_texels.size() / 4;
                out.h_texels.insert(out.h_texels.end(), b.rgba, b.rgba + 4 * n);
                stored[b.rgba] = off;
            }
            BitmapRec r;
            std::memset(&r, 0, sizeof(r));
            r.offset = off;
            r.w = b.width;
            r.h = b.height;
            r.filter = b.filter;
            r.wrap = b.wrap;
            out.h_bitmaps[i] = r;
        }
    }
    if (s->env_type == PT_ENV_MAP) build_env_map(s, out);
    for (int r = 0; r < 4; ++r) {
        out.camera.s2c[r] = f4{s->camera.sample_to_camera[4 * r], s->camera.sample_to_camera[4 * r + 1],
                                s->camera.sample_to_camera[4 * r + 2], s->camera.sample_to_camera[4 * r + 3]};
        out.camera.c2w[r] = f4{s->camera.camera_to_world[4 * r], s->camera.camera_to_world[4 * r + 1],
                                s->camera.camera_to_world[4 * r + 2], s->camera.camera_to_world[4 * r + 3]};
    }
    out.width = s->width;
    out.height = s->height;
    out.max_depth = s->max_depth ? s->max_depth : 1;
    out.num_tris = nt;
    out.num_spheres = nsph;

    out.has_normals = s->normals != nullptr;
    out.has_texcoords = s->texcoords != nullptr;
    out.feat = scene_features(out);
    return PT_OK;
}

// --- batched views: host side ------------------------------------------------------
constexpr uint32_t kBatchDefaultPaths = 32u << 20;

uint32_t batch_path_bits(size_t views) {
    uint32_t vb = 0;
    while (((size_t)1 << vb) < views) ++vb;
    return 31u - vb;
}

void batch_free_device(pt_ctx *ctx) {
    BatchState &b = ctx->batch;
    dfree(b.d_tri);
    dfree(b.d_pos);
    dfree(b.d_nrm);
    dfree(b.d_uv);
    dfree(b.d_shapes);
    dfree(b.d_spheres);
    dfree(b.d_materials);
    dfree(b.d_emitters);
    dfree(b.d_cdf);
    dfree(b.d_bitmaps);
    dfree(b.d_texels);
    dfree(b.d_env_tables);
    dfree(b.d_nodes);
    dfree(b.d_leaf);
    dfree(b.d_views);
    dfree(b.d_roots);
    b.dirty = true;
}

void batch_free_all(pt_ctx *ctx) {
    BatchState &b = ctx->batch;
    batch_free_device(ctx);
    for (auto &v : b.views) {
        dfree(v.d_accum);
        dfree(v.d_frame);
        dfree(v.d_accum64);
    }
    dfree(b.d_launch);
    dfree(b.d_rays);
    dfree(b.d_gate_acc);
    dfree(b.d_gate_off);
    if (b.h_launch) (void)hipHostFree(b.h_launch);
    b.h_launch = nullptr;
    b.h_launch_slots = 0;
    b = BatchState();
}

int batch_upload(pt_ctx *ctx) {
    BatchState &b = ctx->batch;
    if (!b.dirty) return PT_OK;
    int rc;
    const size_t nv = b.pos4.size() / 4;
    if ((rc = upload(ctx, b.d_tri, b.tri4.data(), b.tri4.size())) != PT_OK) return rc;
    if ((rc = upload(ctx, b.d_pos, reinterpret_cast<const float4 *>(b.pos4.data()), nv)) != PT_OK) return rc;
    if (b.any_normals &&
        (rc = upload(ctx, b.d_nrm, reinterpret_cast<const float4 *>(b.nrm4.data()), nv)) != PT_OK)
        return rc;
    if (b.any_texcoords &&
        (rc = upload(ctx, b.d_uv, reinterpret_cast<const float2 *>(b.uv2.data()), nv)) != PT_OK)
        return rc;
    if ((rc = upload(ctx, b.d_shapes, b.shapes.data(), b.shapes.size())) != PT_OK) return rc;
    if ((rc = upload(ctx, b.d_spheres, b.spheres.data(), b.spheres.size())) != PT_OK) return rc;
    if ((rc = upload(ctx, b.d_materials, b.materials.data(), b.materials.size())) != PT_OK) return rc;
    if ((rc = upload(ctx, b.d_emitters, b.emitters.data(), b.emitters.size())) != PT_OK) return rc;
    if ((rc = upload(ctx, b.d_cdf, b.cdf.data(), b.cdf.size())) != PT_OK) return rc;
    if ((rc = upload(ctx, b.d_bitmaps, b.bitmaps.data(), b.bitmaps.size())) != PT_OK) return rc;
    if ((rc = upload(ctx, b.d_texels, reinterpret_cast<const float4 *>(b.texels.data()), b.texels.size() / 4)) != PT_OK)
        return rc;
    if ((rc = upload(ctx, b.d_env_tables, b.env_tables.data(), b.env_tables.size())) != PT_OK) return rc;
    if ((rc = upload(ctx, b.d_nodes, b.nodes.data(), b.nodes.size())) != PT_OK) return rc;
    if ((rc = upload(ctx, b.d_leaf, b.leaf.data(), b.leaf.size())) != PT_OK) return rc;
    std::vector<BatchViewDev> vd(b.views.size());
    std::vector<uint32_t> roots(b.views.size());
    for (size_t i = 0; i < b.views.size(); ++i) {
        const BatchView &v = b.views[i];
        BatchViewDev &d = vd[i];
        std::memset(&d, 0, sizeof(d));
        d.camera = v.camera;
        d.env = v.env;
        env_table_pointers(d.env, b.d_env_tables ? b.d_env_tables + v.env_tables : nullptr);
        d.width = v.width;
        d.height = v.height;
        d.max_depth = v.max_depth;
        d.root = v.root;
        d.emitter_begin = v.emitter_begin;
        d.emitter_count = v.emitter_count;
        d.accum = v.d_accum;
        d.frame = v.d_frame;
        d.accum64 = v.d_accum64;
        roots[i] = v.root;
    }
    if ((rc = upload(ctx, b.d_views, vd.data(), vd.size())) != PT_OK) return rc;
    if ((rc = upload(ctx, b.d_roots, roots.data(), roots.size())) != PT_OK) return rc;
    if (!b.d_launch) PT_HIP(hipMalloc((void **)&b.d_launch, sizeof(LaunchDesc)));
    if (!b.d_rays) {
        PT_HIP(hipMalloc((void **)&b.d_rays, 2 * sizeof(unsigned long long)));
        PT_HIP(hipMemset(b.d_rays, 0, 2 * sizeof(unsigned long long)));
    }
    b.dirty = false;
    return PT_OK;
}

// Brings the BVH up to date with pending shape updates: leaf records from
// the current geometry, then node boxes level by level from the deepest.
int bvh_refit(pt_ctx *ctx) {
    if (!ctx->bvh_dirty || !ctx->has_bvh) return PT_OK;
    const auto t0 = std::chrono::steady_clock::now();
    (void)hipSetDevice(ctx->device);
    PT_HIP(hipMemsetAsync(ctx->d_refit_err, 0, sizeof(uint32_t), ctx->stream));
    if (ctx->num_leaf) {
        const uint32_t grid = std::min<uint32_t>((ctx->num_leaf + 255u) / 256u, (uint32_t)ctx->num_cu * 16u);
        hipLaunchKernelGGL(k_refit_leaves, dim3(grid), dim3(256), 0, ctx->stream, ctx->d_leaf, ctx->num_leaf,
                           ctx->d_tri, ctx->d_pos, ctx->d_sph4);
    }
    for (size_t l = ctx->level_off.size() - 1; l-- > 0;) {
        const uint32_t begin = ctx->level_off[l], count = ctx->level_off[l + 1] - begin;
        if (count == 0) continue;
        hipLaunchKernelGGL(k_refit_nodes, dim3((count + 63u) / 64u), dim3(64), 0, ctx->stream, ctx->d_nodes,
                           ctx->d_node_box, ctx->d_leaf, ctx->d_levels + begin, count, ctx->d_refit_err);
    }
    uint32_t failed = 0;
    PT_HIP(hipGetLastError());
    PT_HIP(hipMemcpyAsync(&failed, ctx->d_refit_err, sizeof(uint32_t), hipMemcpyDeviceToHost, ctx->stream));
    PT_HIP(hipStreamSynchronize(ctx->stream));
    if (failed) {
        ctx->error = "BVH refit: a node box cannot be quantized; rebuild with pt_build_bvh";
        ctx->has_bvh = false;
        return PT_ERR_CAPACITY;
    }
    ctx->bvh_dirty = false;
    ctx->stats.refits += 1;
    ctx->stats.ms_refit += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return PT_OK;
}

struct BatchShape {
    uint32_t capacity, shade_block, shade_bpc, trace_bpc;
    uint32_t train_paths;  // training launch width (0 = kTrainLaunchPaths)
};

// Training samples for `steps` optimizer steps of `batch` samples (capped at
// 2^28 samples, 12 GB; the optimizer then runs fewer steps).
uint32_t guide_sample_count(uint64_t steps, uint32_t batch) {
    return (uint32_t)std::min<uint64_t>(steps * batch, 1ull << 28);
}

// Guided batched render: network outputs per queue slot and the record sink
// (epoch and phase of each launch are read from its LaunchDesc).
constexpr uint32_t kMixLaunches = 8;  // training launches whose samples one optimizer step mixes

struct BatchGuide {
    bool infer = false, train = false;
    GuideView view{};
    GuideShadeArgs ga{};
    uint32_t gdepth = 0;  // vertices below this depth may sample the guide
    // With several views, training samples go to a ring that holds about the
    // last kMixLaunches training launches, and every optimizer step draws its
    // batch from the whole ring, so each step mixes several views instead of
    // fitting the shared MLP to one view after another.
    uint32_t ring_n = 0;  // slots (a multiple of the batch size); 0 = one launch at a time
    uint32_t step = 0;    // optimizer steps taken on the ring
};

// The kernels of one launch (the sequence a hipGraph captures).  As in
// pt_render, the launch's paths run as concurrent chains, each on its own
// stream with its own queue part (a fixed share of the queues: slot ranges
// come from the device launch descriptor), counters and spill area; under
// capture the chains fork from and join the pass stream through events, so
// the graph holds them as parallel branches.  Accumulation follows the join.
// Needs ensure_dep_events(ctx, kMaxChains) and queues of batch_queue_paths.
uint32_t batch_chains(const BatchShape &sh) { return chain_count(sh.capacity, false); }
uint32_t batch_queue_paths(const BatchShape &sh) { return batch_queue_slots(sh.capacity); }
void batch_enqueue_launch(pt_ctx *ctx, const BatchShape &sh, uint32_t max_depth, const ShadeParams &sp,
                          uint32_t path_bits, bool fp64, const BatchGuide &bg, bool timing = false,
                          uint64_t launch_group = 0) {
    BatchState &b = ctx->batch;
    SceneView sv;
    sv.tri = b.d_tri;
    sv.vpos = reinterpret_cast<const f4 *>(b.d_pos);
    sv.vnrm = reinterpret_cast<const f4 *>(b.d_nrm);
    sv.vuv = reinterpret_cast<const f2 *>(b.d_uv);
    sv.shapes = b.d_shapes;
    sv.spheres = b.d_spheres;
    sv.materials = b.d_materials;
    sv.emitters = b.d_emitters;
    sv.emitter_cdf = b.d_cdf;
    sv.num_emitters = 0;
    sv.pad = 0;
    sv.tex = TexArrays{b.d_bitmaps, reinterpret_cast<const f4 *>(b.d_texels)};
    std::memset(&sv.env, 0, sizeof(sv.env));
    AovBuffers aov{nullptr, nullptr, nullptr, nullptr, 0u, 1u};
    const uint32_t tg = (uint32_t)ctx->num_cu * sh.trace_bpc;
    const uint32_t sg = (uint32_t)ctx->num_cu * sh.shade_bpc;
    const uint32_t nchain = batch_chains(sh);
    const size_t share = batch_share(batch_queue_paths(sh), nchain);  // queue slots per chain
    const hipStream_t streams[kMaxChains] = {ctx->stream, ctx->stream2, ctx->stream3, ctx->stream4};
    struct Chain {
        hipStream_t stream;
        uint32_t *counters, *ext, *shd, *twork, *swork;
        PathQueue q[2];
        ShadowQueue sq;
        DeviceBvh bvh;
        GuideShadeArgs ga;
        uint4 *gout;
    } ch[kMaxChains];
    for (uint32_t c = 0; c < nchain; ++c) {
        Chain &k = ch[c];
        k.stream = streams[c];
        k.counters = ctx->d_counters + (size_t)c * ctx->num_counters;
        k.ext = k.counters;
        k.shd = k.counters + (max_depth + 2);
        k.twork = k.counters + 2 * (max_depth + 2);
        k.swork = k.counters + 3 * (max_depth + 2);
        k.q[0] = queue_at(ctx->q[0], c * share);
        k.q[1] = queue_at(ctx->q[1], c * share);
        k.sq = shadow_queue_at(ctx->sq, c * share);
        k.bvh = DeviceBvh{b.d_nodes, b.d_leaf, ctx->d_spill + (size_t)c * ctx->spill_threads * kSpillEntries};
        k.gout = ctx->d_gout ? ctx->d_gout + 6 * c * share : nullptr;
        k.ga = bg.ga;
    }
    if (nchain > 1) (void)hipEventRecord(ctx->dep_events[0], ctx->stream);
    for (uint32_t c = 1; c < nchain; ++c) (void)hipStreamWaitEvent(ch[c].stream, ctx->dep_events[0], 0);
    for (uint32_t c = 0; c < nchain; ++c)
        hipLaunchKernelGGL(k_raygen_batch, dim3((uint32_t)ctx->num_cu * 32u), dim3(256), 0, ch[c].stream, ch[c].q[0],
                           b.d_launch, b.d_views, path_bits, ch[c].counters, ctx->num_counters, c, nchain);
    int cur = 0;
    for (uint32_t bb = 0; bb < max_depth; ++bb) {
        const int nxt = cur ^ 1;
        const uint64_t group = launch_group ? (launch_group | (bb + 1)) : 0u;
        const bool infer_here = bg.infer && bb < bg.gdepth && bb + 1 < max_depth;
        for (uint32_t c = 0; c < nchain; ++c) {
            const Chain &k = ch[c];
            {
                EventTimer t(ctx, timing, k.stream, group);
                hipLaunchKernelGGL(trace_kernel(false, !b.spheres.empty()), dim3(tg), dim3(kTraceBlock), 0, k.stream,
                                   k.q[cur], k.bvh, k.ext + bb, k.twork + bb, ctx->d_stat, (const uint32_t *)b.d_roots,
                                   path_bits);
                t.stop(kKindTrace);
            }
            if (bg.infer || bg.train) {
                GuideShadeArgs ga = k.ga;
                ga.gout = infer_here ? k.gout : nullptr;
                if (infer_here) {
                    EventTimer t(ctx, timing, k.stream, group);
                    hipLaunchKernelGGL(k_guide_infer, dim3(ctx->num_cu * ctx->infer_blocks_per_cu), dim3(kInferBlock),
                                       0, k.stream, bg.view, k.q[cur], k.ext + bb, reinterpret_cast<uint2 *>(k.gout),
                                       ctx->d_stat + 4, path_bits, InferScene{sv.tri, sv.shapes, sv.spheres, sv.materials});
                    t.stop(kKindGuideInfer);
                }
                EventTimer t(ctx, timing, k.stream, group);
                hipLaunchKernelGGL(k_shade_guided<true>, dim3((uint32_t)ctx->num_cu * sh.shade_bpc), dim3(256), 0,
                                   k.stream, sv, sp, k.q[cur], k.ext + bb, k.q[nxt], k.ext + bb + 1, k.sq, k.shd + bb,
                                   ctx->d_out_l, aov, ga, (const BatchViewDev *)b.d_views, path_bits,
                                   (const LaunchDesc *)b.d_launch);
                t.stop(kKindShade);
            } else {
                EventTimer t(ctx, timing, k.stream, group);
                hipLaunchKernelGGL(shade_kernel(true, b.feat), dim3(sg), dim3(sh.shade_block), 0, k.stream, sv, sp,
                                   k.q[cur], k.ext + bb, k.q[nxt], k.ext + bb + 1, k.sq, k.shd + bb, ctx->d_out_l, aov,
                                   (const BatchViewDev *)b.d_views, path_bits);
                t.stop(kKindShade);
            }
            {
                EventTimer t(ctx, timing, k.stream, group);
                hipLaunchKernelGGL(shadow_kernel(false, !b.spheres.empty()), dim3(shadow_grid(ctx, !b.spheres.empty())),
                                   dim3(kTraceBlock), 0, k.stream, k.sq, k.bvh, k.shd + bb, k.swork + bb, k.q[nxt],
                                   ctx->d_out_l, ctx->d_stat + 2, (const uint32_t *)b.d_roots, path_bits);
                t.stop(kKindShadow);
            }
        }
        cur = nxt;
    }
    for (uint32_t c = 1; c < nchain; ++c) {
        (void)hipEventRecord(ctx->dep_events[c], ch[c].stream);
        (void)hipStreamWaitEvent(ctx->stream, ctx->dep_events[c], 0);
    }
    hipLaunchKernelGGL(k_accumulate_batch, dim3((uint32_t)ctx->num_cu * 16u), dim3(256), 0, ctx->stream,
                       ctx->d_out_l, b.d_launch, b.d_views, fp64 ? 1u : 0u);
    for (uint32_t c = 0; c < nchain; ++c)
        hipLaunchKernelGGL(k_batch_count, dim3(1), dim3(64), 0, ctx->stream, ch[c].counters, max_depth, b.d_rays);
}

// Width of a launch of guide-training frames: by default at most 2^23 paths
// (four 1080p frames, of different views during the round-robin) and at least
// one frame of the largest view; the network is updated after every training
// launch.  Measured on 3 door-room + 3 terrain views at 1080p: with one-frame
// (2^21-path) launches the door views' guided frames 8..15 were only
// 1.2-1.3x better than plain ones and the gate switched the guide of one door
// view off in three of four runs; with four-frame launches (a sample ring of
// four times as many frames) they were 2.0-2.5x better and every door view
// kept its guide, at the same render time.
constexpr uint32_t kTrainLaunchPaths = 1u << 23;
uint32_t batch_train_capacity(const BatchState &b, const BatchShape &sh) {
    uint32_t max_npix = 1;
    for (const auto &v : b.views) max_npix = std::max(max_npix, v.width * v.height);
    return std::min(sh.capacity, std::max(sh.train_paths ? sh.train_paths : kTrainLaunchPaths, max_npix));
}

// Checks and allocations of a guided batched render: network outputs per
// queue slot, training records for the training launch width, samples for
// the optimizer steps of the fullest training launch.
int batch_guide_setup(pt_ctx *ctx, const pt_batch_params &p, const BatchShape &sh, uint32_t max_depth,
                      const std::vector<LaunchDesc> &plan, BatchGuide &bg) {
    if (!bg.infer && !bg.train) return PT_OK;
    GuideState &gs = ctx->guide;
    if (!gs.enabled || !gs.for_batch || gs.grids != ctx->batch.views.size()) {
        ctx->error = "pt_batch_render: guided rendering needs pt_guide_enable_batch";
        return PT_ERR_NO_GUIDE;
    }
    if (p.mis_mode != PT_MIS_CONSISTENT) {
        ctx->error = "pt_batch_render: guided rendering needs PT_MIS_CONSISTENT";
        return PT_ERR_INVALID_ARG;
    }
    bg.view = gs.view();
    std::memset(&bg.ga, 0, sizeof(bg.ga));
    bg.ga.fixed_c = std::min(gs.cfg.fixed_selection, 1.f);
    bg.gdepth = gs.cfg.guided_depth ? gs.cfg.guided_depth : max_depth;
    if (bg.infer) {
        if (ctx->gout_cap < batch_queue_paths(sh)) {
            dfree(ctx->d_gout);
            PT_HIP(hipMalloc((void **)&ctx->d_gout, (size_t)batch_queue_paths(sh) * 6 * sizeof(uint4)));
            ctx->gout_cap = batch_queue_paths(sh);
        }
        bg.ga.gout = ctx->d_gout;
    }
    if (bg.train) {
        const pt_guide_config &gc = gs.cfg;
        const uint32_t train_cap = batch_train_capacity(ctx->batch, sh);
        const uint64_t slots = ((uint64_t)train_cap + gc.train_stride - 1) / gc.train_stride;
        const uint64_t records = slots * max_depth;
        if (records > 0xFFFFFFFFull / 2) {
            ctx->error = "pt_batch_render: too many guide training records per launch (raise train_stride)";
            return PT_ERR_CAPACITY;
        }
        uint32_t max_frames = 1;
        for (const auto &v : ctx->batch.views)
            max_frames = std::max(max_frames, std::min(kMaxLaunchEntries, train_cap / (v.width * v.height)));
        if (!gs.ensure_records((uint32_t)records, ctx->error) ||
            !gs.ensure_samples(guide_sample_count((uint64_t)gc.steps_per_frame * max_frames, gc.batch_size),
                               ctx->error))
            return PT_ERR_HIP;
        // Ring of about kMixLaunches average training launches (their sample
        // budgets), a multiple of the batch size.
        uint64_t frames = 0, launches = 0;
        for (const LaunchDesc &l : plan)
            if (l.train_frames) {
                frames += l.train_frames;
                ++launches;
            }
        if (ctx->batch.views.size() > 1 && launches > 1) {
            const uint64_t per = std::max<uint64_t>(1, (frames * gc.steps_per_frame + launches - 1) / launches);
            const uint64_t batches = std::min<uint64_t>(per * std::min<uint64_t>(kMixLaunches, launches),
                                                        (1ull << 28) / gc.batch_size);
            bg.ring_n = (uint32_t)(batches * gc.batch_size);
            if (!gs.ensure_samples(bg.ring_n, ctx->error)) return PT_ERR_HIP;
        }
        bg.ga.rs.rec = reinterpret_cast<f4 *>(gs.rec);
        bg.ga.rs.capacity = gs.rec_cap;
        bg.ga.rs.stride = gc.train_stride;
        bg.ga.rs.phase = 0;  // per launch from its LaunchDesc
        bg.ga.rs.max_depth = max_depth;
        bg.ga.rs.epoch = 0;
    }
    return PT_OK;
}

// Renders the batch's frames with one launch shape.  Returns PT_OK and the
// wall time in *ms.
int batch_run(pt_ctx *ctx, const pt_batch_params &p, const BatchShape &sh, bool graph, double *ms,
              uint32_t *launches) {
    BatchState &b = ctx->batch;
    const auto t0 = std::chrono::steady_clock::now();
    int rc = batch_upload(ctx);
    if (rc != PT_OK) return rc;
    const uint32_t path_bits = batch_path_bits(b.views.size());
    uint32_t max_depth = 1, max_npix = 0;
    for (const auto &v : b.views) {
        max_depth = std::max(max_depth, v.max_depth);
        max_npix = std::max(max_npix, v.width * v.height);
    }
    if (sh.capacity < max_npix || sh.capacity > (1u << path_bits)) {
        ctx->error = "pt_batch_render: paths_per_launch must hold one frame of the largest view and fit the "
                     "path-id range";
        return PT_ERR_CAPACITY;
    }
    if ((rc = ensure_queues(ctx, batch_queue_paths(sh), max_depth)) != PT_OK) return rc;
    if ((rc = ensure_spill(ctx)) != PT_OK) return rc;
    BatchGuide bg;
    bg.infer = (p.render_flags & PT_RENDER_GUIDE) != 0;
    bg.train = (p.render_flags & PT_RENDER_GUIDE_TRAIN) != 0;
    const bool gate = bg.train && (p.batch_flags & PT_BATCH_GUIDE_GATE) != 0;
    const uint32_t nviews = (uint32_t)b.views.size();
    PlanState ps;
    ps.next.assign(nviews, 0u);
    ps.T.assign(nviews, 0u);
    if (bg.train)
        for (uint32_t vi = 0; vi < nviews; ++vi) ps.T[vi] = training_frames(b.views[vi].spp, p.guide_train_spp);
    ps.train_end = ps.T;
    const uint32_t train_cap = batch_train_capacity(b, sh);
    std::vector<LaunchDesc> plan;
    batch_plan(b.views, sh.capacity, true, plan, ps, train_cap, gate);
    if (bg.train && b.gate_views < b.views.size()) {
        dfree(b.d_gate_acc);
        dfree(b.d_gate_off);
        PT_HIP(hipMalloc((void **)&b.d_gate_acc, b.views.size() * kGateDoublesPerView * sizeof(double)));
        PT_HIP(hipMalloc((void **)&b.d_gate_off, sizeof(uint32_t)));
        b.gate_views = (uint32_t)b.views.size();
    }
    // Launch descriptors are copied to the device from pinned slots, slot k
    // for the plan's launch k; (re)staged only while no copy is in flight (at
    // the start of a render or after a stream synchronization).
    auto stage_plan = [&]() -> int {
        if (plan.size() > b.h_launch_slots) {
            if (b.h_launch) (void)hipHostFree(b.h_launch);
            b.h_launch = nullptr;
            b.h_launch_slots = 0;
            PT_HIP(hipHostMalloc((void **)&b.h_launch, plan.size() * sizeof(LaunchDesc), hipHostMallocDefault));
            b.h_launch_slots = (uint32_t)plan.size();
        }
        std::memcpy(b.h_launch, plan.data(), plan.size() * sizeof(LaunchDesc));
        return PT_OK;
    };
    if ((rc = stage_plan()) != PT_OK) return rc;
    ShadeParams spar;
    spar.max_depth = max_depth;
    spar.rr_depth = p.rr_depth;
    spar.mis_mode = p.mis_mode;
    spar.aov = 0;
    const bool fp64 = (p.render_flags & PT_RENDER_FP64_ACCUM) != 0;
    if ((rc = batch_guide_setup(ctx, p, sh, max_depth, plan, bg)) != PT_OK) return rc;
    if ((rc = ensure_dep_events(ctx, kMaxChains)) != PT_OK) return rc;
    // Per-kernel event timing (direct launches only; around the training
    // steps in both modes).
    const bool timing = (p.render_flags & PT_RENDER_TIMING) != 0;
    hipGraph_t g = nullptr;
    hipGraphExec_t ge = nullptr;
    if (graph) {
        PT_HIP(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
        batch_enqueue_launch(ctx, sh, max_depth, spar, path_bits, fp64, bg);
        hipError_t e = hipStreamEndCapture(ctx->stream, &g);
        if (e == hipSuccess) e = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
        if (e != hipSuccess) {
            if (g) (void)hipGraphDestroy(g);
            ctx->error = std::string("pt_batch_render: graph capture failed: ") + hipGetErrorString(e);
            return PT_ERR_HIP;
        }
    }
    rc = PT_OK;
    GuideState &gs = ctx->guide;
    if (bg.train)  // a training render decides the gate afresh (views start guided)
        hipLaunchKernelGGL(k_guide_gate_reset, dim3((uint32_t)std::min<size_t>(
                               (nviews * kGateDoublesPerView + 255) / 256, 1024)), dim3(256), 0, ctx->stream,
                           b.d_views, nviews, b.d_gate_acc, b.d_gate_off);
    uint32_t executed = 0;
    std::vector<BatchViewDev> gated_views;
    for (size_t k = 0; k < plan.size() && rc == PT_OK; ++k) {
        LaunchDesc &hl = b.h_launch[k];
        const bool train_here = bg.train && hl.train_frames > 0;
        if (train_here) {
            gs.epoch += 1u;
            hl.epoch = gs.epoch;
            hl.phase = gs.phase;
        }
        hipError_t e = hipMemcpyAsync(b.d_launch, &hl, sizeof(LaunchDesc), hipMemcpyHostToDevice, ctx->stream);
        if (e == hipSuccess) {
            if (graph) e = hipGraphLaunch(ge, ctx->stream);
            else {
                batch_enqueue_launch(ctx, sh, max_depth, spar, path_bits, fp64, bg, timing,
                                     (uint64_t)(executed + 1) << 16);
                e = hipGetLastError();
            }
        }
        bool decide = false;
        if (e == hipSuccess && gate && train_here) {
            uint32_t eval_px = 0;
            for (uint32_t en = 0; en < hl.n_entries; ++en) {
                if (hl.e[en].view & kEntryEval) eval_px = std::max(eval_px, hl.e[en].pixels);
                if (hl.e[en].view & kEntryDecide) decide = true;
            }
            if (eval_px) {
                const uint32_t bx = std::min((eval_px + 255u) / 256u, 64u);
                hipLaunchKernelGGL(k_eval_error, dim3(bx, hl.n_entries), dim3(256), 0, ctx->stream, ctx->d_out_l,
                                   (const LaunchDesc *)b.d_launch, (const BatchViewDev *)b.d_views, b.d_gate_acc);
            }
            if (decide)
                hipLaunchKernelGGL(k_guide_gate, dim3((hl.n_entries + 63) / 64), dim3(64), 0, ctx->stream,
                                   b.d_views, (const LaunchDesc *)b.d_launch, b.d_gate_acc, b.d_gate_off);
            e = hipGetLastError();
        }
        if (e == hipSuccess && train_here) {
            EventTimer tt(ctx, timing);
            // Samples from this launch's training frames, then the optimizer
            // steps (Adam over the hash grids of the views that trained).
            uint32_t v_lo = 0xFFFFFFFFu, v_hi = 0;
            for (uint32_t en = 0; en < hl.n_entries; ++en)
                if (hl.e[en].view & kEntryTrain) {
                    const uint32_t v = hl.e[en].view & kEntryViewMask;
                    v_lo = std::min(v_lo, v);
                    v_hi = std::max(v_hi, v + 1u);
                }
            const GuideRecords r{gs.rec, gs.rec_cap, gs.cfg.train_stride, hl.phase, max_depth, hl.epoch};
            const uint32_t steps = gs.cfg.steps_per_frame * hl.train_frames;
            bool ok;
            if (bg.ring_n) {
                ok = gs.collect(ctx->d_out_l, hl.total_paths, r, ctx->error, bg.ring_n,
                                guide_sample_count(steps, gs.cfg.batch_size)) &&
                     gs.train_ring(steps, 0.f, bg.step, bg.ring_n, ctx->error);
                bg.step += steps;
            } else {
                ok = gs.ensure_samples(guide_sample_count(steps, gs.cfg.batch_size), ctx->error) &&
                     gs.collect(ctx->d_out_l, hl.total_paths, r, ctx->error) &&
                     gs.train(steps, 0.f, v_lo, v_hi, ctx->error);
            }
            if (!ok) {
                rc = PT_ERR_HIP;
                break;
            }
            gs.phase = (gs.phase + 1u) % gs.cfg.train_stride;
            ctx->stats.guide_steps += steps;
            tt.stop(kKindGuide);
        }
        if (e != hipSuccess) {
            ctx->error = std::string("pt_batch_render: ") + hipGetErrorString(e);
            rc = PT_ERR_HIP;
            break;
        }
        ++executed;
        plan_advance(ps, hl, b.views);
        if (decide && k + 1 < plan.size()) {
            // Views the gate switched off stop training now: read the
            // decisions and plan the remaining frames again.
            gated_views.resize(nviews);
            e = hipStreamSynchronize(ctx->stream);
            if (e == hipSuccess)
                e = hipMemcpy(gated_views.data(), b.d_views, nviews * sizeof(BatchViewDev), hipMemcpyDeviceToHost);
            if (e != hipSuccess) {
                ctx->error = std::string("pt_batch_render: ") + hipGetErrorString(e);
                rc = PT_ERR_HIP;
                break;
            }
            bool changed = false;
            for (uint32_t vi = 0; vi < nviews; ++vi)
                if (gated_views[vi].guide_off) changed |= plan_gate_off(ps, vi);
            if (changed) {
                batch_plan(b.views, sh.capacity, true, plan, ps, train_cap, gate);
                if ((rc = stage_plan()) != PT_OK) break;
                k = (size_t)-1;  // continue with the new plan's first launch
            }
        }
    }
    hipError_t e = hipStreamSynchronize(ctx->stream);
    if (ge) (void)hipGraphExecDestroy(ge);
    if (g) (void)hipGraphDestroy(g);
    if (timing) collect_event_times(ctx);
    if (rc != PT_OK) return rc;
    if (e != hipSuccess) {
        ctx->error = std::string("pt_batch_render: ") + hipGetErrorString(e);
        return PT_ERR_HIP;
    }
    if (gate) {
        PT_HIP(hipMemcpy(&b.stats.guide_gated_off, b.d_gate_off, sizeof(uint32_t), hipMemcpyDeviceToHost));
        gated_views.resize(b.views.size());
        PT_HIP(hipMemcpy(gated_views.data(), b.d_views, gated_views.size() * sizeof(BatchViewDev),
                         hipMemcpyDeviceToHost));
    }
    uint64_t paths = 0;
    for (size_t vi = 0; vi < b.views.size(); ++vi) {
        BatchView &v = b.views[vi];
        paths += (uint64_t)v.width * v.height * v.spp;
        const uint32_t trained = training_frames(v.spp, p.guide_train_spp);
        // A view that kept its guide restarted its accumulation after training
        // (when it had frames after training).
        if (gate && trained >= kGateMinFrames && trained < v.spp && !gated_views[vi].guide_off)
            v.sample_cnt = v.spp - trained;
        else
            v.sample_cnt += v.spp;
        v.random_seed += v.spp;
    }
    *ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    if (launches) *launches = executed;
    b.stats.paths += paths;
    return PT_OK;
}

int batch_reset_films(pt_ctx *ctx) {
    for (auto &v : ctx->batch.views) {
        const size_t npix = (size_t)v.width * v.height;
        v.sample_cnt = 0;
        v.random_seed = 0;
        PT_HIP(hipMemsetAsync(v.d_accum, 0, npix * sizeof(float4), ctx->stream));
        PT_HIP(hipMemsetAsync(v.d_frame, 0, npix * sizeof(float4), ctx->stream));
        PT_HIP(hipMemsetAsync(v.d_accum64, 0, npix * 3 * sizeof(double), ctx->stream));
    }
    PT_HIP(hipStreamSynchronize(ctx->stream));
    return PT_OK;
}

BatchShape batch_shape(pt_ctx *ctx, const pt_batch_params &p) {
    const BatchState &b = ctx->batch;
    BatchShape sh;
    sh.capacity = p.paths_per_launch ? p.paths_per_launch : (b.tuned ? b.tuning.paths_per_launch : kBatchDefaultPaths);
    sh.shade_block = p.shade_block ? p.shade_block : (b.tuned ? b.tuning.shade_block : 256u);
    sh.train_paths = p.train_paths_per_launch;
    sh.shade_bpc = p.shade_blocks_per_cu ? p.shade_blocks_per_cu : (b.tuned ? b.tuning.shade_blocks_per_cu : 8u);
    sh.trace_bpc = p.trace_blocks_per_cu ? p.trace_blocks_per_cu
                                         : (b.tuned ? b.tuning.trace_blocks_per_cu : ctx->trace_blocks_per_cu);
    const uint32_t path_bits = batch_path_bits(b.views.size());
    sh.capacity = std::min(sh.capacity, 1u << path_bits);
    // The traversal spill area is sized for the resident grid.
    sh.trace_bpc = std::max(1u, std::min(sh.trace_bpc, ctx->trace_blocks_per_cu));
    sh.shade_bpc = std::max(1u, sh.shade_bpc);
    return sh;
}

}  // namespace

namespace {
// A host allocation that fails inside an entry point is reported as
// PT_ERR_OOM (with the entry point named in pt_last_error) instead of
// escaping the C interface.
template <typename F>
int guard_oom(pt_ctx *ctx, const char *who, F &&body) {
    try {
        return body();
    } catch (const std::bad_alloc &) {
        if (ctx) ctx->error = std::string(who) + ": out of host memory";
        return PT_ERR_OOM;
    }
}
}  // namespace

extern "C" {

int pt_create(int hip_device, pt_ctx **out) {
    if (!out) return PT_ERR_INVALID_ARG;
    *out = nullptr;
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess || n == 0) {
        g_create_error = std::string("no HIP device available: ") + hipGetErrorString(e);
        return PT_ERR_HIP;
    }
    if (hip_device < 0 || hip_device >= n) {
        g_create_error = "device index out of range";
        return PT_ERR_INVALID_ARG;
    }
    pt_ctx *ctx = new (std::nothrow) pt_ctx();
    if (!ctx) return PT_ERR_OOM;
    ctx->device = hip_device;
    if (hipSetDevice(hip_device) != hipSuccess || hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess) {
        g_create_error = "failed to set device / create stream";
        delete ctx;
        return PT_ERR_HIP;
    }
    if (hipStreamCreateWithFlags(&ctx->stream2, hipStreamNonBlocking) != hipSuccess ||
        hipStreamCreateWithFlags(&ctx->stream3, hipStreamNonBlocking) != hipSuccess ||
        hipStreamCreateWithFlags(&ctx->stream4, hipStreamNonBlocking) != hipSuccess) {
        g_create_error = "failed to create the chain streams";
        for (hipStream_t st : {ctx->stream, ctx->stream2, ctx->stream3})
            if (st) (void)hipStreamDestroy(st);
        delete ctx;
        return PT_ERR_HIP;
    }
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, hip_device) == hipSuccess && prop.multiProcessorCount > 0)
        ctx->num_cu = prop.multiProcessorCount;
    {
        // Persistent traversal grids: exactly the blocks of each kernel that
        // can be resident.
        auto resident = [](const void *kernel, uint32_t &out) {
            int n = 0;
            if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, kernel, kTraceBlock, 0) == hipSuccess && n > 0)
                out = (uint32_t)n;
        };
        resident(reinterpret_cast<const void *>(k_trace<false, true>), ctx->trace_blocks_per_cu);
        resident(reinterpret_cast<const void *>(k_trace<false, false>), ctx->trace_blocks_per_cu_tri);
        resident(reinterpret_cast<const void *>(k_shadow<false, true>), ctx->shadow_blocks_per_cu);
        resident(reinterpret_cast<const void *>(k_shadow<false, false>), ctx->shadow_blocks_per_cu_tri);
        int c = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&c, k_guide_infer, kInferBlock, 0) == hipSuccess && c > 0)
            ctx->infer_blocks_per_cu = (uint32_t)c;
        c = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&c, k_shade<false, kFeatAll>, 256, 0) == hipSuccess && c > 0)
            ctx->shade_blocks_per_cu = (uint32_t)c;
        c = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&c, k_shade_guided<false>, 256, 0) == hipSuccess && c > 0)
            ctx->shade_guided_blocks_per_cu = (uint32_t)c;

    }
    {
        const hipError_t e = resolve_scratch(hip_device);
        if (e != hipSuccess) {
            g_create_error = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
            (void)hipStreamDestroy(ctx->stream);
            delete ctx;
            return PT_ERR_HIP;
        }
    }
    if (hipMalloc((void **)&ctx->d_stat, 5 * sizeof(unsigned long long)) != hipSuccess ||
        hipMemset(ctx->d_stat, 0, 5 * sizeof(unsigned long long)) != hipSuccess) {
        g_create_error = "hipMalloc failed";
        (void)hipStreamDestroy(ctx->stream);
        delete ctx;
        return PT_ERR_HIP;
    }
    *out = ctx;
    return PT_OK;
}

void pt_destroy(pt_ctx *ctx) {
    if (!ctx) return;
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    for (hipStream_t st : {ctx->stream2, ctx->stream3, ctx->stream4})
        if (st) (void)hipStreamSynchronize(st);
    ctx->guide.release();
    batch_free_all(ctx);
    free_scene(ctx);
    free_queues(ctx);
    dfree(ctx->d_gout);
    dfree(ctx->d_spill);
    dfree(ctx->d_stat);
    dfree(ctx->d_refit_err);
    dfree(ctx->d_count_hist);
    for (hipEvent_t e : ctx->event_pool) (void)hipEventDestroy(e);
    for (hipEvent_t e : ctx->dep_events) (void)hipEventDestroy(e);
    for (hipStream_t st : {ctx->stream4, ctx->stream3, ctx->stream2, ctx->stream})
        if (st) (void)hipStreamDestroy(st);
    delete ctx;
}

const char *pt_last_error(const pt_ctx *ctx) { return ctx ? ctx->error.c_str() : g_create_error.c_str(); }

static int pt_set_scene_unguarded(pt_ctx *ctx, const pt_scene_desc *s);
int pt_set_scene(pt_ctx *ctx, const pt_scene_desc *s) {
    return guard_oom(ctx, "pt_set_scene", [&] { return pt_set_scene_unguarded(ctx, s); });
}
static int pt_set_scene_unguarded(pt_ctx *ctx, const pt_scene_desc *s) {
    if (!ctx) return PT_ERR_INVALID_ARG;
    SceneHostData d;
    int rc = convert_scene(s, "pt_set_scene", d, ctx->error);
    if (rc != PT_OK) return rc;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    free_scene(ctx);
    // The guide's position normalization follows the scene bounds.
    ctx->guide.release();
    ctx->h_pos4.swap(d.h_pos4);
    ctx->h_nrm4.swap(d.h_nrm4);
    ctx->h_uv2.swap(d.h_uv2);
    ctx->h_tri4.swap(d.h_tri4);
    ctx->h_shapes.swap(d.h_shapes);
    ctx->h_spheres.swap(d.h_spheres);
    ctx->h_sph4.swap(d.h_sph4);
    ctx->h_materials.swap(d.h_materials);
    ctx->h_emitters.swap(d.h_emitters);
    ctx->h_cdf.swap(d.h_cdf);
    ctx->shape_host.swap(d.shape_host);
    ctx->vert_owner.swap(d.vert_owner);
    ctx->h_pos0.assign(s->positions, s->positions + 3 * (size_t)s->num_vertices);
    if (s->normals) ctx->h_nrm0.assign(s->normals, s->normals + 3 * (size_t)s->num_vertices);
    else ctx->h_nrm0.clear();
    ctx->h_sph_m0.resize(12 * (size_t)s->num_spheres);
    for (uint32_t k = 0; k < s->num_spheres; ++k)
        std::memcpy(&ctx->h_sph_m0[12 * (size_t)k], s->spheres[k].to_world, 12 * sizeof(float));
    ctx->has_env = s->env_type != PT_ENV_NONE;
    ctx->bvh_dirty = false;
    ctx->env = d.env;
    ctx->camera = d.camera;
    ctx->width = d.width;
    ctx->height = d.height;
    ctx->scene_max_depth = d.max_depth;
    ctx->num_tris = d.num_tris;
    ctx->num_spheres = d.num_spheres;
    ctx->shade_feat = d.feat;
    for (int k = 0; k < 3; ++k) {
        ctx->scene_lo[k] = d.lo[k];
        ctx->scene_hi[k] = d.hi[k];
    }
    const uint32_t nv = s->num_vertices;

    if ((rc = upload(ctx, ctx->d_tri, ctx->h_tri4.data(), ctx->h_tri4.size())) != PT_OK) return rc;
    if ((rc = upload(ctx, ctx->d_pos, reinterpret_cast<const float4 *>(ctx->h_pos4.data()), (size_t)nv)) != PT_OK) return rc;
    if (s->normals &&
        (rc = upload(ctx, ctx->d_nrm, reinterpret_cast<const float4 *>(ctx->h_nrm4.data()), (size_t)nv)) != PT_OK)
        return rc;
    if (s->texcoords &&
        (rc = upload(ctx, ctx->d_uv, reinterpret_cast<const float2 *>(ctx->h_uv2.data()), (size_t)nv)) != PT_OK)
        return rc;
    if ((rc = upload(ctx, ctx->d_shapes, ctx->h_shapes.data(), ctx->h_shapes.size())) != PT_OK) return rc;
    if ((rc = upload(ctx, ctx->d_spheres, ctx->h_spheres.data(), ctx->h_spheres.size())) != PT_OK) return rc;
    if ((rc = upload(ctx, ctx->d_sph4, reinterpret_cast<const float4 *>(ctx->h_sph4.data()), ctx->h_sph4.size() / 4)) !=
        PT_OK)
        return rc;
    if ((rc = upload(ctx, ctx->d_materials, ctx->h_materials.data(), ctx->h_materials.size())) != PT_OK) return rc;
    if ((rc = upload(ctx, ctx->d_emitters, ctx->h_emitters.data(), ctx->h_emitters.size())) != PT_OK) return rc;
    if ((rc = upload(ctx, ctx->d_cdf, ctx->h_cdf.data(), ctx->h_cdf.size())) != PT_OK) return rc;
    if ((rc = upload(ctx, ctx->d_bitmaps, d.h_bitmaps.data(), d.h_bitmaps.size())) != PT_OK) return rc;
    if ((rc = upload(ctx, ctx->d_texels, reinterpret_cast<const float4 *>(d.h_texels.data()), d.h_texels.size() / 4)) !=
        PT_OK)
        return rc;
    if ((rc = upload(ctx, ctx->d_env_tables, d.h_env_tables.data(), d.h_env_tables.size())) != PT_OK) return rc;
    env_table_pointers(ctx->env, ctx->d_env_tables);

    const size_t npix = (size_t)ctx->width * ctx->height;
    PT_HIP(hipMalloc((void **)&ctx->d_accum, npix * sizeof(float4)));
    PT_HIP(hipMalloc((void **)&ctx->d_frame, npix * sizeof(float4)));
    PT_HIP(hipMalloc((void **)&ctx->d_accum64, npix * 3 * sizeof(double)));
    PT_HIP(hipMalloc((void **)&ctx->d_albedo, npix * 3 * sizeof(float)));
    PT_HIP(hipMalloc((void **)&ctx->d_normal, npix * 3 * sizeof(float)));
    PT_HIP(hipMalloc((void **)&ctx->d_test, npix * sizeof(float)));
    PT_HIP(hipMalloc((void **)&ctx->d_prim, npix * sizeof(uint32_t)));
    PT_HIP(hipMemset(ctx->d_accum, 0, npix * sizeof(float4)));
    PT_HIP(hipMemset(ctx->d_frame, 0, npix * sizeof(float4)));
    PT_HIP(hipMemset(ctx->d_accum64, 0, npix * 3 * sizeof(double)));
    PT_HIP(hipMemset(ctx->d_albedo, 0, npix * 3 * sizeof(float)));
    PT_HIP(hipMemset(ctx->d_normal, 0, npix * 3 * sizeof(float)));
    PT_HIP(hipMemset(ctx->d_test, 0, npix * sizeof(float)));
    PT_HIP(hipMemset(ctx->d_prim, 0xFF, npix * sizeof(uint32_t)));
    ctx->has_scene = true;
    ctx->has_bvh = false;
    ctx->sample_cnt = 0;
    ctx->random_seed = 0;
    ctx->stats = pt_stats{};
    PT_HIP(hipMemset(ctx->d_stat, 0, 5 * sizeof(unsigned long long)));
    return PT_OK;
}

static int pt_build_bvh_unguarded(pt_ctx *ctx, const pt_bvh_opts *opts);
int pt_build_bvh(pt_ctx *ctx, const pt_bvh_opts *opts) {
    return guard_oom(ctx, "pt_build_bvh", [&] { return pt_build_bvh_unguarded(ctx, opts); });
}
static int pt_build_bvh_unguarded(pt_ctx *ctx, const pt_bvh_opts *opts) {
    if (!ctx) return PT_ERR_INVALID_ARG;
    if (!ctx->has_scene) {
        ctx->error = "pt_build_bvh: no scene";
        return PT_ERR_NO_SCENE;
    }
    (void)hipSetDevice(ctx->device);
    BvhBuildOptions bo;
    if (opts && opts->num_bins) bo.num_bins = (int)opts->num_bins;
    if (opts && opts->threads) bo.threads = (int)opts->threads;
    std::vector<WideNode> nodes;
    std::vector<LeafTri> tris;
    std::string err;
    if (!build_wide_bvh(ctx->h_pos4.data(), (uint32_t)(ctx->h_pos4.size() / 4), ctx->h_tri4.data(), ctx->num_tris,
                        ctx->h_sph4.data(), ctx->num_spheres, bo, nodes, tris, ctx->bvh_stats, err)) {
        ctx->error = "pt_build_bvh: " + err;
        return PT_ERR_INVALID_ARG;
    }
    int rc;
    if ((rc = upload(ctx, ctx->d_nodes, nodes.data(), nodes.size())) != PT_OK) return rc;
    if ((rc = upload(ctx, ctx->d_leaf, tris.data(), tris.size())) != PT_OK) return rc;
    if ((rc = ensure_spill(ctx)) != PT_OK) return rc;
    // Node levels for refits: internal children of node i are child_base ..
    // child_base + popcount(imask) - 1, one level below.
    {
        std::vector<uint32_t> order, next, cur{0u};
        ctx->level_off.assign(1, 0u);
        while (!cur.empty()) {
            next.clear();
            for (uint32_t i : cur) {
                order.push_back(i);
                const uint32_t ni = (uint32_t)__builtin_popcount(nodes[i].imask);
                for (uint32_t j = 0; j < ni; ++j) next.push_back(nodes[i].child_base + j);
            }
            ctx->level_off.push_back((uint32_t)order.size());
            cur.swap(next);
        }
        if ((rc = upload(ctx, ctx->d_levels, order.data(), order.size())) != PT_OK) return rc;
        dfree(ctx->d_node_box);
        PT_HIP(hipMalloc((void **)&ctx->d_node_box, nodes.size() * 2 * sizeof(float4)));
        if (!ctx->d_refit_err) PT_HIP(hipMalloc((void **)&ctx->d_refit_err, sizeof(uint32_t)));
    }
    ctx->num_nodes = (uint32_t)nodes.size();
    ctx->num_leaf = (uint32_t)tris.size();
    ctx->bvh_dirty = false;
    ctx->has_bvh = true;
    return PT_OK;
}

int pt_set_camera(pt_ctx *ctx, const float s2c[16], const float c2w[16]) {
    if (!ctx || !s2c || !c2w) return PT_ERR_INVALID_ARG;
    for (int r = 0; r < 4; ++r) {
        ctx->camera.s2c[r] = f4{s2c[4 * r], s2c[4 * r + 1], s2c[4 * r + 2], s2c[4 * r + 3]};
        ctx->camera.c2w[r] = f4{c2w[4 * r], c2w[4 * r + 1], c2w[4 * r + 2], c2w[4 * r + 3]};
    }
    return pt_reset_accumulation(ctx);
}

int pt_set_random_seed(pt_ctx *ctx, uint32_t frame) {
    if (!ctx) return PT_ERR_INVALID_ARG;
    ctx->random_seed = frame;
    return PT_OK;
}

int pt_reset_accumulation(pt_ctx *ctx) {
    if (!ctx) return PT_ERR_INVALID_ARG;
    ctx->sample_cnt = 0;
    ctx->random_seed = 0;
    ctx->stats.samples = 0;
    if (ctx->d_accum64) {
        (void)hipSetDevice(ctx->device);
        PT_HIP(hipMemsetAsync(ctx->d_accum64, 0, (size_t)ctx->width * ctx->height * 3 * sizeof(double), ctx->stream));
        PT_HIP(hipStreamSynchronize(ctx->stream));
    }
    return PT_OK;
}

static int pt_set_shape_transform_unguarded(pt_ctx *ctx, uint32_t shape, const float m[12]);
int pt_set_shape_transform(pt_ctx *ctx, uint32_t shape, const float m[12]) {
    return guard_oom(ctx, "pt_set_shape_transform", [&] { return pt_set_shape_transform_unguarded(ctx, shape, m); });
}
static int pt_set_shape_transform_unguarded(pt_ctx *ctx, uint32_t shape, const float m[12]) {
    if (!ctx || !m) return PT_ERR_INVALID_ARG;
    if (!ctx->has_scene) {
        ctx->error = "pt_set_shape_transform: no scene";
        return PT_ERR_NO_SCENE;
    }
    if (shape >= ctx->shape_host.size()) {
        ctx->error = "pt_set_shape_transform: shape index out of range";
        return PT_ERR_INVALID_ARG;
    }
    for (int i = 0; i < 12; ++i)
        if (!std::isfinite(m[i])) {
            ctx->error = "pt_set_shape_transform: non-finite transform";
            return PT_ERR_INVALID_ARG;
        }
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    const ShapeHost &sh = ctx->shape_host[shape];
    int rc;
    if (sh.sphere) {
        // New placement m * M0 of each sphere (3x4 composition, float).
        std::vector<float> placed(12 * (size_t)sh.count);
        for (uint32_t j = 0; j < sh.count; ++j) {
            const float *a = &ctx->h_sph_m0[12 * (size_t)(sh.first + j)];
            float *o = &placed[12 * (size_t)j];
            for (int r = 0; r < 3; ++r) {
                for (int c = 0; c < 3; ++c)
                    o[4 * r + c] = m[4 * r] * a[c] + m[4 * r + 1] * a[4 + c] + m[4 * r + 2] * a[8 + c];
                o[4 * r + 3] = m[4 * r] * a[3] + m[4 * r + 1] * a[7] + m[4 * r + 2] * a[11] + m[4 * r + 3];
            }
            if (!is_similarity(o)) {
                ctx->error = "pt_set_shape_transform: a sphere must stay a similarity transform";
                return PT_ERR_UNSUPPORTED;
            }
        }
        for (uint32_t j = 0; j < sh.count; ++j) {
            const uint32_t k = sh.first + j;
            sphere_center_radius(&placed[12 * (size_t)j], &ctx->h_sph4[4 * (size_t)k]);
            sphere_inverse_rows(&placed[12 * (size_t)j], ctx->h_spheres[k]);
            if (sh.emitter) {
                const int32_t ei = ctx->h_shapes[shape].emitter_offset + (int32_t)j;
                ctx->h_emitters[ei] = sphere_emitter(&ctx->h_sph4[4 * (size_t)k], sh.radiance, sh.radiance_weight);
            }
        }
        if (sh.count) {
            PT_HIP(hipMemcpy(ctx->d_spheres + sh.first, &ctx->h_spheres[sh.first], sh.count * sizeof(SphereRec),
                             hipMemcpyHostToDevice));
            PT_HIP(hipMemcpy(ctx->d_sph4 + sh.first, &ctx->h_sph4[4 * (size_t)sh.first], sh.count * sizeof(float4),
                             hipMemcpyHostToDevice));
        }
    } else {
        if (sh.shared) {
            ctx->error = "pt_set_shape_transform: the shape shares vertices with another shape";
            return PT_ERR_UNSUPPORTED;
        }
        // Normals by the inverse transpose of the 3x3 part (cofactors / det).
        const double a00 = m[0], a01 = m[1], a02 = m[2], a10 = m[4], a11 = m[5], a12 = m[6];
        const double a20 = m[8], a21 = m[9], a22 = m[10];
        const double c00 = a11 * a22 - a12 * a21, c01 = a02 * a21 - a01 * a22, c02 = a01 * a12 - a02 * a11;
        const double c10 = a12 * a20 - a10 * a22, c11 = a00 * a22 - a02 * a20, c12 = a02 * a10 - a00 * a12;
        const double c20 = a10 * a21 - a11 * a20, c21 = a01 * a20 - a00 * a21, c22 = a00 * a11 - a01 * a10;
        const double det = a00 * c00 + a01 * c10 + a02 * c20;
        if (!(det != 0.0)) {
            ctx->error = "pt_set_shape_transform: singular transform";
            return PT_ERR_INVALID_ARG;
        }
        // (A^-1)^T: row i = column i of A^-1 = (c0i, c1i, c2i) / det.
        const float it[9] = {(float)(c00 / det), (float)(c10 / det), (float)(c20 / det),
                             (float)(c01 / det), (float)(c11 / det), (float)(c21 / det),
                             (float)(c02 / det), (float)(c12 / det), (float)(c22 / det)};
        const bool normals = !ctx->h_nrm0.empty();
        for (uint32_t v = sh.vmin; sh.count && v <= sh.vmax; ++v) {
            if (ctx->vert_owner[v] != shape) continue;
            const float *p = &ctx->h_pos0[3 * (size_t)v];
            float *q = &ctx->h_pos4[4 * (size_t)v];
            q[0] = m[0] * p[0] + m[1] * p[1] + m[2] * p[2] + m[3];
            q[1] = m[4] * p[0] + m[5] * p[1] + m[6] * p[2] + m[7];
            q[2] = m[8] * p[0] + m[9] * p[1] + m[10] * p[2] + m[11];
            if (normals) {
                const float *n = &ctx->h_nrm0[3 * (size_t)v];
                float *o = &ctx->h_nrm4[4 * (size_t)v];
                o[0] = it[0] * n[0] + it[1] * n[1] + it[2] * n[2];
                o[1] = it[3] * n[0] + it[4] * n[1] + it[5] * n[2];
                o[2] = it[6] * n[0] + it[7] * n[1] + it[8] * n[2];
            }
        }
        ShapeRec &rec = ctx->h_shapes[shape];
        rec.flags = sh.flags0 ^ (det < 0.0 ? (uint32_t)kShapeNegFaceNormal : 0u);
        if (sh.count) {
            const size_t nv = (size_t)sh.vmax - sh.vmin + 1;
            PT_HIP(hipMemcpy(ctx->d_pos + sh.vmin, &ctx->h_pos4[4 * (size_t)sh.vmin], nv * sizeof(float4),
                             hipMemcpyHostToDevice));
            if (normals && ctx->d_nrm)
                PT_HIP(hipMemcpy(ctx->d_nrm + sh.vmin, &ctx->h_nrm4[4 * (size_t)sh.vmin], nv * sizeof(float4),
                                 hipMemcpyHostToDevice));
        }
        PT_HIP(hipMemcpy(ctx->d_shapes + shape, &rec, sizeof(ShapeRec), hipMemcpyHostToDevice));
        if (sh.emitter) {
            const bool has_n = normals && (rec.flags & kShapeHasNormals);
            for (uint32_t j = 0; j < sh.count; ++j) {
                const uint32_t *t = &ctx->h_tri4[4 * (size_t)(sh.first + j)];
                f3 p[3], n[3];
                f2 uv[3];
                for (int k = 0; k < 3; ++k) {
                    const float *q = &ctx->h_pos4[4 * (size_t)t[k]];
                    p[k] = mk3(q[0], q[1], q[2]);
                    n[k] = has_n ? mk3(ctx->h_nrm4[4 * (size_t)t[k]], ctx->h_nrm4[4 * (size_t)t[k] + 1],
                                       ctx->h_nrm4[4 * (size_t)t[k] + 2])
                                 : mk3(0.f);
                    uv[k] = ctx->h_uv2.empty() ? mk2(0.f, 0.f)
                                               : mk2(ctx->h_uv2[2 * (size_t)t[k]], ctx->h_uv2[2 * (size_t)t[k] + 1]);
                }
                ctx->h_emitters[rec.emitter_offset + (int32_t)j] =
                    tri_emitter(p, n, has_n, (rec.flags & kShapeNegFaceNormal) != 0, uv, sh.radiance,
                                sh.radiance_weight);
            }
        }
    }
    if (sh.emitter) {
        float env_prob;
        emitter_probabilities(ctx->h_emitters, ctx->has_env, ctx->h_cdf, env_prob);
        if ((rc = upload(ctx, ctx->d_emitters, ctx->h_emitters.data(), ctx->h_emitters.size())) != PT_OK) return rc;
        if ((rc = upload(ctx, ctx->d_cdf, ctx->h_cdf.data(), ctx->h_cdf.size())) != PT_OK) return rc;
    }
    ctx->bvh_dirty = true;
    return pt_reset_accumulation(ctx);
}

namespace {
int denoise_params(const pt_denoise_params *p, DenoiseParams &out, uint32_t &flags, std::string &err) {
    pt_denoise_params d{};
    if (p) d = *p;
    if (d.iterations > 10 || !(d.sigma_luminance >= 0.f) || !(d.sigma_normal >= 0.f) || !(d.sigma_albedo >= 0.f)) {
        err = "denoise: iterations must be <= 10 and sigmas non-negative";
        return PT_ERR_INVALID_ARG;
    }
    out.iterations = d.iterations ? d.iterations : 5u;
    out.sigma_luminance = d.sigma_luminance > 0.f ? d.sigma_luminance : 4.f;
    out.sigma_normal = d.sigma_normal > 0.f ? d.sigma_normal : 64.f;
    out.sigma_albedo = d.sigma_albedo > 0.f ? d.sigma_albedo : 0.1f;
    flags = d.flags;
    return PT_OK;
}

int ensure_dn_scratch(pt_ctx *ctx, size_t px) {
    if (ctx->d_dn_scratch && ctx->dn_scratch_px >= px) return PT_OK;
    dfree(ctx->d_dn_scratch);
    PT_HIP(hipMalloc((void **)&ctx->d_dn_scratch, 2 * px * sizeof(float4)));
    ctx->dn_scratch_px = px;
    return PT_OK;
}
}  // namespace

int pt_denoise(pt_ctx *ctx, const pt_denoise_params *params) {
    if (!ctx) return PT_ERR_INVALID_ARG;
    if (!ctx->has_scene) {
        ctx->error = "pt_denoise: no scene";
        return PT_ERR_NO_SCENE;
    }
    DenoiseParams prm;
    uint32_t flags;
    int rc = denoise_params(params, prm, flags, ctx->error);
    if (rc != PT_OK) return rc;
    (void)hipSetDevice(ctx->device);
    const size_t px = (size_t)ctx->width * ctx->height;
    if ((rc = ensure_dn_scratch(ctx, px)) != PT_OK) return rc;
    PT_HIP(denoise_launch(ctx->stream, ctx->width, ctx->height, ctx->d_accum,
                          (flags & PT_DENOISE_ALBEDO) ? ctx->d_albedo : nullptr,
                          (flags & PT_DENOISE_NORMAL) ? ctx->d_normal : nullptr, ctx->d_frame, ctx->d_dn_scratch, prm));
    PT_HIP(hipStreamSynchronize(ctx->stream));
    return PT_OK;
}

static int pt_denoise_image_unguarded(pt_ctx *ctx, uint32_t w, uint32_t h, const float *rgba, const float *albedo, const float *normal,
                     float *out, const pt_denoise_params *params);
int pt_denoise_image(pt_ctx *ctx, uint32_t w, uint32_t h, const float *rgba, const float *albedo, const float *normal,
                     float *out, const pt_denoise_params *params) {
    return guard_oom(ctx, "pt_denoise_image", [&] {
        return pt_denoise_image_unguarded(ctx, w, h, rgba, albedo, normal, out, params);
    });
}
static int pt_denoise_image_unguarded(pt_ctx *ctx, uint32_t w, uint32_t h, const float *rgba, const float *albedo, const float *normal,
                     float *out, const pt_denoise_params *params) {
    if (!ctx || !rgba || !out || w == 0 || h == 0 || (uint64_t)w * h > (1ull << 28)) return PT_ERR_INVALID_ARG;
    DenoiseParams prm;
    uint32_t flags;
    int rc = denoise_params(params, prm, flags, ctx->error);
    if (rc != PT_OK) return rc;
    (void)hipSetDevice(ctx->device);
    const size_t px = (size_t)w * h;
    float4 *d_in = nullptr, *d_out = nullptr, *d_scr = nullptr;
    float *d_a = nullptr, *d_n = nullptr;
    const bool ua = albedo && (flags & PT_DENOISE_ALBEDO), un = normal && (flags & PT_DENOISE_NORMAL);
    auto cleanup = [&]() {
        dfree(d_in);
        dfree(d_out);
        dfree(d_scr);
        dfree(d_a);
        dfree(d_n);
    };
    hipError_t e = hipMalloc((void **)&d_in, px * sizeof(float4));
    if (e == hipSuccess) e = hipMalloc((void **)&d_out, px * sizeof(float4));
    if (e == hipSuccess) e = hipMalloc((void **)&d_scr, 2 * px * sizeof(float4));
    if (e == hipSuccess && ua) e = hipMalloc((void **)&d_a, px * 3 * sizeof(float));
    if (e == hipSuccess && un) e = hipMalloc((void **)&d_n, px * 3 * sizeof(float));
    if (e == hipSuccess) e = hipMemcpy(d_in, rgba, px * sizeof(float4), hipMemcpyHostToDevice);
    if (e == hipSuccess && ua) e = hipMemcpy(d_a, albedo, px * 3 * sizeof(float), hipMemcpyHostToDevice);
    if (e == hipSuccess && un) e = hipMemcpy(d_n, normal, px * 3 * sizeof(float), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = denoise_launch(ctx->stream, w, h, d_in, d_a, d_n, d_out, d_scr, prm);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e == hipSuccess) e = hipMemcpy(out, d_out, px * sizeof(float4), hipMemcpyDeviceToHost);
    cleanup();
    if (e != hipSuccess) {
        ctx->error = std::string("pt_denoise_image: ") + hipGetErrorString(e);
        return PT_ERR_HIP;
    }
    return PT_OK;
}

int pt_refit_bvh(pt_ctx *ctx) {
    if (!ctx) return PT_ERR_INVALID_ARG;
    if (!ctx->has_bvh) {
        ctx->error = "pt_refit_bvh: BVH not built";
        return PT_ERR_NO_BVH;
    }
    return bvh_refit(ctx);
}

int pt_set_counting(pt_ctx *ctx, int enable) {
    if (!ctx) return PT_ERR_INVALID_ARG;
    ctx->counting = enable != 0;
    return PT_OK;
}

static int pt_render_unguarded(pt_ctx *ctx, const pt_render_params *p);
int pt_render(pt_ctx *ctx, const pt_render_params *p) {
    return guard_oom(ctx, "pt_render", [&] { return pt_render_unguarded(ctx, p); });
}
static int pt_render_unguarded(pt_ctx *ctx, const pt_render_params *p) {
    if (!ctx || !p) return PT_ERR_INVALID_ARG;
    if (!ctx->has_scene) {
        ctx->error = "pt_render: no scene";
        return PT_ERR_NO_SCENE;
    }
    if (!ctx->has_bvh) {
        ctx->error = "pt_render: BVH not built";
        return PT_ERR_NO_BVH;
    }
    {
        const int rr = bvh_refit(ctx);
        if (rr != PT_OK) return rr;
    }
    if (p->mis_mode > PT_MIS_CONSISTENT) {
        ctx->error = "pt_render: unknown MIS mode";
        return PT_ERR_INVALID_ARG;
    }
    const bool guide_infer = (p->flags & PT_RENDER_GUIDE) != 0;
    const bool guide_train = (p->flags & PT_RENDER_GUIDE_TRAIN) != 0;
    const bool guided = guide_infer || guide_train;
    if (guided && (!ctx->guide.enabled || ctx->guide.for_batch)) {
        ctx->error = "pt_render: guided rendering needs pt_guide_enable";
        return PT_ERR_NO_GUIDE;
    }
    if (guided && p->mis_mode != PT_MIS_CONSISTENT) {
        ctx->error = "pt_render: guided rendering needs PT_MIS_CONSISTENT";
        return PT_ERR_INVALID_ARG;
    }
    (void)hipSetDevice(ctx->device);
    const auto wall0 = std::chrono::steady_clock::now();
    const uint32_t npix = ctx->width * ctx->height;
    const uint32_t max_depth = p->max_depth ? p->max_depth : ctx->scene_max_depth;
    if (max_depth > 250) {
        ctx->error = "pt_render: max_depth above 250";
        return PT_ERR_INVALID_ARG;
    }
    uint32_t per_launch = p->spp_per_launch;
    // Default launch width: ~128M paths in flight (about 30 GB of queues, a
    // tenth of the HBM), which amortizes the latency-bound tail of the last
    // bounces over many frames (terrain 1080p, 64 spp: 748, 794 and 820
    // Msamples/s at 16, 32 and 64 frames per launch).
    // While training the guide, launches are kept short so the network is
    // updated every few frames: 2M paths, or up to four frames (8M paths) of
    // a larger film -- one-frame launches cost each 1080p frame a whole
    // launch's latency-bound last bounces (door room, 256 training frames:
    // 4.1-4.2 s in one-frame launches, 3.7 s in four-frame launches, with the
    // same final error).
    if (per_launch == 0) {
        const uint32_t train_paths =
            std::max(2u * 1024u * 1024u, std::min(8u * 1024u * 1024u, 4u * npix));
        const uint32_t paths = guide_train ? train_paths : 128u * 1024u * 1024u;
        per_launch = std::max(1u, std::min(64u, paths / std::max(1u, npix)));
    }
    per_launch = std::min(per_launch, std::max(1u, p->spp));
    if ((uint64_t)per_launch * npix > 0x7FFFFFFFull) {
        ctx->error = "pt_render: spp_per_launch * pixels exceeds the queue index range";
        return PT_ERR_CAPACITY;
    }
    int rc = ensure_queues(ctx, per_launch * npix, max_depth);
    if (rc != PT_OK) return rc;
    if (ctx->counting) PT_HIP(hipMemsetAsync(ctx->d_stat, 0, 4 * sizeof(unsigned long long), ctx->stream));
    // Network queries of this call only (batched renders count into the same
    // device word).
    if (guide_infer) PT_HIP(hipMemsetAsync(ctx->d_stat + 4, 0, sizeof(unsigned long long), ctx->stream));

    const bool timing = (p->flags & PT_RENDER_TIMING) != 0;
    const SceneView sv = scene_view(ctx);
    if ((rc = ensure_dep_events(ctx, 4)) != PT_OK) return rc;  // fork and the joins of the chains
    ShadeParams sp;
    sp.max_depth = max_depth;
    sp.rr_depth = p->rr_depth;
    sp.mis_mode = p->mis_mode;
    sp.aov = (p->flags & PT_RENDER_AOV) ? 1u : 0u;
    const uint32_t tg = trace_grid(ctx), sg = shade_grid(ctx, guided, guide_train);

    const uint32_t batches = (p->spp + per_launch - 1) / per_launch;
    const uint32_t hist_elems = std::max(1u, batches) * 4u * ctx->num_counters;  // one set per chain
    if (hist_elems > ctx->count_hist_slots || !ctx->d_count_hist) {
        dfree(ctx->d_count_hist);
        PT_HIP(hipMalloc((void **)&ctx->d_count_hist, (size_t)hist_elems * sizeof(uint32_t)));
        ctx->count_hist_slots = hist_elems;
    }

    GuideShadeArgs ga;
    std::memset(&ga, 0, sizeof(ga));
    GuideView gview;
    std::memset(&gview, 0, sizeof(gview));
    if (guided) {
        gview = ctx->guide.view();
        ga.fixed_c = std::min(ctx->guide.cfg.fixed_selection, 1.f);
    }
    if (guide_infer) {
        if (ctx->gout_cap < per_launch * npix) {
            dfree(ctx->d_gout);
            PT_HIP(hipMalloc((void **)&ctx->d_gout, (size_t)per_launch * npix * 6 * sizeof(uint4)));
            ctx->gout_cap = per_launch * npix;
        }
        ga.gout = ctx->d_gout;
    }
    if (guide_train) {
        const pt_guide_config &gc = ctx->guide.cfg;
        const uint64_t slots = ((uint64_t)per_launch * npix + gc.train_stride - 1) / gc.train_stride;
        const uint64_t records = slots * max_depth;
        if (records > 0xFFFFFFFFull / 2) {
            ctx->error = "pt_render: too many guide training records per launch (raise train_stride)";
            return PT_ERR_CAPACITY;
        }
        if (!ctx->guide.ensure_records((uint32_t)records, ctx->error) ||
            !ctx->guide.ensure_samples(guide_sample_count((uint64_t)gc.steps_per_frame * per_launch, gc.batch_size),
                                       ctx->error))
            return PT_ERR_HIP;
    }

    // Chains: a launch's queue slots are split into up to kMaxChains
    // contiguous ranges of at least kChainPaths (10^6) paths, each traced, shaded and
    // shadow-tested on its own stream with its own part of the queues, its
    // own counters and spill area.  A chain's bounces serialize and every
    // kernel drains before the next starts; the other chains fill the idle
    // CUs meanwhile.  Terrain 1080p, one chain against four: 815 -> 970
    // Msamples/s at 20 frames per launch, 861 -> 1058-1071 at 64.  More
    // streams than the hardware queues (four) cost: two chains with a second
    // stream each for their shadow rays reached 885 and 967, three such 743.
    // Every path's work is unchanged, so is the image.
    uint32_t done = 0, batch = 0;
    uint32_t sample_cnt = ctx->sample_cnt, seed = ctx->random_seed;
    std::vector<uint32_t> chains_per_launch;
    const uint32_t gdepth = ctx->guide.cfg.guided_depth ? ctx->guide.cfg.guided_depth : max_depth;
    const hipStream_t chain_stream[kMaxChains] = {ctx->stream, ctx->stream2, ctx->stream3, ctx->stream4};
    while (done < p->spp) {
        const uint32_t frames = std::min(per_launch, p->spp - done);
        const uint32_t total = frames * npix;
        const uint32_t nchain = chain_count(total, (p->flags & PT_RENDER_ONE_CHAIN) != 0);
        // With at least as many frames as chains, a chain takes a share of the
        // frames of every pixel, so the chains run through the same pixels at
        // the same time and share L2 (terrain 1080p, four chains of 5 frames:
        // 968 Msamples/s at 20 frames per launch, against 914 for four bands
        // of pixels in all 20 frames); otherwise a band of pixels.
        ChainPart part[kMaxChains];
        chain_parts(frames, npix, nchain, part);
        AovBuffers aov{nullptr, nullptr, nullptr, nullptr, frames - 1, npix};
        if (sp.aov) {
            aov.albedo = ctx->d_albedo;
            aov.normal = ctx->d_normal;
            aov.test = ctx->d_test;
            aov.prim = ctx->d_prim;
        }
        if (guide_train) {
            GuideState &gs = ctx->guide;
            gs.epoch += 1u;
            ga.rs.rec = reinterpret_cast<f4 *>(gs.rec);
            ga.rs.capacity = gs.rec_cap;
            ga.rs.stride = gs.cfg.train_stride;
            ga.rs.phase = gs.phase;
            ga.rs.max_depth = max_depth;
            ga.rs.epoch = gs.epoch;
        }
        struct Chain {
            uint32_t frame0, frames, pix0, pixels;  // frames x tile-order pixel places of the launch
            uint32_t slot0, paths;                  // its slots [slot0, slot0 + paths) of the queues
            hipStream_t stream;
            uint32_t *counters, *ext, *shd, *twork, *swork;
            PathQueue q[2];
            ShadowQueue sq;
            DeviceBvh bvh;
            uint4 *gout;
        } ch[kMaxChains];
        for (uint32_t c = 0; c < nchain; ++c) {
            Chain &k = ch[c];
            k.frame0 = part[c].frame0;
            k.frames = part[c].frames;
            k.pix0 = part[c].pix0;
            k.pixels = part[c].pixels;
            k.slot0 = part[c].slot0;
            k.paths = part[c].paths;
            k.stream = chain_stream[c];
            k.counters = ctx->d_counters + (size_t)c * ctx->num_counters;
            k.ext = k.counters;                          // extension queue sizes
            k.shd = k.counters + (max_depth + 2);        // shadow queue sizes
            k.twork = k.counters + 2 * (max_depth + 2);  // trace work cursors
            k.swork = k.counters + 3 * (max_depth + 2);  // shadow work cursors
            k.q[0] = queue_at(ctx->q[0], k.slot0);
            k.q[1] = queue_at(ctx->q[1], k.slot0);
            k.sq = shadow_queue_at(ctx->sq, k.slot0);
            k.bvh = device_bvh_area(ctx, c);
            k.gout = ctx->d_gout ? ctx->d_gout + 6 * (size_t)k.slot0 : nullptr;
        }
        // The other chains start after everything before on the pass stream.
        if (nchain > 1) PT_HIP(hipEventRecord(ctx->dep_events[0], ctx->stream));
        for (uint32_t c = 1; c < nchain; ++c) PT_HIP(hipStreamWaitEvent(ch[c].stream, ctx->dep_events[0], 0));
        // Timer groups: a bounce's kernels of all chains time as one launch.
        const uint64_t launch_group = (uint64_t)(batch + 1) << 16;
        for (uint32_t c = 0; c < nchain; ++c) {
            const Chain &k = ch[c];
            EventTimer t(ctx, timing, k.stream, launch_group | 0xFFFFu);
            const uint32_t grid = std::min<uint32_t>((k.paths + 255) / 256, (uint32_t)ctx->num_cu * 32u);
            hipLaunchKernelGGL(k_raygen, dim3(grid), dim3(256), 0, k.stream, k.q[0], ctx->camera, ctx->width,
                               ctx->height, k.frames, seed, k.counters, ctx->num_counters, k.frame0, k.pix0, k.pixels);
            t.stop(kKindRaygen);
        }
        int cur = 0;
        for (uint32_t b = 0; b < max_depth; ++b) {
            const int nxt = cur ^ 1;
            // The guide is sampled at vertices b < guided_depth; the vertex of
            // the last bounce never samples a direction.
            const bool infer_here = guide_infer && b < gdepth && b + 1 < max_depth;
            const uint64_t group = launch_group | (b + 1);
            for (uint32_t c = 0; c < nchain; ++c) {
                const Chain &k = ch[c];
                {
                    EventTimer t(ctx, timing, k.stream, group);
                    hipLaunchKernelGGL(trace_kernel(ctx->counting, ctx->num_spheres > 0), dim3(tg), dim3(kTraceBlock),
                                       0, k.stream, k.q[cur], k.bvh, k.ext + b, k.twork + b, ctx->d_stat,
                                       (const uint32_t *)nullptr, 31u);
                    t.stop(kKindTrace);
                }
                GuideShadeArgs gac = ga;
                gac.gout = infer_here ? k.gout : nullptr;
                if (infer_here) {
                    EventTimer t(ctx, timing, k.stream, group);
                    hipLaunchKernelGGL(k_guide_infer, dim3(ctx->num_cu * ctx->infer_blocks_per_cu), dim3(kInferBlock),
                                       0, k.stream, gview, k.q[cur], k.ext + b, reinterpret_cast<uint2 *>(k.gout),
                                       ctx->d_stat + 4, 31u, InferScene{sv.tri, sv.shapes, sv.spheres, sv.materials});
                    t.stop(kKindGuideInfer);
                }
                {
                    EventTimer t(ctx, timing, k.stream, group);
                    if (guided)
                        hipLaunchKernelGGL(k_shade_guided<false>, dim3(sg), dim3(256), 0, k.stream, sv, sp, k.q[cur],
                                           k.ext + b, k.q[nxt], k.ext + b + 1, k.sq, k.shd + b, ctx->d_out_l, aov, gac,
                                           (const BatchViewDev *)nullptr, 31u, (const LaunchDesc *)nullptr);
                    else
                        hipLaunchKernelGGL(shade_kernel(false, ctx->shade_feat), dim3(sg), dim3(256), 0, k.stream, sv,
                                           sp, k.q[cur], k.ext + b, k.q[nxt], k.ext + b + 1, k.sq, k.shd + b,
                                           ctx->d_out_l, aov, (const BatchViewDev *)nullptr, 31u);
                    t.stop(kKindShade);
                }
                {
                    EventTimer t(ctx, timing, k.stream, group);
                    hipLaunchKernelGGL(shadow_kernel(ctx->counting, ctx->num_spheres > 0),
                                       dim3(shadow_grid(ctx, ctx->num_spheres > 0)), dim3(kTraceBlock), 0, k.stream,
                                       k.sq, k.bvh, k.shd + b, k.swork + b, k.q[nxt], ctx->d_out_l, ctx->d_stat + 2,
                                       (const uint32_t *)nullptr, 31u);
                    t.stop(kKindShadow);
                }
            }
            cur = nxt;
        }
        for (uint32_t c = 1; c < nchain; ++c) {
            PT_HIP(hipEventRecord(ctx->dep_events[c], ch[c].stream));
            PT_HIP(hipStreamWaitEvent(ctx->stream, ctx->dep_events[c], 0));
        }
        {
            EventTimer t(ctx, timing);
            const uint32_t grid = std::min<uint32_t>((npix + 255) / 256, (uint32_t)ctx->num_cu * 16u);
            hipLaunchKernelGGL(k_accumulate, dim3(grid), dim3(256), 0, ctx->stream, ctx->d_out_l, ctx->d_accum,
                               ctx->d_frame, (p->flags & PT_RENDER_FP64_ACCUM) ? ctx->d_accum64 : nullptr, npix, frames,
                               sample_cnt, p->accumulate);
            t.stop(kKindAccumulate);
        }
        if (guide_train) {
            // Training samples from this launch's paths, then the optimizer
            // steps; the next launch samples with the updated network.
            EventTimer t(ctx, timing);
            GuideState &gs = ctx->guide;
            GuideRecords r{gs.rec, gs.rec_cap, ga.rs.stride, ga.rs.phase, max_depth, gs.epoch};
            if (!gs.ensure_samples(guide_sample_count((uint64_t)gs.cfg.steps_per_frame * frames, gs.cfg.batch_size),
                                   ctx->error) ||
                !gs.collect(ctx->d_out_l, total, r, ctx->error) ||
                !gs.train(gs.cfg.steps_per_frame * frames, 0.f, ctx->error))
                return PT_ERR_HIP;
            gs.phase = (gs.phase + 1u) % gs.cfg.train_stride;
            ctx->stats.guide_steps += (uint64_t)gs.cfg.steps_per_frame * frames;
            t.stop(kKindGuide);
        }
        PT_HIP(hipGetLastError());
        // Keep this launch's queue counters (both chains' sets); read back once
        // after the last launch.
        PT_HIP(hipMemcpyAsync(ctx->d_count_hist + (size_t)batch * kMaxChains * ctx->num_counters, ctx->d_counters,
                              (size_t)nchain * ctx->num_counters * sizeof(uint32_t), hipMemcpyDeviceToDevice,
                              ctx->stream));
        chains_per_launch.push_back(nchain);
        if (p->accumulate) sample_cnt += frames;
        seed += frames;
        done += frames;
        ++batch;
    }
    std::vector<uint32_t> hist((size_t)batch * kMaxChains * ctx->num_counters);
    PT_HIP(hipMemcpyAsync(hist.data(), ctx->d_count_hist, hist.size() * sizeof(uint32_t), hipMemcpyDeviceToHost,
                          ctx->stream));
    PT_HIP(hipStreamSynchronize(ctx->stream));
    for (uint32_t k = 0; k < batch; ++k)
        for (uint32_t c = 0; c < chains_per_launch[k]; ++c) {
            const uint32_t *cn = hist.data() + ((size_t)k * kMaxChains + c) * ctx->num_counters;
            for (uint32_t b = 0; b < max_depth; ++b) {
                ctx->stats.rays_extension += cn[b];
                ctx->stats.rays_shadow += cn[max_depth + 2 + b];
            }
        }
    if (guide_infer) {
        unsigned long long nq = 0;
        PT_HIP(hipMemcpy(&nq, ctx->d_stat + 4, sizeof(nq), hipMemcpyDeviceToHost));
        ctx->stats.guide_queries += nq;
    }
    if (guide_train) {
        unsigned long long kept = 0;
        PT_HIP(hipMemcpy(&kept, &ctx->guide.ctl->total_samples, sizeof(kept), hipMemcpyDeviceToHost));
        ctx->stats.guide_samples = kept;
    }
    ctx->sample_cnt = sample_cnt;
    ctx->random_seed = seed;
    ctx->stats.samples += (uint64_t)p->spp * npix;
    if (timing) collect_event_times(ctx);
    if (ctx->counting) {
        unsigned long long st[4];
        PT_HIP(hipMemcpy(st, ctx->d_stat, sizeof(st), hipMemcpyDeviceToHost));
        ctx->stats.node_visits += st[0];
        ctx->stats.tri_tests += st[1];
        ctx->stats.shadow_node_visits += st[2];
        ctx->stats.shadow_tri_tests += st[3];
    }
    ctx->stats.ms_total += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - wall0).count();
    return PT_OK;
}

// Carves one call's buffers out of the scratch array: counters, the
// traversal spill area for `spill_threads` lanes, then two path queues, the
// shadow queue and the radiance slots for as many paths as fit.
struct NoAllocLayout {
    PathQueue q[2];
    ShadowQueue sq;
    float4 *out_l;
    uint32_t *counters;
    uint2 *spill;
    uint32_t capacity;  // paths
};

bool noalloc_layout(unsigned char *base, uint32_t num_counters, uint32_t spill_threads, NoAllocLayout &L) {
    size_t off = 0;
    auto take = [&](size_t bytes) {
        unsigned char *p = base + off;
        off += (bytes + 255) & ~(size_t)255;
        return p;
    };
    L.counters = reinterpret_cast<uint32_t *>(take((size_t)num_counters * sizeof(uint32_t)));
    L.spill = reinterpret_cast<uint2 *>(take((size_t)spill_threads * kSpillEntries * sizeof(uint2)));
    if (off >= kNoAllocScratchBytes) return false;
    // 10 path-queue arrays, 3 shadow arrays and the radiance slots: 14
    // float4 per path, each array rounded up to 256 bytes.
    const size_t left = kNoAllocScratchBytes - off - 14 * 256;
    L.capacity = (uint32_t)std::min<size_t>(left / (14 * sizeof(float4)), 0x7FFFFFFFull);
    const size_t n = L.capacity;
    for (auto &q : L.q) {
        q.ray0 = reinterpret_cast<float4 *>(take(n * sizeof(float4)));
        q.ray1 = reinterpret_cast<float4 *>(take(n * sizeof(float4)));
        q.tl = reinterpret_cast<float4 *>(take(n * sizeof(float4)));
        q.lr = reinterpret_cast<float4 *>(take(n * sizeof(float4)));
        q.hit = reinterpret_cast<float4 *>(take(n * sizeof(float4)));
    }
    L.sq.s0 = reinterpret_cast<float4 *>(take(n * sizeof(float4)));
    L.sq.s1 = reinterpret_cast<float4 *>(take(n * sizeof(float4)));
    L.sq.s2 = reinterpret_cast<float4 *>(take(n * sizeof(float4)));
    L.out_l = reinterpret_cast<float4 *>(take(n * sizeof(float4)));
    return off <= kNoAllocScratchBytes;
}

// Traversal blocks per CU of the allocation-free render (its spill area is
// sized for this grid).
constexpr uint32_t kNoAllocTraceBlocksPerCu = 8;

int pt_render_noalloc(pt_ctx *ctx, const pt_render_params *p) {
    if (!ctx || !p) return PT_ERR_INVALID_ARG;
    if (!ctx->has_scene) {
        ctx->error = "pt_render_noalloc: no scene";
        return PT_ERR_NO_SCENE;
    }
    if (!ctx->has_bvh) {
        ctx->error = "pt_render_noalloc: BVH not built";
        return PT_ERR_NO_BVH;
    }
    if (ctx->bvh_dirty) {
        ctx->error = "pt_render_noalloc: shape updates pending (call pt_refit_bvh first)";
        return PT_ERR_INVALID_ARG;
    }
    if (p->mis_mode > PT_MIS_CONSISTENT) {
        ctx->error = "pt_render_noalloc: unknown MIS mode";
        return PT_ERR_INVALID_ARG;
    }
    if (p->flags & (PT_RENDER_GUIDE | PT_RENDER_GUIDE_TRAIN | PT_RENDER_TIMING)) {
        ctx->error = "pt_render_noalloc: guided rendering and timing are not available without allocation";
        return PT_ERR_UNSUPPORTED;
    }
    const int dev = ctx->device;
    unsigned char *base = (dev >= 0 && dev < kMaxScratchDevices) ? g_scratch_ptr[dev] : nullptr;
    if (!base) {
        ctx->error = "pt_render_noalloc: scratch memory of this device was not resolved";
        return PT_ERR_HIP;
    }
    const uint32_t npix = ctx->width * ctx->height;
    const uint32_t max_depth = p->max_depth ? p->max_depth : ctx->scene_max_depth;
    if (max_depth > 250) {
        ctx->error = "pt_render_noalloc: max_depth above 250";
        return PT_ERR_INVALID_ARG;
    }
    const uint32_t num_counters = 4 * (max_depth + 2);
    const uint32_t tbpc = std::min(kNoAllocTraceBlocksPerCu, ctx->num_spheres > 0 ? ctx->trace_blocks_per_cu
                                                                                 : ctx->trace_blocks_per_cu_tri);
    const uint32_t sbpc = std::min(kNoAllocTraceBlocksPerCu, ctx->num_spheres > 0 ? ctx->shadow_blocks_per_cu
                                                                                 : ctx->shadow_blocks_per_cu_tri);
    const uint32_t tg = (uint32_t)ctx->num_cu * tbpc, shg = (uint32_t)ctx->num_cu * sbpc;
    NoAllocLayout L;
    if (!noalloc_layout(base, num_counters, std::max(tg, shg) * kTraceBlock, L) || L.capacity < npix) {
        ctx->error = "pt_render_noalloc: one frame of " + std::to_string(npix) + " paths exceeds the " +
                     std::to_string(kNoAllocScratchBytes >> 20) + " MiB scratch (" + std::to_string(L.capacity) +
                     " paths of 224 bytes)";
        return PT_ERR_CAPACITY;
    }
    uint32_t per_launch = p->spp_per_launch ? p->spp_per_launch : L.capacity / npix;
    per_launch = std::max(1u, std::min(per_launch, std::max(1u, p->spp)));
    if ((uint64_t)per_launch * npix > L.capacity) {
        ctx->error = "pt_render_noalloc: spp_per_launch * pixels exceeds the scratch capacity of " +
                     std::to_string(L.capacity) + " paths";
        return PT_ERR_CAPACITY;
    }
    (void)hipSetDevice(ctx->device);
    const SceneView sv = scene_view(ctx);
    const DeviceBvh bvh{ctx->d_nodes, ctx->d_leaf, L.spill};
    ShadeParams sp;
    sp.max_depth = max_depth;
    sp.rr_depth = p->rr_depth;
    sp.mis_mode = p->mis_mode;
    sp.aov = (p->flags & PT_RENDER_AOV) ? 1u : 0u;
    const uint32_t sg = shade_grid(ctx, false, false);
    uint32_t done = 0, sample_cnt = ctx->sample_cnt, seed = ctx->random_seed;
    while (done < p->spp) {
        const uint32_t frames = std::min(per_launch, p->spp - done);
        const uint32_t total = frames * npix;
        uint32_t *ext = L.counters, *shd = L.counters + (max_depth + 2);
        uint32_t *twork = L.counters + 2 * (max_depth + 2), *swork = L.counters + 3 * (max_depth + 2);
        // k_raygen clears every counter before the first traversal reads them.
        const uint32_t rgrid = std::min<uint32_t>((total + 255) / 256, (uint32_t)ctx->num_cu * 32u);
        hipLaunchKernelGGL(k_raygen, dim3(rgrid), dim3(256), 0, ctx->stream, L.q[0], ctx->camera, ctx->width,
                           ctx->height, frames, seed, L.counters, num_counters, 0u, 0u, npix);
        AovBuffers aov{nullptr, nullptr, nullptr, nullptr, frames - 1, npix};
        if (sp.aov) {
            aov.albedo = ctx->d_albedo;
            aov.normal = ctx->d_normal;
            aov.test = ctx->d_test;
            aov.prim = ctx->d_prim;
        }
        int cur = 0;
        for (uint32_t b = 0; b < max_depth; ++b) {
            const int nxt = cur ^ 1;
            hipLaunchKernelGGL(trace_kernel(false, ctx->num_spheres > 0), dim3(tg), dim3(kTraceBlock), 0, ctx->stream,
                               L.q[cur], bvh, ext + b, twork + b, ctx->d_stat, (const uint32_t *)nullptr, 31u);
            hipLaunchKernelGGL(shade_kernel(false, ctx->shade_feat), dim3(sg), dim3(256), 0, ctx->stream, sv, sp,
                               L.q[cur], ext + b, L.q[nxt], ext + b + 1, L.sq, shd + b, L.out_l, aov,
                               (const BatchViewDev *)nullptr, 31u);
            hipLaunchKernelGGL(shadow_kernel(false, ctx->num_spheres > 0), dim3(shg), dim3(kTraceBlock), 0,
                               ctx->stream, L.sq, bvh, shd + b, swork + b, L.q[nxt], L.out_l, ctx->d_stat + 2,
                               (const uint32_t *)nullptr, 31u);
            cur = nxt;
        }
        const uint32_t agrid = std::min<uint32_t>((npix + 255) / 256, (uint32_t)ctx->num_cu * 16u);
        hipLaunchKernelGGL(k_accumulate, dim3(agrid), dim3(256), 0, ctx->stream, L.out_l, ctx->d_accum, ctx->d_frame,
                           (p->flags & PT_RENDER_FP64_ACCUM) ? ctx->d_accum64 : nullptr, npix, frames, sample_cnt,
                           p->accumulate);
        PT_HIP(hipGetLastError());
        if (p->accumulate) sample_cnt += frames;
        seed += frames;
        done += frames;
    }
    ctx->sample_cnt = sample_cnt;
    ctx->random_seed = seed;
    ctx->stats.samples += (uint64_t)p->spp * npix;
    return PT_OK;
}

int pt_read_buffer(pt_ctx *ctx, int which, void *dst, size_t bytes) {
    if (!ctx || !dst) return PT_ERR_INVALID_ARG;
    if (!ctx->has_scene) {
        ctx->error = "pt_read_buffer: no scene";
        return PT_ERR_NO_SCENE;
    }
    (void)hipSetDevice(ctx->device);
    const size_t npix = (size_t)ctx->width * ctx->height;
    const void *src = nullptr;
    size_t need = 0;
    switch (which) {
        case PT_BUF_FRAME: src = ctx->d_frame; need = npix * 16; break;
        case PT_BUF_ACCUM: src = ctx->d_accum; need = npix * 16; break;
        case PT_BUF_ALBEDO: src = ctx->d_albedo; need = npix * 12; break;
        case PT_BUF_NORMAL: src = ctx->d_normal; need = npix * 12; break;
        case PT_BUF_TEST: src = ctx->d_test; need = npix * 4; break;
        case PT_BUF_PRIM_ID: src = ctx->d_prim; need = npix * 4; break;
        case PT_BUF_ACCUM64: src = ctx->d_accum64; need = npix * 24; break;
        default: ctx->error = "pt_read_buffer: unknown buffer"; return PT_ERR_INVALID_ARG;
    }
    if (bytes < need) {
        ctx->error = "pt_read_buffer: destination too small";
        return PT_ERR_INVALID_ARG;
    }
    PT_HIP(hipStreamSynchronize(ctx->stream));
    PT_HIP(hipMemcpy(dst, src, need, hipMemcpyDeviceToHost));
    return PT_OK;
}

int pt_get_stats(pt_ctx *ctx, pt_stats *out) {
    if (!ctx || !out) return PT_ERR_INVALID_ARG;
    *out = ctx->stats;
    out->bvh_nodes = ctx->bvh_stats.wide_nodes;
    out->bvh_leaf_tris = ctx->bvh_stats.leaf_tris;
    out->bvh_depth = ctx->bvh_stats.depth;
    out->bvh_build_ms = ctx->bvh_stats.build_seconds * 1000.0;
    return PT_OK;
}

int pt_bandwidth_probe(pt_ctx *ctx, uint64_t bytes, uint32_t iters, double *gbps) {
    if (!ctx || !gbps || bytes < 4096 || iters == 0) return PT_ERR_INVALID_ARG;
    (void)hipSetDevice(ctx->device);
    const size_t n = (size_t)(bytes / sizeof(float4));
    float4 *a = nullptr, *b = nullptr;
    PT_HIP(hipMalloc((void **)&a, n * sizeof(float4)));
    if (hipMalloc((void **)&b, n * sizeof(float4)) != hipSuccess) {
        (void)hipFree(a);
        ctx->error = "pt_bandwidth_probe: hipMalloc failed";
        return PT_ERR_HIP;
    }
    int rc = PT_OK;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    do {
        if (hipMemsetAsync(a, 0, n * sizeof(float4), ctx->stream) != hipSuccess ||
            hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
            rc = PT_ERR_HIP;
            break;
        }
        const uint32_t grid = (uint32_t)ctx->num_cu * 32u;
        hipLaunchKernelGGL(k_copy, dim3(grid), dim3(256), 0, ctx->stream, a, b, n);  // warm
        (void)hipEventRecord(e0, ctx->stream);
        for (uint32_t i = 0; i < iters; ++i)
            hipLaunchKernelGGL(k_copy, dim3(grid), dim3(256), 0, ctx->stream, a, b, n);
        (void)hipEventRecord(e1, ctx->stream);
        if (hipEventSynchronize(e1) != hipSuccess) {
            rc = PT_ERR_HIP;
            break;
        }
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, e0, e1);
        *gbps = 2.0 * (double)n * sizeof(float4) * iters / (ms * 1e-3) / 1e9;
    } while (0);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    (void)hipFree(a);
    (void)hipFree(b);
    if (rc != PT_OK) ctx->error = "pt_bandwidth_probe: HIP call failed";
    return rc;
}

static int pt_trace_rays_unguarded(pt_ctx *ctx, const float *rays, uint32_t count, int any_hit, void *hits);
int pt_trace_rays(pt_ctx *ctx, const float *rays, uint32_t count, int any_hit, void *hits) {
    return guard_oom(ctx, "pt_trace_rays", [&] { return pt_trace_rays_unguarded(ctx, rays, count, any_hit, hits); });
}
static int pt_trace_rays_unguarded(pt_ctx *ctx, const float *rays, uint32_t count, int any_hit, void *hits) {
    if (!ctx || (!rays && count) || (!hits && count)) return PT_ERR_INVALID_ARG;
    if (!ctx->has_bvh) {
        ctx->error = "pt_trace_rays: BVH not built";
        return PT_ERR_NO_BVH;
    }
    {
        const int rr = bvh_refit(ctx);
        if (rr != PT_OK) return rr;
    }
    if (count == 0) return PT_OK;
    (void)hipSetDevice(ctx->device);
    float *d_rays = nullptr;
    uint4 *d_hits = nullptr;
    PT_HIP(hipMalloc((void **)&d_rays, (size_t)count * 8 * sizeof(float)));
    if (hipMalloc((void **)&d_hits, (size_t)count * sizeof(uint4)) != hipSuccess) {
        (void)hipFree(d_rays);
        ctx->error = "pt_trace_rays: hipMalloc failed";
        return PT_ERR_HIP;
    }
    int rc = PT_OK;
    do {
        if (hipMemcpy(d_rays, rays, (size_t)count * 8 * sizeof(float), hipMemcpyHostToDevice) != hipSuccess) {
            rc = PT_ERR_HIP;
            break;
        }
        const DeviceBvh bvh = device_bvh(ctx);
        const uint32_t grid = std::min<uint32_t>((count + kTraceBlock - 1) / kTraceBlock, trace_grid(ctx));
        hipLaunchKernelGGL(trace_rays_kernel(any_hit != 0, ctx->num_spheres > 0), dim3(grid), dim3(kTraceBlock), 0,
                           ctx->stream, d_rays, count, bvh, d_hits);
        if (hipGetLastError() != hipSuccess || hipStreamSynchronize(ctx->stream) != hipSuccess) {
            rc = PT_ERR_HIP;
            break;
        }
        if (hipMemcpy(hits, d_hits, (size_t)count * sizeof(uint4), hipMemcpyDeviceToHost) != hipSuccess) rc = PT_ERR_HIP;
    } while (0);
    (void)hipFree(d_rays);
    (void)hipFree(d_hits);
    if (rc != PT_OK) ctx->error = "pt_trace_rays: HIP call failed";
    return rc;
}

// ---- learned guiding ----------------------------------------------------------

namespace {

int guide_ready(pt_ctx *ctx, const char *who) {
    if (!ctx) return PT_ERR_INVALID_ARG;
    if (!ctx->guide.enabled) {
        ctx->error = std::string(who) + ": guide not enabled";
        return PT_ERR_NO_GUIDE;
    }
    (void)hipSetDevice(ctx->device);
    return PT_OK;
}

}  // namespace

static int pt_guide_enable_unguarded(pt_ctx *ctx, const pt_guide_config *config);
int pt_guide_enable(pt_ctx *ctx, const pt_guide_config *config) {
    return guard_oom(ctx, "pt_guide_enable", [&] { return pt_guide_enable_unguarded(ctx, config); });
}
static int pt_guide_enable_unguarded(pt_ctx *ctx, const pt_guide_config *config) {
    if (!ctx) return PT_ERR_INVALID_ARG;
    if (!ctx->has_scene) {
        ctx->error = "pt_guide_enable: no scene";
        return PT_ERR_NO_SCENE;
    }
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    pt_guide_config c;
    std::memset(&c, 0, sizeof(c));
    if (config) c = *config;
    if (!ctx->guide.init(c, ctx->scene_lo, ctx->scene_hi, 1u, ctx->stream, ctx->error)) {
        ctx->guide.release();
        return ctx->error.rfind("pt_guide_enable", 0) == 0 ? PT_ERR_INVALID_ARG : PT_ERR_HIP;
    }
    ctx->guide.for_batch = false;
    return PT_OK;
}

static int pt_guide_enable_batch_unguarded(pt_ctx *ctx, const pt_guide_config *config);
int pt_guide_enable_batch(pt_ctx *ctx, const pt_guide_config *config) {
    return guard_oom(ctx, "pt_guide_enable_batch", [&] { return pt_guide_enable_batch_unguarded(ctx, config); });
}
static int pt_guide_enable_batch_unguarded(pt_ctx *ctx, const pt_guide_config *config) {
    if (!ctx) return PT_ERR_INVALID_ARG;
    const BatchState &b = ctx->batch;
    if (b.views.empty()) {
        ctx->error = "pt_guide_enable_batch: no batch views";
        return PT_ERR_NO_SCENE;
    }
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    pt_guide_config c;
    std::memset(&c, 0, sizeof(c));
    if (config) c = *config;
    std::vector<float> lo(3 * b.views.size()), hi(3 * b.views.size());
    for (size_t v = 0; v < b.views.size(); ++v)
        for (int k = 0; k < 3; ++k) {
            lo[3 * v + k] = b.views[v].lo[k];
            hi[3 * v + k] = b.views[v].hi[k];
        }
    if (!ctx->guide.init(c, lo.data(), hi.data(), (uint32_t)b.views.size(), ctx->stream, ctx->error)) {
        ctx->guide.release();
        if (ctx->error.rfind("pt_guide_enable", 0) == 0) {
            ctx->error.insert(std::strlen("pt_guide_enable"), "_batch");
            return PT_ERR_INVALID_ARG;
        }
        return PT_ERR_HIP;
    }
    ctx->guide.for_batch = true;
    return PT_OK;
}

int pt_guide_query_grid(pt_ctx *ctx, uint32_t grid, const float *positions, uint32_t n, float *outputs) {
    int rc = guide_ready(ctx, "pt_guide_query_grid");
    if (rc != PT_OK) return rc;
    if (n && (!positions || !outputs)) {
        ctx->error = "pt_guide_query_grid: null buffer";
        return PT_ERR_INVALID_ARG;
    }
    if (grid >= ctx->guide.grids) {
        ctx->error = "pt_guide_query_grid: no such hash grid";
        return PT_ERR_INVALID_ARG;
    }
    return ctx->guide.query(positions, n, outputs, grid, ctx->error) ? PT_OK : PT_ERR_HIP;
}

int pt_guide_grid_bounds(pt_ctx *ctx, uint32_t grid, float lo[3], float inv[3]) {
    int rc = guide_ready(ctx, "pt_guide_grid_bounds");
    if (rc != PT_OK) return rc;
    if (!lo || !inv) return PT_ERR_INVALID_ARG;
    return ctx->guide.grid_box(grid, lo, inv, ctx->error) ? PT_OK : PT_ERR_INVALID_ARG;
}

int pt_guide_disable(pt_ctx *ctx) {
    if (!ctx) return PT_ERR_INVALID_ARG;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    ctx->guide.release();
    return PT_OK;
}

int pt_guide_get_info(pt_ctx *ctx, pt_guide_info *out) {
    if (!ctx || !out) return PT_ERR_INVALID_ARG;
    (void)hipSetDevice(ctx->device);
    return ctx->guide.info(*out, ctx->error) ? PT_OK : PT_ERR_HIP;
}

int pt_guide_query(pt_ctx *ctx, const float *positions, uint32_t n, float *outputs) {
    int rc = guide_ready(ctx, "pt_guide_query");
    if (rc != PT_OK) return rc;
    if (n && (!positions || !outputs)) {
        ctx->error = "pt_guide_query: null buffer";
        return PT_ERR_INVALID_ARG;
    }
    return ctx->guide.query(positions, n, outputs, 0u, ctx->error) ? PT_OK : PT_ERR_HIP;
}

static int guide_check_samples(pt_ctx *ctx, const char *who, const float *samples, uint32_t n) {
    if (!samples || n == 0 || n % 64u != 0u || n > (1u << 22)) {
        ctx->error = std::string(who) + ": samples must be a positive multiple of 64 (at most 2^22)";
        return PT_ERR_INVALID_ARG;
    }
    for (uint64_t i = 0; i < n; ++i) {
        const float *v = samples + 12 * i;
        for (int k = 0; k < 12; ++k)
            if (!std::isfinite(v[k])) {
                ctx->error = std::string(who) + ": non-finite sample value";
                return PT_ERR_INVALID_ARG;
            }
    }
    return PT_OK;
}

int pt_guide_train(pt_ctx *ctx, const float *samples, uint32_t n, float fp_norm, float *loss) {
    int rc = guide_ready(ctx, "pt_guide_train");
    if (rc != PT_OK) return rc;
    if ((rc = guide_check_samples(ctx, "pt_guide_train", samples, n)) != PT_OK) return rc;
    GuideState &g = ctx->guide;
    const uint32_t saved = g.cfg.batch_size;
    g.cfg.batch_size = n;
    bool ok = g.upload_samples(samples, n, ctx->error) && g.train(1, fp_norm > 0.f ? fp_norm : 1.f, ctx->error);
    g.cfg.batch_size = saved;
    if (!ok) return PT_ERR_HIP;
    ctx->stats.guide_steps += 1;
    pt_guide_info info;
    if (!g.info(info, ctx->error)) return PT_ERR_HIP;
    if (loss) *loss = info.last_loss;
    return PT_OK;
}

int pt_guide_gradients(pt_ctx *ctx, const float *samples, uint32_t n, float fp_norm, float *table_grad,
                       float *weight_grad, float *loss) {
    int rc = guide_ready(ctx, "pt_guide_gradients");
    if (rc != PT_OK) return rc;
    if ((rc = guide_check_samples(ctx, "pt_guide_gradients", samples, n)) != PT_OK) return rc;
    GuideState &g = ctx->guide;
    if (!g.upload_samples(samples, n, ctx->error)) return PT_ERR_HIP;
    if (!g.gradients(n, fp_norm > 0.f ? fp_norm : 1.f, table_grad, weight_grad, loss, ctx->error))
        return ctx->error.find("non-finite") != std::string::npos ? PT_ERR_INVALID_ARG : PT_ERR_HIP;
    return PT_OK;
}

int pt_guide_get_params(pt_ctx *ctx, float *table, float *weights) {
    int rc = guide_ready(ctx, "pt_guide_get_params");
    if (rc != PT_OK) return rc;
    return ctx->guide.get_params(table, weights, ctx->error) ? PT_OK : PT_ERR_HIP;
}

int pt_guide_read_samples(pt_ctx *ctx, float *samples, uint32_t max_samples, uint32_t *count) {
    int rc = guide_ready(ctx, "pt_guide_read_samples");
    if (rc != PT_OK) return rc;
    if (!count || (max_samples && !samples)) {
        ctx->error = "pt_guide_read_samples: null buffer";
        return PT_ERR_INVALID_ARG;
    }
    return ctx->guide.read_samples(samples, max_samples, *count, ctx->error) ? PT_OK : PT_ERR_HIP;
}

int pt_guide_set_params(pt_ctx *ctx, const float *table, const float *weights) {
    int rc = guide_ready(ctx, "pt_guide_set_params");
    if (rc != PT_OK) return rc;
    return ctx->guide.set_params(table, weights, ctx->error) ? PT_OK : PT_ERR_HIP;
}

// ---- batched views ------------------------------------------------------------

int pt_batch_clear(pt_ctx *ctx) {
    if (!ctx) return PT_ERR_INVALID_ARG;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    if (ctx->guide.for_batch) ctx->guide.release();
    batch_free_all(ctx);
    return PT_OK;
}

static int pt_batch_add_view_unguarded(pt_ctx *ctx, const pt_scene_desc *scene, uint32_t spp, uint32_t *view_index);
int pt_batch_add_view(pt_ctx *ctx, const pt_scene_desc *scene, uint32_t spp, uint32_t *view_index) {
    return guard_oom(ctx, "pt_batch_add_view", [&] {
        return pt_batch_add_view_unguarded(ctx, scene, spp, view_index);
    });
}
static int pt_batch_add_view_unguarded(pt_ctx *ctx, const pt_scene_desc *scene, uint32_t spp, uint32_t *view_index) {
    if (!ctx) return PT_ERR_INVALID_ARG;
    if (spp == 0) {
        ctx->error = "pt_batch_add_view: spp must be positive";
        return PT_ERR_INVALID_ARG;
    }
    BatchState &b = ctx->batch;
    if (b.views.size() >= 1024) {
        ctx->error = "pt_batch_add_view: at most 1024 views";
        return PT_ERR_CAPACITY;
    }
    SceneHostData d;
    int rc = convert_scene(scene, "pt_batch_add_view", d, ctx->error);
    if (rc != PT_OK) return rc;
    (void)hipSetDevice(ctx->device);
    const auto t0 = std::chrono::steady_clock::now();
    std::vector<WideNode> nodes;
    std::vector<LeafTri> leaf;
    BvhBuildStats bst;
    std::string err;
    if (!build_wide_bvh(d.h_pos4.data(), (uint32_t)(d.h_pos4.size() / 4), d.h_tri4.data(), d.num_tris,
                        d.h_sph4.data(), d.num_spheres, BvhBuildOptions(), nodes, leaf, bst, err)) {
        ctx->error = "pt_batch_add_view: " + err;
        return PT_ERR_INVALID_ARG;
    }
    const uint64_t nv0 = b.pos4.size() / 4, nt0 = b.tri4.size() / 4, nsph0 = b.spheres.size();
    const uint64_t node0 = b.nodes.size(), leaf0 = b.leaf.size();
    const uint64_t nv = d.h_pos4.size() / 4;
    if (nv0 + nv > 0xFFFFFFFFull || nt0 + d.num_tris >= kPrimSphere || nsph0 + d.num_spheres >= kPrimSphere ||
        node0 + nodes.size() > 0xFFFFFFFFull || leaf0 + leaf.size() > 0xFFFFFFFFull) {
        ctx->error = "pt_batch_add_view: batch exceeds 32-bit index ranges";
        return PT_ERR_CAPACITY;
    }
    const uint32_t shape0 = (uint32_t)b.shapes.size(), mat0 = (uint32_t)b.materials.size();
    const uint32_t em0 = (uint32_t)b.emitters.size();
    // Relocate into the shared arrays.
    for (auto &n : nodes) {
        n.child_base += (uint32_t)node0;
        n.tri_base += (uint32_t)leaf0;
    }
    for (auto &t : leaf)
        t.v0.w = u2f(f2u(t.v0.w) + (uint32_t)(leaf_is_sphere(t) ? nsph0 : nt0));
    b.nodes.insert(b.nodes.end(), nodes.begin(), nodes.end());
    b.leaf.insert(b.leaf.end(), leaf.begin(), leaf.end());
    b.feat |= d.feat;
    b.pos4.insert(b.pos4.end(), d.h_pos4.begin(), d.h_pos4.end());
    if (d.has_normals && !b.any_normals) b.nrm4.assign(nv0 * 4, 0.f);
    if (d.has_texcoords && !b.any_texcoords) b.uv2.assign(nv0 * 2, 0.f);
    b.any_normals = b.any_normals || d.has_normals;
    b.any_texcoords = b.any_texcoords || d.has_texcoords;
    if (b.any_normals) {
        if (d.has_normals) b.nrm4.insert(b.nrm4.end(), d.h_nrm4.begin(), d.h_nrm4.end());
        else b.nrm4.resize(b.nrm4.size() + nv * 4, 0.f);
    }
    if (b.any_texcoords) {
        if (d.has_texcoords) b.uv2.insert(b.uv2.end(), d.h_uv2.begin(), d.h_uv2.end());
        else b.uv2.resize(b.uv2.size() + nv * 2, 0.f);
    }
    for (size_t t = 0; t < d.num_tris; ++t) {
        for (int k = 0; k < 3; ++k) b.tri4.push_back(d.h_tri4[4 * t + k] + (uint32_t)nv0);
        b.tri4.push_back(d.h_tri4[4 * t + 3] + shape0);
    }
    for (ShapeRec r : d.h_shapes) {
        r.material += mat0;
        r.first_tri += (uint32_t)((r.flags & PT_SHAPE_SPHERE) ? nsph0 : nt0);  // emitter offsets stay view-local
        b.shapes.push_back(r);
    }
    for (SphereRec sp : d.h_spheres) {
        sp.shape += shape0;
        b.spheres.push_back(sp);
    }
    // Bitmap indices and texel offsets become batch-global.
    const uint32_t bmp0 = (uint32_t)b.bitmaps.size();
    const uint64_t texel0 = b.texels.size() / 4;
    for (MaterialRec &m : d.h_materials)
        for (TextureRec &t : m.tex)
            if (t.type == kTexBitmap) t.bitmap += bmp0;
    for (EmitterRec &e : d.h_emitters)
        if (e.radiance.type == kTexBitmap) e.radiance.bitmap += bmp0;
    for (BitmapRec &r : d.h_bitmaps) r.offset += texel0;
    if (d.env.type == kEmitterEnvMap) d.env.bitmap += bmp0;
    const size_t env0 = b.env_tables.size();
    b.bitmaps.insert(b.bitmaps.end(), d.h_bitmaps.begin(), d.h_bitmaps.end());
    b.texels.insert(b.texels.end(), d.h_texels.begin(), d.h_texels.end());
    b.env_tables.insert(b.env_tables.end(), d.h_env_tables.begin(), d.h_env_tables.end());
    b.materials.insert(b.materials.end(), d.h_materials.begin(), d.h_materials.end());
    b.emitters.insert(b.emitters.end(), d.h_emitters.begin(), d.h_emitters.end());
    b.cdf.insert(b.cdf.end(), d.h_cdf.begin(), d.h_cdf.end());
    if (ctx->guide.for_batch) {
        // The batch guide has one hash grid per view: a new view disables it.
        (void)hipStreamSynchronize(ctx->stream);
        ctx->guide.release();
    }
    BatchView v;
    v.width = d.width;
    v.height = d.height;
    v.spp = spp;
    for (int k = 0; k < 3; ++k) {
        v.lo[k] = d.lo[k];
        v.hi[k] = d.hi[k];
    }
    v.max_depth = d.max_depth;
    v.camera = d.camera;
    v.env = d.env;
    v.env_tables = env0;
    v.root = (uint32_t)node0;
    v.emitter_begin = em0;
    v.emitter_count = (uint32_t)d.h_emitters.size();
    v.tri_count = d.num_tris;
    const size_t npix = (size_t)v.width * v.height;
    PT_HIP(hipMalloc((void **)&v.d_accum, npix * sizeof(float4)));
    PT_HIP(hipMalloc((void **)&v.d_frame, npix * sizeof(float4)));
    PT_HIP(hipMalloc((void **)&v.d_accum64, npix * 3 * sizeof(double)));
    PT_HIP(hipMemset(v.d_accum, 0, npix * sizeof(float4)));
    PT_HIP(hipMemset(v.d_frame, 0, npix * sizeof(float4)));
    PT_HIP(hipMemset(v.d_accum64, 0, npix * 3 * sizeof(double)));
    b.views.push_back(v);
    b.dirty = true;
    b.stats.views = (uint32_t)b.views.size();
    b.stats.triangles += d.num_tris;
    b.stats.build_ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    if (view_index) *view_index = (uint32_t)b.views.size() - 1u;
    return PT_OK;
}

static int pt_batch_render_unguarded(pt_ctx *ctx, const pt_batch_params *params);
int pt_batch_render(pt_ctx *ctx, const pt_batch_params *params) {
    return guard_oom(ctx, "pt_batch_render", [&] { return pt_batch_render_unguarded(ctx, params); });
}
static int pt_batch_render_unguarded(pt_ctx *ctx, const pt_batch_params *params) {
    if (!ctx || !params) return PT_ERR_INVALID_ARG;
    BatchState &b = ctx->batch;
    if (b.views.empty()) {
        ctx->error = "pt_batch_render: no views";
        return PT_ERR_NO_SCENE;
    }
    if (params->mis_mode > PT_MIS_CONSISTENT || (params->shade_block != 0 && params->shade_block != 64 &&
                                                  params->shade_block != 128 && params->shade_block != 256)) {
        ctx->error = "pt_batch_render: bad mis_mode or shade_block";
        return PT_ERR_INVALID_ARG;
    }
    (void)hipSetDevice(ctx->device);
    if ((params->batch_flags & PT_BATCH_AUTOTUNE) && !b.tuned) {
        pt_batch_tuning t;
        int rc = pt_batch_autotune(ctx, params, &t);
        if (rc != PT_OK) return rc;
    }
    const BatchShape sh = batch_shape(ctx, *params);
    double ms = 0.0;
    uint32_t launches = 0;
    int rc = batch_run(ctx, *params, sh, (params->batch_flags & PT_BATCH_GRAPH) != 0, &ms, &launches);
    if (rc != PT_OK) return rc;
    b.stats.launches += launches;
    b.stats.ms_total += ms;
    return PT_OK;
}

int pt_batch_reset(pt_ctx *ctx) {
    if (!ctx) return PT_ERR_INVALID_ARG;
    (void)hipSetDevice(ctx->device);
    return batch_reset_films(ctx);
}

int pt_batch_read(pt_ctx *ctx, uint32_t view, int which, void *dst, size_t bytes) {
    if (!ctx || !dst) return PT_ERR_INVALID_ARG;
    BatchState &b = ctx->batch;
    if (view >= b.views.size()) {
        ctx->error = "pt_batch_read: no such view";
        return PT_ERR_INVALID_ARG;
    }
    const BatchView &v = b.views[view];
    const size_t npix = (size_t)v.width * v.height;
    const void *src = nullptr;
    size_t need = 0;
    switch (which) {
        case PT_BUF_FRAME: src = v.d_frame; need = npix * 16; break;
        case PT_BUF_ACCUM: src = v.d_accum; need = npix * 16; break;
        case PT_BUF_ACCUM64: src = v.d_accum64; need = npix * 24; break;
        default: ctx->error = "pt_batch_read: unsupported buffer"; return PT_ERR_INVALID_ARG;
    }
    if (bytes < need) {
        ctx->error = "pt_batch_read: destination too small";
        return PT_ERR_INVALID_ARG;
    }
    (void)hipSetDevice(ctx->device);
    PT_HIP(hipStreamSynchronize(ctx->stream));
    PT_HIP(hipMemcpy(dst, src, need, hipMemcpyDeviceToHost));
    return PT_OK;
}

static int pt_batch_autotune_unguarded(pt_ctx *ctx, const pt_batch_params *params, pt_batch_tuning *out);
int pt_batch_autotune(pt_ctx *ctx, const pt_batch_params *params, pt_batch_tuning *out) {
    return guard_oom(ctx, "pt_batch_autotune", [&] { return pt_batch_autotune_unguarded(ctx, params, out); });
}
static int pt_batch_autotune_unguarded(pt_ctx *ctx, const pt_batch_params *params, pt_batch_tuning *out) {
    if (!ctx || !params) return PT_ERR_INVALID_ARG;
    BatchState &b = ctx->batch;
    if (b.views.empty()) {
        ctx->error = "pt_batch_autotune: no views";
        return PT_ERR_NO_SCENE;
    }
    (void)hipSetDevice(ctx->device);
    // Probe: every view with a share of its frames, ~96M paths in all.
    uint64_t total = 0;
    uint32_t max_npix = 0;
    for (const auto &v : b.views) {
        total += (uint64_t)v.width * v.height * v.spp;
        max_npix = std::max(max_npix, v.width * v.height);
    }
    const double share = std::min(1.0, (96.0 * (1 << 20)) / (double)std::max<uint64_t>(total, 1));
    std::vector<uint32_t> saved(b.views.size());
    uint64_t probe_paths = 0;
    for (size_t i = 0; i < b.views.size(); ++i) {
        saved[i] = b.views[i].spp;
        b.views[i].spp = std::max(1u, (uint32_t)std::lround(share * saved[i]));
        probe_paths += (uint64_t)b.views[i].width * b.views[i].height * b.views[i].spp;
    }
    const uint32_t path_bits = batch_path_bits(b.views.size());
    pt_batch_params p = *params;
    p.render_flags &= ~(uint32_t)(PT_RENDER_FP64_ACCUM | PT_RENDER_GUIDE_TRAIN);  // probes leave the guide as is
    uint32_t candidates = 0;
    int rc = PT_OK;
    auto time_shape = [&](const BatchShape &sh, double &ns) -> int {
        double ms = 0.0;
        int r = batch_run(ctx, p, sh, (params->batch_flags & PT_BATCH_GRAPH) != 0, &ms, nullptr);
        if (r != PT_OK) return r;
        ++candidates;
        ns = ms * 1e6 / (double)probe_paths;
        return PT_OK;
    };
    BatchShape best{kBatchDefaultPaths, 256u, 8u, ctx->trace_blocks_per_cu, params->train_paths_per_launch};
    best.capacity = std::min(best.capacity, 1u << path_bits);
    best.capacity = std::max(best.capacity, max_npix);
    double best_ns = 0.0, dummy = 0.0, default_ns = 0.0;
    // warm-up (allocations, code loading), then the default shape
    if ((rc = time_shape(best, dummy)) == PT_OK) rc = time_shape(best, best_ns);
    default_ns = best_ns;
    const uint32_t caps[] = {4u << 20, 8u << 20, 16u << 20, 32u << 20, 64u << 20};
    for (uint32_t c : caps) {
        if (rc != PT_OK) break;
        if (c < max_npix || c > (1u << path_bits) || c == best.capacity) continue;
        BatchShape sh = best;
        sh.capacity = c;
        double ns = 0.0;
        if ((rc = time_shape(sh, ns)) == PT_OK && ns < best_ns) {
            best_ns = ns;
            best = sh;
        }
    }
    const uint32_t shades[][2] = {{256, 4}, {256, 16}, {256, 64}, {128, 8}, {128, 16}, {128, 128}, {64, 16}};
    for (auto &sb : shades) {
        if (rc != PT_OK) break;
        BatchShape sh = best;
        sh.shade_block = sb[0];
        sh.shade_bpc = sb[1];
        double ns = 0.0;
        if ((rc = time_shape(sh, ns)) == PT_OK && ns < best_ns) {
            best_ns = ns;
            best = sh;
        }
    }
    for (uint32_t tb : {ctx->trace_blocks_per_cu / 2u, (ctx->trace_blocks_per_cu * 3u) / 4u}) {
        if (rc != PT_OK) break;
        if (tb == 0 || tb == best.trace_bpc) continue;
        BatchShape sh = best;
        sh.trace_bpc = tb;
        double ns = 0.0;
        if ((rc = time_shape(sh, ns)) == PT_OK && ns < best_ns) {
            best_ns = ns;
            best = sh;
        }
    }
    for (size_t i = 0; i < b.views.size(); ++i) b.views[i].spp = saved[i];
    if (rc != PT_OK) return rc;
    if ((rc = batch_reset_films(ctx)) != PT_OK) return rc;
    b.stats.paths = 0;
    b.tuned = true;
    b.tuning.paths_per_launch = best.capacity;
    b.tuning.shade_block = best.shade_block;
    b.tuning.shade_blocks_per_cu = best.shade_bpc;
    b.tuning.trace_blocks_per_cu = best.trace_bpc;
    b.tuning.candidates = candidates;
    b.tuning.probe_paths = (uint32_t)std::min<uint64_t>(probe_paths, 0xFFFFFFFFull);
    b.tuning.best_ns_per_path = best_ns;
    b.tuning.default_ns_per_path = default_ns;
    if (out) *out = b.tuning;
    return PT_OK;
}

int pt_batch_get_gate(pt_ctx *ctx, uint32_t view, pt_batch_gate_info *out) {
    if (!ctx || !out) return PT_ERR_INVALID_ARG;
    BatchState &b = ctx->batch;
    if (view >= b.views.size()) {
        ctx->error = "pt_batch_get_gate: view out of range";
        return PT_ERR_INVALID_ARG;
    }
    std::memset(out, 0, sizeof(*out));
    if (!b.d_gate_acc || view >= b.gate_views) return PT_OK;  // no gated render yet
    (void)hipSetDevice(ctx->device);
    std::vector<double> acc(kGateDoublesPerView);
    BatchViewDev dv;
    PT_HIP(hipStreamSynchronize(ctx->stream));
    PT_HIP(hipMemcpy(acc.data(), b.d_gate_acc + (size_t)view * kGateDoublesPerView,
                     kGateDoublesPerView * sizeof(double), hipMemcpyDeviceToHost));
    PT_HIP(hipMemcpy(&dv, (const BatchViewDev *)b.d_views + view, sizeof(dv), hipMemcpyDeviceToHost));
    for (uint32_t plain = 0; plain < 2; ++plain) {
        std::vector<double> x;  // per-frame errors, as k_guide_gate forms them
        for (uint32_t slot = 0; slot < kGateSlots; ++slot) {
            const size_t k = gate_slot(0u, plain, slot);
            if (acc[k + 1] > 0.0) x.push_back(acc[k] / acc[k + 1]);
        }
        double mean = 0.0, median = 0.0;
        if (!x.empty()) {
            std::sort(x.begin(), x.end());
            for (double e : x) mean += e;
            mean /= (double)x.size();
            const size_t n = x.size();
            median = (n & 1u) ? x[n / 2] : 0.5 * (x[n / 2 - 1] + x[n / 2]);
        }
        (plain ? out->plain_mean : out->guided_mean) = mean;
        (plain ? out->plain_median : out->guided_median) = median;
        (plain ? out->plain_frames : out->guided_frames) = (uint32_t)x.size();
    }
    out->guide_off = dv.guide_off;
    return PT_OK;
}

int pt_batch_get_stats(pt_ctx *ctx, pt_batch_stats *out) {
    if (!ctx || !out) return PT_ERR_INVALID_ARG;
    BatchState &b = ctx->batch;
    *out = b.stats;
    out->views = (uint32_t)b.views.size();
    if (b.d_rays) {
        (void)hipSetDevice(ctx->device);
        unsigned long long r[2] = {0, 0};
        PT_HIP(hipStreamSynchronize(ctx->stream));
        PT_HIP(hipMemcpy(r, b.d_rays, sizeof(r), hipMemcpyDeviceToHost));
        out->rays_extension = r[0];
        out->rays_shadow = r[1];
    }
    return PT_OK;
}

}  // extern "C"

