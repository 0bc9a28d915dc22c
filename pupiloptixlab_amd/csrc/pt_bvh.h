// Compressed 8-wide BVH (80-byte nodes) and its closest-hit / any-hit
// traversal, shared by the host (tests, statistics) and the gfx950 kernels.
//
// Replaces the OptiX acceleration structures and optixTrace of the reference
// (world/gas_manager.cpp:69-245, world/ias_manager.cpp:29-114,
// example/path_tracer/main.cu:80-85,161-166, render/emitter.h:91-100).
// The scene is flattened to world space, so there is a single level.
//
// Result contract (independent of tree shape): the closest hit is the
// lexicographic minimum of (t, primitive id) over all triangles whose
// watertight intersection has t in [tmin, tmax] and all spheres with a root
// in [tmin, tmax] (the nearest such root); a shadow query reports whether
// any such primitive exists.  Node tests are widened by a relative margin so
// rounding in the slab test never culls a primitive the exact test would
// accept.
#pragma once

#include <cstring>

#include "pt_math.h"

namespace pupil {

struct WideNode {           // 80 bytes, read as five 16-byte loads
    float px, py, pz;       // quantization origin
    uint8_t ex, ey, ez;     // per-axis scale exponent (biased by 127)
    uint8_t imask;          // bit k set: slot k holds an internal child
    uint32_t child_base;    // index of the first internal child node
    uint32_t tri_base;      // index of the first leaf triangle
    uint8_t meta[8];        // per slot: 0 = empty; leaf: (unary count << 5) | tri offset;
                            // internal: (1 << 5) | (24 + slot)
    uint8_t qlo_x[8], qlo_y[8], qlo_z[8];
    uint8_t qhi_x[8], qhi_y[8], qhi_z[8];
};
static_assert(sizeof(WideNode) == 80, "wide node must be 80 bytes");

// A node as traversal holds it: its 80 bytes as twenty 32-bit words, decoded
// with shifts so that registers hold whole words.  Words: 0-2 px, py, pz;
// 3 ex | ey << 8 | ez << 16 | imask << 24; 4 child_base; 5 tri_base; 6-7
// meta; 8-9 qlo_x; 10-11 qlo_y; 12-13 qlo_z; 14-15 qhi_x; 16-17 qhi_y;
// 18-19 qhi_z (little-endian byte order within a word).
struct NodeWords {
    uint32_t w[20];
};
static_assert(sizeof(NodeWords) == sizeof(WideNode), "node words must cover the node");

// Leaf primitives in traversal order, three float4 each.  Triangle: v0.xyz +
// primitive id bits, v1.xyz, v2.xyz (v1.w = 0).  Sphere: v0 = (center,
// primitive id bits), v1 = (radius, 0, 0, tag bits 1), v2 unused.
struct LeafTri {
    f4 v0;
    f4 v1;
    f4 v2;
};

// Primitive ids: triangles are numbered from 0; sphere k has id
// kPrimSphere | k.
constexpr uint32_t kPrimSphere = 1u << 30;
constexpr uint32_t kPrimIndexMask = kPrimSphere - 1u;
PT_HD bool leaf_is_sphere(const LeafTri &l) { return f2u(l.v1.w) != 0u; }

constexpr float kNodeMargin = 1e-4f;     // relative widening of node intervals
// The widened test tn (1 - m) <= tf (1 + m) as one product: tn <= tf r with
// r = (1 + m) / (1 - m) rounded up (tn >= tmin >= 0).
constexpr float kNodeMarginRatio = 1.0002001f;
constexpr uint32_t kMaxBvhDepth = 64;    // builder guarantees depth <= this

struct Hit {
    float t;
    uint32_t prim;   // 0xFFFFFFFF = miss
    float b1, b2;    // barycentric weights of vertex 1 and vertex 2
};

// Per-ray constants of the watertight test (Woop, Benthin, Wald 2013).
struct RayPre {
    int kx, ky, kz;
    float sx, sy, sz;
};
PT_HD float comp(f3 v, int k) { return k == 0 ? v.x : (k == 1 ? v.y : v.z); }
PT_HD RayPre ray_precompute(f3 d) {
    RayPre r;
    float ax = fabsf(d.x), ay = fabsf(d.y), az = fabsf(d.z);
    int kz = (ax > ay) ? (ax > az ? 0 : 2) : (ay > az ? 1 : 2);
    int kx = kz + 1;
    if (kx == 3) kx = 0;
    int ky = kx + 1;
    if (ky == 3) ky = 0;
    float dkz = comp(d, kz);
    if (dkz < 0.f) {
        int tmp = kx;
        kx = ky;
        ky = tmp;
    }
    r.kx = kx;
    r.ky = ky;
    r.kz = kz;
    r.sx = comp(d, kx) / dkz;
    r.sy = comp(d, ky) / dkz;
    r.sz = 1.f / dkz;
    return r;
}

// Watertight ray/triangle test.  Returns true with t = T/det and the
// barycentric weights of v1 and v2 when the ray crosses the triangle (either
// facing); the caller applies the t range.
PT_HD bool intersect_triangle(f3 o, const RayPre &rp, f3 v0, f3 v1, f3 v2, float &t, float &b1, float &b2) {
    f3 A = v0 - o, B = v1 - o, C = v2 - o;
    float Akz = comp(A, rp.kz), Bkz = comp(B, rp.kz), Ckz = comp(C, rp.kz);
    float Ax = comp(A, rp.kx) - rp.sx * Akz;
    float Ay = comp(A, rp.ky) - rp.sy * Akz;
    float Bx = comp(B, rp.kx) - rp.sx * Bkz;
    float By = comp(B, rp.ky) - rp.sy * Bkz;
    float Cx = comp(C, rp.kx) - rp.sx * Ckz;
    float Cy = comp(C, rp.ky) - rp.sy * Ckz;
    float U = Cx * By - Cy * Bx;
    float V = Ax * Cy - Ay * Cx;
    float W = Bx * Ay - By * Ax;
    if (U == 0.f || V == 0.f || W == 0.f) {
        double CxBy = (double)Cx * (double)By, CyBx = (double)Cy * (double)Bx;
        U = (float)(CxBy - CyBx);
        double AxCy = (double)Ax * (double)Cy, AyCx = (double)Ay * (double)Cx;
        V = (float)(AxCy - AyCx);
        double BxAy = (double)Bx * (double)Ay, ByAx = (double)By * (double)Ax;
        W = (float)(BxAy - ByAx);
    }
    if ((U < 0.f || V < 0.f || W < 0.f) && (U > 0.f || V > 0.f || W > 0.f)) return false;
    float det = U + V + W;
    if (det == 0.f) return false;
    float Az = rp.sz * Akz, Bz = rp.sz * Bkz, Cz = rp.sz * Ckz;
    float T = U * Az + V * Bz + W * Cz;
    t = T / det;
    b1 = V / det;
    b2 = W / det;
    return true;
}

// Ray/sphere test after Haines et al., "Precision Improvements for
// Ray/Sphere Intersection" (Ray Tracing Gems, 2019, ch. 7): the discriminant
// from the squared distance of the center to the ray line, and the roots as
// c/q and q/a with q = -(b + sign(b) sqrt(disc)), which avoids cancellation.
// Returns the nearest root in [tmin, tmax].  With t >= tmin a ray
// leaving the surface (root near 0) finds the far side only when it points
// inward, as a surface-bounded sphere intersector does.
PT_HD bool intersect_sphere(f3 o, f3 d, f3 c, float r, float tmin, float tmax, float &t) {
    const f3 f = o - c;
    const float a = dot(d, d);
    const float b = dot(f, d);  // half the linear coefficient
    const f3 l = f - (b / a) * d;
    const float r2 = r * r;
    const float disc = a * (r2 - dot(l, l));
    if (!(disc >= 0.f)) return false;
    const float sq = sqrtf(disc);
    const float q = b >= 0.f ? -(b + sq) : -(b - sq);
    if (q == 0.f) return false;
    const float cc = dot(f, f) - r2;
    float t0 = cc / q, t1 = q / a;
    if (t0 > t1) {
        const float tmp = t0;
        t0 = t1;
        t1 = tmp;
    }
    if (t0 >= tmin && t0 <= tmax) {
        t = t0;
        return true;
    }
    if (t1 >= tmin && t1 <= tmax) {
        t = t1;
        return true;
    }
    return false;
}

PT_HD float exp2_from_biased(uint8_t e) { return u2f(((uint32_t)e) << 23); }
PT_HD float qbyte(const uint32_t *w, int k) { return (float)((w[k >> 2] >> (8 * (k & 3))) & 0xFFu); }

// Exponent n with x = m 2^n, m in [0.5, 1), for positive finite x (frexp);
// subnormal x reports -1022.
PT_HD int frexp_exponent(double x) {
    union { double d; uint64_t u; } v;
    v.d = x;
    const int e = (int)((v.u >> 52) & 0x7FFu);
    return e == 0 ? -1022 : e - 1022;
}

// Quantization of one axis of a wide node: the smallest power-of-two step
// 2^e (e >= -126) for which 8-bit codes exist whose decode p + q 2^e (in
// float) contains every child interval [clo, chi].  Returns false when no
// exponent up to 127 works.  The same rule serves the host builder and the
// GPU refit.
PT_HD bool quantize_node_axis(float p, float node_hi, const float *clo, const float *chi, int n, uint8_t &e_out,
                              uint8_t *qlo, uint8_t *qhi) {
    const double extent = (double)node_hi - (double)p;
    int e = -126;
    if (extent > 0.0) {
        const int ex = frexp_exponent(extent / 255.0);
        e = ex > -126 ? ex : -126;
    }
    for (; e <= 127; ++e) {
        const float sc = u2f((uint32_t)(e + 127) << 23);
        bool ok = true;
        for (int i = 0; i < n && ok; ++i) {
            const double lo = floor(((double)clo[i] - (double)p) / (double)sc);
            const double hi = ceil(((double)chi[i] - (double)p) / (double)sc);
            int ql = (int)fmax(0.0, fmin(255.0, lo));
            int qh = (int)fmax(0.0, fmin(255.0, hi));
            while (ql > 0 && !(p + (float)ql * sc <= clo[i])) --ql;
            if (!(p + (float)ql * sc <= clo[i])) ok = false;
            while (ok && qh < 255 && !(p + (float)qh * sc >= chi[i])) ++qh;
            if (ok && !(p + (float)qh * sc >= chi[i])) ok = false;
            if (ok) {
                qlo[i] = (uint8_t)ql;
                qhi[i] = (uint8_t)qh;
            }
        }
        if (ok) {
            e_out = (uint8_t)(e + 127);
            return true;
        }
    }
    return false;
}

struct TraverseStats {
    uint32_t nodes;
    uint32_t tris;
};

// Traversal over the compressed wide BVH.  `Stack` supplies push/pop/empty of
// 64-bit node groups (x = child_base, y = hit bits << 24 | imask).  Groups
// hold the not-yet-visited internal children of a node; internal hit bits are
// stored at position (slot ^ octant) so popping the lowest bit visits slots
// near-to-far for the ray's direction octant.
// `root` is the index of the tree's root node (several trees can share one
// node array).  With kSpheres false the sphere test is compiled out (for
// trees without sphere leaves).
//
// The traversal of one ray is trav_init followed by trav_step until it
// returns false; one step visits one node and tests its leaf primitives.
// traverse_bvh runs a ray to the end; the GPU closest-hit kernel interleaves
// rays per lane (a lane whose ray ends takes the next one).
struct TravRay {
    f3 o, d;
    RayPre rp;
    float idx, idy, idz;
    float tmin, tmax, tbest;
    uint32_t oct;
    uint32_t g_base, g_bits;  // current node group
    bool have;                // `n` holds (or is loading) the next node to visit
    NodeWords n;
};

// Next node to visit: the nearest remaining child of the current group, or of
// the group on top of the stack; the rest of the group is pushed.  The choice
// never depends on hits found later, so the next node is picked -- and, for
// closest-hit rays, its load issued -- before the current node's primitives
// are tested, and the two memory latencies overlap.
template <typename Stack>
PT_HD bool trav_next(TravRay &r, Stack &stack, uint32_t &node_index) {
    if ((r.g_bits & 0xFF000000u) == 0) {
        if (stack.empty()) return false;
        stack.pop(r.g_base, r.g_bits);
    }
    const uint32_t bitpos = __builtin_ctz(r.g_bits >> 24);
    r.g_bits &= ~(1u << (24 + bitpos));
    const uint32_t slot = bitpos ^ r.oct;
    const uint32_t imask = r.g_bits & 0xFFu;
    node_index = r.g_base + __builtin_popcount(imask & ((1u << slot) - 1u));
    if (r.g_bits & 0xFF000000u) stack.push(r.g_base, r.g_bits);
    return true;
}

template <typename Stack, typename NodeLoad>
PT_HD void trav_init(TravRay &r, NodeLoad load_node, f3 o, f3 d, float tmin, float tmax, Stack &stack, Hit &hit,
                     uint32_t root) {
    r.o = o;
    r.d = d;
    r.rp = ray_precompute(d);
    r.idx = 1.f / (fabsf(d.x) > 1e-12f ? d.x : copysignf(1e-12f, d.x));
    r.idy = 1.f / (fabsf(d.y) > 1e-12f ? d.y : copysignf(1e-12f, d.y));
    r.idz = 1.f / (fabsf(d.z) > 1e-12f ? d.z : copysignf(1e-12f, d.z));
    r.oct = (d.x < 0.f ? 1u : 0u) | (d.y < 0.f ? 2u : 0u) | (d.z < 0.f ? 4u : 0u);
    r.tmin = tmin;
    r.tmax = tmax;
    r.tbest = tmax;
    hit.t = tmax;
    hit.prim = 0xFFFFFFFFu;
    hit.b1 = 0.f;
    hit.b2 = 0.f;
    // Root group: the root sits in slot 0 of a virtual parent.
    r.g_base = root;
    r.g_bits = (1u << (24 + (0 ^ r.oct))) | 1u;
    uint32_t node_index = 0;
    r.have = trav_next(r, stack, node_index);
    if (r.have) load_node(node_index, r.n);
}

// One node of the ray's traversal and its leaf primitives.  Returns whether
// the ray has more to visit (false also once an any-hit ray has found a hit).
template <bool kAnyHit, bool kSpheres = true, typename Stack, typename NodeLoad, typename TriLoad>
PT_HD bool trav_step(TravRay &r, NodeLoad load_node, TriLoad load_tri, Stack &stack, Hit &hit,
                     TraverseStats *stats) {
    if (!r.have) return false;
    const uint32_t *w = r.n.w;
    const uint32_t imask = w[3] >> 24;
    const bool neg_x = r.idx < 0.f, neg_y = r.idy < 0.f, neg_z = r.idz < 0.f;  // sign of each slab slope
    uint32_t t_base, t_bits;
    {
        if (stats) stats->nodes++;
        // Slab distances of the quantized child boxes: with a = 2^e / d and
        // b = (origin - o) / d per axis, plane q is at t = q a + b (one FMA).
        // Rounding differs from decoding the box first by a few ulps, which
        // the relative node margin covers; triangle tests stay exact.
        // The near plane of an axis is the low bound when the slope a is
        // positive and the high bound otherwise (fma is monotone in q), so
        // the byte arrays are swapped once per node instead of taking a
        // min and a max per child (Ylitie et al. 2017); the distances are
        // the same as min(t0, t1) / max(t0, t1).
        const float ax = u2f((w[3] & 0xFFu) << 23) * r.idx, bx = (u2f(w[0]) - r.o.x) * r.idx;
        const float ay = u2f(((w[3] >> 8) & 0xFFu) << 23) * r.idy, by = (u2f(w[1]) - r.o.y) * r.idy;
        const float az = u2f(((w[3] >> 16) & 0xFFu) << 23) * r.idz, bz = (u2f(w[2]) - r.o.z) * r.idz;
        uint32_t nqx[2], fqx[2], nqy[2], fqy[2], nqz[2], fqz[2];
        for (int i = 0; i < 2; ++i) {  // selects, not branches
            nqx[i] = neg_x ? w[14 + i] : w[8 + i];
            fqx[i] = neg_x ? w[8 + i] : w[14 + i];
            nqy[i] = neg_y ? w[16 + i] : w[10 + i];
            fqy[i] = neg_y ? w[10 + i] : w[16 + i];
            nqz[i] = neg_z ? w[18 + i] : w[12 + i];
            fqz[i] = neg_z ? w[12 + i] : w[18 + i];
        }
        // A hit internal child sets bit 24 + (slot ^ octant), a hit leaf its
        // unary triangle count at its offset, so one word carries both.
        uint32_t bits = 0;
        for (int k = 0; k < 8; ++k) {
            const uint32_t meta = (w[6 + (k >> 2)] >> (8 * (k & 3))) & 0xFFu;
            if (meta == 0) continue;
            const float txn = fmaf(qbyte(nqx, k), ax, bx), txf = fmaf(qbyte(fqx, k), ax, bx);
            const float tyn = fmaf(qbyte(nqy, k), ay, by), tyf = fmaf(qbyte(fqy, k), ay, by);
            const float tzn = fmaf(qbyte(nqz, k), az, bz), tzf = fmaf(qbyte(fqz, k), az, bz);
            float tn = fmaxf(fmaxf(txn, tyn), fmaxf(tzn, r.tmin));
            float tf = fminf(fminf(txf, tyf), fminf(tzf, r.tbest));
            if (tn <= tf * kNodeMarginRatio) {
                if ((imask >> k) & 1u) bits |= (1u << 24) << (k ^ r.oct);
                else bits |= (meta >> 5) << (meta & 31u);
            }
        }
        r.g_base = w[4];
        r.g_bits = (bits & 0xFF000000u) | imask;
        t_base = w[5];
        t_bits = bits & 0x00FFFFFFu;
    }
    // Closest-hit rays prefetch the next node; any-hit rays usually stop at
    // their first primitive hit, so for them the load waits (fewer registers).
    uint32_t node_index = 0;
    const bool have_next = trav_next(r, stack, node_index);
    if (!kAnyHit && have_next) load_node(node_index, r.n);

    while (t_bits) {
        uint32_t i = __builtin_ctz(t_bits);
        t_bits &= t_bits - 1u;
        LeafTri tri;
        load_tri(t_base + i, tri);
        if (stats) stats->tris++;
        float t, b1 = 0.f, b2 = 0.f;
        bool hit_prim;
        if (kSpheres && leaf_is_sphere(tri)) {
            hit_prim = intersect_sphere(r.o, r.d, f3{tri.v0.x, tri.v0.y, tri.v0.z}, tri.v1.x, r.tmin, r.tmax, t);
        } else {
            hit_prim = intersect_triangle(r.o, r.rp, f3{tri.v0.x, tri.v0.y, tri.v0.z},
                                          f3{tri.v1.x, tri.v1.y, tri.v1.z}, f3{tri.v2.x, tri.v2.y, tri.v2.z}, t,
                                          b1, b2);
        }
        if (hit_prim) {
            uint32_t prim = f2u(tri.v0.w);
            if (t >= r.tmin && (t < r.tbest || (t == r.tbest && prim < hit.prim))) {
                hit.t = t;
                hit.prim = prim;
                hit.b1 = b1;
                hit.b2 = b2;
                if (kAnyHit) {
                    r.have = false;
                    return false;
                }
                r.tbest = t;
            }
        }
    }
    if (kAnyHit && have_next) load_node(node_index, r.n);
    r.have = have_next;
    return have_next;
}

template <bool kAnyHit, bool kSpheres = true, typename Stack, typename NodeLoad, typename TriLoad>
PT_HD bool traverse_bvh(NodeLoad load_node, TriLoad load_tri, f3 o, f3 d, float tmin, float tmax,
                        Stack &stack, Hit &hit, TraverseStats *stats, uint32_t root = 0u) {
    TravRay r;
    trav_init(r, load_node, o, d, tmin, tmax, stack, hit, root);
    while (trav_step<kAnyHit, kSpheres>(r, load_node, load_tri, stack, hit, stats)) {
    }
    return hit.prim != 0xFFFFFFFFu;
}

}  // namespace pupil
