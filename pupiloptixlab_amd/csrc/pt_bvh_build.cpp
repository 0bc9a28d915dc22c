#include "pt_bvh_build.h"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstring>
#include <functional>
#include <thread>

namespace pupil {
namespace {

struct Aabb {
    float lo[3] = {INFINITY, INFINITY, INFINITY};
    float hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    void grow(const float *p) {
        for (int a = 0; a < 3; ++a) {
            lo[a] = std::min(lo[a], p[a]);
            hi[a] = std::max(hi[a], p[a]);
        }
    }
    void grow(const Aabb &b) {
        for (int a = 0; a < 3; ++a) {
            lo[a] = std::min(lo[a], b.lo[a]);
            hi[a] = std::max(hi[a], b.hi[a]);
        }
    }
    bool valid() const { return lo[0] <= hi[0]; }
    double area() const {
        if (!valid()) return 0.0;
        double dx = (double)hi[0] - lo[0], dy = (double)hi[1] - lo[1], dz = (double)hi[2] - lo[2];
        return 2.0 * (dx * dy + dy * dz + dz * dx);
    }
};

struct PrimRef {
    Aabb box;
    float c[3];
    uint32_t prim;
};

struct BNode {
    Aabb box;
    uint32_t left = 0xFFFFFFFFu, right = 0xFFFFFFFFu;
    uint32_t first = 0, count = 0;
    bool leaf() const { return left == 0xFFFFFFFFu; }
};

constexpr int kMaxBins = 64;

class BinaryBuilder {
public:
    BinaryBuilder(std::vector<PrimRef> &refs, int bins, int threads)
        : refs_(refs), bins_(bins), threads_(threads) {
        nodes_.resize(refs.size() * 2 + 1);
    }

    uint32_t build() {
        uint32_t root = alloc();
        active_threads_ = 1;
        build_node(root, 0, (uint32_t)refs_.size());
        nodes_.resize(counter_.load());
        return root;
    }
    std::vector<BNode> &nodes() { return nodes_; }

private:
    uint32_t alloc() { return counter_.fetch_add(1); }

    void build_node(uint32_t ni, uint32_t first, uint32_t count) {
        BNode &n = nodes_[ni];
        Aabb box, cbox;
        for (uint32_t i = first; i < first + count; ++i) {
            box.grow(refs_[i].box);
            cbox.grow(refs_[i].c);
        }
        n.box = box;
        n.first = first;
        n.count = count;
        if (count == 1) return;

        uint32_t mid = first + count / 2;
        bool found = false;
        double best_cost = INFINITY;
        int best_axis = -1, best_split = -1;
        Aabb bin_box[kMaxBins];
        uint32_t bin_cnt[kMaxBins];
        double right_area[kMaxBins];
        uint32_t right_cnt[kMaxBins];
        for (int axis = 0; axis < 3; ++axis) {
            float cmin = cbox.lo[axis], cmax = cbox.hi[axis];
            if (!(cmax > cmin)) continue;
            double scale = (double)bins_ / ((double)cmax - (double)cmin);
            for (int b = 0; b < bins_; ++b) {
                bin_box[b] = Aabb();
                bin_cnt[b] = 0;
            }
            for (uint32_t i = first; i < first + count; ++i) {
                int b = (int)(((double)refs_[i].c[axis] - cmin) * scale);
                b = std::min(std::max(b, 0), bins_ - 1);
                bin_box[b].grow(refs_[i].box);
                bin_cnt[b]++;
            }
            Aabb acc;
            uint32_t cnt = 0;
            for (int b = bins_ - 1; b > 0; --b) {
                acc.grow(bin_box[b]);
                cnt += bin_cnt[b];
                right_area[b] = acc.area();
                right_cnt[b] = cnt;
            }
            acc = Aabb();
            cnt = 0;
            for (int b = 0; b < bins_ - 1; ++b) {
                acc.grow(bin_box[b]);
                cnt += bin_cnt[b];
                if (cnt == 0 || right_cnt[b + 1] == 0) continue;
                double cost = acc.area() * cnt + right_area[b + 1] * right_cnt[b + 1];
                if (cost < best_cost) {
                    best_cost = cost;
                    best_axis = axis;
                    best_split = b + 1;
                    found = true;
                }
            }
        }
        if (found) {
            float cmin = cbox.lo[best_axis], cmax = cbox.hi[best_axis];
            double scale = (double)bins_ / ((double)cmax - (double)cmin);
            auto it = std::partition(refs_.begin() + first, refs_.begin() + first + count, [&](const PrimRef &r) {
                int b = (int)(((double)r.c[best_axis] - cmin) * scale);
                b = std::min(std::max(b, 0), bins_ - 1);
                return b < best_split;
            });
            mid = (uint32_t)(it - refs_.begin());
            if (mid == first || mid == first + count) mid = first + count / 2;
        }
        uint32_t l = alloc(), r = alloc();
        n.left = l;
        n.right = r;
        uint32_t lc = mid - first, rc = first + count - mid;
        if (count > 65536 && active_threads_.load() < threads_) {
            active_threads_++;
            std::thread th([this, l, first, lc] { build_node(l, first, lc); });
            build_node(r, mid, rc);
            th.join();
            active_threads_--;
        } else {
            build_node(l, first, lc);
            build_node(r, mid, rc);
        }
    }

    std::vector<PrimRef> &refs_;
    int bins_;
    int threads_;
    std::vector<BNode> nodes_;
    std::atomic<uint32_t> counter_{0};
    std::atomic<int> active_threads_{0};
};

// Builds the quantized 8-wide tree from the binary one.
class WideCollapser {
public:
    WideCollapser(const std::vector<BNode> &bn, const std::vector<PrimRef> &refs, const float *pos4,
                  const uint32_t *idx4, const float *spheres4, std::vector<WideNode> &nodes,
                  std::vector<LeafTri> &tris)
        : bn_(bn), refs_(refs), pos4_(pos4), idx4_(idx4), sph4_(spheres4), nodes_(nodes), tris_(tris) {}

    bool run(uint32_t root, std::string &err) {
        nodes_.clear();
        tris_.clear();
        nodes_.push_back(WideNode{});
        return emit(root, 0, 1, err);
    }
    uint32_t max_depth = 0;
    double sah = 0.0;

private:
    bool emit(uint32_t bi, uint32_t wi, uint32_t depth, std::string &err) {
        max_depth = std::max(max_depth, depth);
        if (depth > kMaxBvhDepth) {
            err = "wide BVH deeper than the traversal stack supports";
            return false;
        }
        // Candidate children: open the largest-area internal binary node until
        // eight candidates exist.
        uint32_t cand[8];
        int nc = 0;
        if (bn_[bi].leaf()) {
            cand[nc++] = bi;
        } else {
            cand[nc++] = bn_[bi].left;
            cand[nc++] = bn_[bi].right;
        }
        while (nc < 8) {
            int best = -1;
            double best_area = -1.0;
            for (int i = 0; i < nc; ++i) {
                const BNode &c = bn_[cand[i]];
                if (c.leaf()) continue;
                double a = c.box.area();
                if (a > best_area) {
                    best_area = a;
                    best = i;
                }
            }
            if (best < 0) break;
            uint32_t b = cand[best];
            cand[best] = bn_[b].left;
            cand[nc++] = bn_[b].right;
        }

        // Octant-ordered slot assignment: slot s is visited first by rays whose
        // direction-sign octant is s, so it gets the child lying furthest
        // "upstream" along that octant's diagonal.
        Aabb nbox;
        for (int i = 0; i < nc; ++i) nbox.grow(bn_[cand[i]].box);
        double pc[3];
        for (int a = 0; a < 3; ++a) pc[a] = 0.5 * ((double)nbox.lo[a] + nbox.hi[a]);
        int slot_of[8];
        bool used_slot[8] = {false, false, false, false, false, false, false, false};
        bool assigned[8] = {false, false, false, false, false, false, false, false};
        double cost[8][8];
        for (int i = 0; i < nc; ++i) {
            const Aabb &b = bn_[cand[i]].box;
            double c[3];
            for (int a = 0; a < 3; ++a) c[a] = 0.5 * ((double)b.lo[a] + b.hi[a]) - pc[a];
            for (int s = 0; s < 8; ++s) {
                double dx = (s & 1) ? -1.0 : 1.0, dy = (s & 2) ? -1.0 : 1.0, dz = (s & 4) ? -1.0 : 1.0;
                cost[i][s] = c[0] * dx + c[1] * dy + c[2] * dz;
            }
        }
        for (int it = 0; it < nc; ++it) {
            int bi2 = -1, bs = -1;
            double bc = INFINITY;
            for (int i = 0; i < nc; ++i) {
                if (assigned[i]) continue;
                for (int s = 0; s < 8; ++s) {
                    if (used_slot[s]) continue;
                    if (cost[i][s] < bc) {
                        bc = cost[i][s];
                        bi2 = i;
                        bs = s;
                    }
                }
            }
            assigned[bi2] = true;
            used_slot[bs] = true;
            slot_of[bi2] = bs;
        }
        int child_in_slot[8];
        for (int s = 0; s < 8; ++s) child_in_slot[s] = -1;
        for (int i = 0; i < nc; ++i) child_in_slot[slot_of[i]] = i;

        WideNode n;
        std::memset(&n, 0, sizeof(n));
        n.px = nbox.lo[0];
        n.py = nbox.lo[1];
        n.pz = nbox.lo[2];
        n.tri_base = (uint32_t)tris_.size();
        // Internal children occupy consecutive node indices in slot order.
        uint32_t num_internal = 0;
        for (int s = 0; s < 8; ++s) {
            int i = child_in_slot[s];
            if (i >= 0 && bn_[cand[i]].count > 3) num_internal++;
        }
        n.child_base = (uint32_t)nodes_.size();
        nodes_.resize(nodes_.size() + num_internal);

        float clo[3][8], chi[3][8];
        int nslots = 0;
        uint8_t slot_list[8];
        uint32_t tri_off = 0;
        uint32_t internal_children[8];
        int ninternal = 0;
        for (int s = 0; s < 8; ++s) {
            int i = child_in_slot[s];
            if (i < 0) continue;
            const BNode &c = bn_[cand[i]];
            for (int a = 0; a < 3; ++a) {
                clo[a][nslots] = c.box.lo[a];
                chi[a][nslots] = c.box.hi[a];
            }
            slot_list[nslots++] = (uint8_t)s;
            if (c.count > 3) {
                n.imask |= (uint8_t)(1u << s);
                n.meta[s] = (uint8_t)((1u << 5) | (24u + s));
                internal_children[ninternal++] = cand[i];
            } else {
                static const uint8_t unary[4] = {0, 1, 3, 7};
                n.meta[s] = (uint8_t)((unary[c.count] << 5) | tri_off);
                for (uint32_t k = 0; k < c.count; ++k) {
                    uint32_t prim = refs_[c.first + k].prim;
                    if (prim & kPrimSphere) {
                        const float *q = sph4_ + 4 * (size_t)(prim & kPrimIndexMask);
                        LeafTri ls;
                        ls.v0 = f4{q[0], q[1], q[2], u2f(prim)};
                        ls.v1 = f4{q[3], 0.f, 0.f, u2f(1u)};
                        ls.v2 = f4{0.f, 0.f, 0.f, 0.f};
                        tris_.push_back(ls);
                        continue;
                    }
                    const uint32_t *ix = idx4_ + 4 * (size_t)prim;
                    const float *p0 = pos4_ + 4 * (size_t)ix[0];
                    const float *p1 = pos4_ + 4 * (size_t)ix[1];
                    const float *p2 = pos4_ + 4 * (size_t)ix[2];
                    LeafTri lt;
                    lt.v0 = f4{p0[0], p0[1], p0[2], u2f(prim)};
                    lt.v1 = f4{p1[0], p1[1], p1[2], 0.f};
                    lt.v2 = f4{p2[0], p2[1], p2[2], 0.f};
                    tris_.push_back(lt);
                }
                tri_off += c.count;
            }
        }
        const float nhi[3] = {nbox.hi[0], nbox.hi[1], nbox.hi[2]};
        const float np[3] = {n.px, n.py, n.pz};
        uint8_t qlo[3][8], qhi[3][8], e[3];
        for (int a = 0; a < 3; ++a) {
            if (!quantize_node_axis(np[a], nhi[a], clo[a], chi[a], nslots, e[a], qlo[a], qhi[a])) {
                err = "failed to quantize a BVH node";
                return false;
            }
        }
        n.ex = e[0];
        n.ey = e[1];
        n.ez = e[2];
        for (int j = 0; j < nslots; ++j) {
            int s = slot_list[j];
            n.qlo_x[s] = qlo[0][j];
            n.qlo_y[s] = qlo[1][j];
            n.qlo_z[s] = qlo[2][j];
            n.qhi_x[s] = qhi[0][j];
            n.qhi_y[s] = qhi[1][j];
            n.qhi_z[s] = qhi[2][j];
        }
        nodes_[wi] = n;
        sah += nbox.area();
        for (int j = 0; j < ninternal; ++j) {
            if (!emit(internal_children[j], n.child_base + (uint32_t)j, depth + 1, err)) return false;
        }
        return true;
    }

    const std::vector<BNode> &bn_;
    const std::vector<PrimRef> &refs_;
    const float *pos4_;
    const uint32_t *idx4_;
    const float *sph4_;
    std::vector<WideNode> &nodes_;
    std::vector<LeafTri> &tris_;
};

}  // namespace

bool build_wide_bvh(const float *positions4, uint32_t num_vertices, const uint32_t *indices4, uint32_t num_tris,
                    const float *spheres4, uint32_t num_spheres, const BvhBuildOptions &opt,
                    std::vector<WideNode> &nodes, std::vector<LeafTri> &tris, BvhBuildStats &stats,
                    std::string &err) {
    auto t0 = std::chrono::steady_clock::now();
    nodes.clear();
    tris.clear();
    stats = BvhBuildStats{};
    if (num_tris >= kPrimSphere || num_spheres >= kPrimSphere) {
        err = "too many primitives";
        return false;
    }
    if (num_tris + (uint64_t)num_spheres == 0) {
        WideNode empty;
        std::memset(&empty, 0, sizeof(empty));
        empty.ex = empty.ey = empty.ez = 127;
        nodes.push_back(empty);
        return true;
    }
    std::vector<PrimRef> refs((size_t)num_tris + num_spheres);
    for (uint32_t k = 0; k < num_spheres; ++k) {
        const float *q = spheres4 + 4 * (size_t)k;
        if (!std::isfinite(q[0]) || !std::isfinite(q[1]) || !std::isfinite(q[2]) || !std::isfinite(q[3]) ||
            !(q[3] > 0.f)) {
            err = "invalid sphere";
            return false;
        }
        PrimRef r;
        for (int a = 0; a < 3; ++a) {
            r.box.lo[a] = q[a] - q[3];
            r.box.hi[a] = q[a] + q[3];
            r.c[a] = q[a];
        }
        r.prim = kPrimSphere | k;
        refs[(size_t)num_tris + k] = r;
    }
    for (uint32_t t = 0; t < num_tris; ++t) {
        PrimRef r;
        for (int k = 0; k < 3; ++k) {
            uint32_t vi = indices4[4 * (size_t)t + k];
            if (vi >= num_vertices) {
                err = "triangle index out of range";
                return false;
            }
            const float *p = positions4 + 4 * (size_t)vi;
            if (!std::isfinite(p[0]) || !std::isfinite(p[1]) || !std::isfinite(p[2])) {
                err = "non-finite vertex position";
                return false;
            }
            r.box.grow(p);
        }
        for (int a = 0; a < 3; ++a) r.c[a] = 0.5f * (r.box.lo[a] + r.box.hi[a]);
        r.prim = t;
        refs[t] = r;
    }
    int threads = opt.threads > 0 ? opt.threads : (int)std::max(1u, std::thread::hardware_concurrency());
    BinaryBuilder bb(refs, std::min(kMaxBins, std::max(4, opt.num_bins)), threads);
    uint32_t root = bb.build();
    stats.binary_nodes = (uint32_t)bb.nodes().size();
    WideCollapser wc(bb.nodes(), refs, positions4, indices4, spheres4, nodes, tris);
    if (!wc.run(root, err)) return false;
    stats.wide_nodes = (uint32_t)nodes.size();
    stats.leaf_tris = (uint32_t)tris.size();
    stats.depth = wc.max_depth;
    stats.sah_cost = wc.sah;
    stats.build_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    return true;
}

}  // namespace pupil
