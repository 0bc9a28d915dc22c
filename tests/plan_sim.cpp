// Host simulation of a batched render's launch sequence, for the planner
// tests: plans the views' frames, walks the launches in order, and after a
// launch holding a view's last evaluation frame applies the gate decision
// given for that view (as the renderer does after reading the decisions
// back) and plans the remaining frames again.
//
// stdin:  capacity train_capacity gate n_views
//         then per view: width height spp sample_cnt random_seed T off
// stdout: per launch "L total_paths train_frames n_entries n_segs", then its
//         entries "E view flags frame path_begin pixels" and segments
//         "S view first_entry frames pix_begin sample_cnt0 restart restart_cnt0".
#include <cstdio>
#include <vector>

#include "pt_batch_plan.h"

using namespace pupil;

struct SimView {
    uint32_t width, height, spp, sample_cnt, random_seed;
};

int main() {
    uint32_t capacity, train_capacity, gate, nv;
    if (std::scanf("%u %u %u %u", &capacity, &train_capacity, &gate, &nv) != 4) return 2;
    std::vector<SimView> views(nv);
    PlanState ps;
    ps.next.assign(nv, 0u);
    ps.T.assign(nv, 0u);
    std::vector<uint32_t> off(nv, 0u);
    for (uint32_t v = 0; v < nv; ++v) {
        SimView &w = views[v];
        if (std::scanf("%u %u %u %u %u %u %u", &w.width, &w.height, &w.spp, &w.sample_cnt, &w.random_seed,
                       &ps.T[v], &off[v]) != 7)
            return 2;
    }
    ps.train_end = ps.T;
    std::vector<LaunchDesc> plan;
    batch_plan(views, capacity, true, plan, ps, train_capacity, gate != 0);
    std::vector<uint8_t> decided(nv, 0);
    for (size_t k = 0; k < plan.size(); ++k) {
        const LaunchDesc &l = plan[k];
        std::printf("L %u %u %u %u\n", l.total_paths, l.train_frames, l.n_entries, l.n_segs);
        bool decide = false;
        for (uint32_t en = 0; en < l.n_entries; ++en) {
            const LaunchEntry &e = l.e[en];
            const uint32_t v = e.view & kEntryViewMask;
            std::printf("E %u %u %u %u %u\n", v, e.view & ~kEntryViewMask, e.frame_seed - views[v].random_seed,
                        e.path_begin, e.pixels);
            if (e.view & kEntryDecide) {
                decided[v] = 1;
                decide = true;
            }
        }
        for (uint32_t g = 0; g < l.n_segs; ++g) {
            const LaunchSeg &s = l.g[g];
            std::printf("S %u %u %u %u %u %u %u\n", s.view, s.first_entry, s.frames, s.pix_begin, s.sample_cnt0,
                        s.restart, s.restart_cnt0);
        }
        plan_advance(ps, l, views);
        if (decide && k + 1 < plan.size()) {
            bool changed = false;
            for (uint32_t v = 0; v < nv; ++v)
                if (decided[v] && off[v]) changed |= plan_gate_off(ps, v);
            if (changed) {
                std::vector<LaunchDesc> next;
                batch_plan(views, capacity, true, next, ps, train_capacity, gate != 0);
                plan.swap(next);
                k = (size_t)-1;
            }
        }
    }
    return 0;
}
