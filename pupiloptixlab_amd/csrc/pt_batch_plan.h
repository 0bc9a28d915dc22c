// Launch layout and frame planning of batched renders (host and device).
//
// A batched render packs the frames of many views into launches of at most a
// queue's capacity of paths; each launch is described by one LaunchDesc that
// is copied to the device before the launch's kernels run.  The planner is
// plain C++ over any container of views with width, height, spp, sample_cnt
// and random_seed, so its rules are testable on the host.
#pragma once

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <utility>
#include <vector>

namespace pupil {

// Launch layout of a batched render (one LaunchDesc per launch, copied to
// the device before the launch's kernels run).
constexpr uint32_t kMaxLaunchEntries = 1024;  // frames per launch
constexpr uint32_t kMaxLaunchSegs = 256;      // runs of consecutive frames of one view per launch
constexpr uint32_t kEntryTrain = 0x80000000u; // LaunchEntry::view flag: the frame trains the guide
constexpr uint32_t kEntryPlain = 0x40000000u; // LaunchEntry::view flag: the frame samples the BSDF only
constexpr uint32_t kEntryEval = 0x20000000u;  // LaunchEntry::view flag: the frame measures its view's error
constexpr uint32_t kEntryDecide = 0x10000000u; // LaunchEntry::view flag: the view's last evaluation frame
constexpr uint32_t kEntryViewMask = 0x0FFFFFFFu;

// Frame `frame_seed` (RNG frame index) of a view occupies the launch's paths
// [path_begin, path_begin + pixels).
struct LaunchEntry {
    uint32_t view, frame_seed, path_begin, pixels;
};
// Consecutive frames [first_entry, first_entry + frames) of one view, folded
// into its film in frame order; its pixels are threads [pix_begin, +pixels).
struct LaunchSeg {
    uint32_t view, first_entry, frames, pix_begin;
    uint32_t sample_cnt0, accumulate;
    // Gated guided renders: the view restarts its accumulation with its first
    // frame after training if it keeps the guide (BatchViewDev::guide_off 0),
    // counting frames from restart_cnt0 instead of sample_cnt0.
    uint32_t restart, restart_cnt0;
};
struct LaunchDesc {
    uint32_t total_paths, n_entries, n_segs, seg_pixels;
    uint32_t epoch, phase, train_frames, pad;  // guide training records of this launch
    LaunchEntry e[kMaxLaunchEntries];
    LaunchSeg g[kMaxLaunchSegs];
};

// Which frames of each view a guided batch render still has to plan: frames
// [next, spp) remain; frames below train_end train the guide.  T is the
// view's configured training length: the gate's evaluation window and the
// restart after training are defined by it, also after a gate decision has
// cut train_end short.
struct PlanState {
    std::vector<uint32_t> next, train_end, T;
};

// Per-view gate (guided batch renders with training).  A view that trains on
// T >= 4 kGateMinFrames frames measures its error on the frames
// [T/4, T/2) (kEntryEval), every other one of them without the guide
// (kEntryPlain); the decision is taken right after the launch holding the
// last of them (kEntryDecide), so a view whose guide does not pay stops
// training and sampling the guide at half its training frames.  A view that
// keeps the guide trains on to T and, with frames after T, restarts its
// accumulation there.
constexpr uint32_t kGateMinFrames = 4;
constexpr uint32_t kGateWindow = 128;  // at most this many evaluation frames (one slot each, mod 128)

inline bool gate_eligible(uint32_t T) { return T / 2 - T / 4 >= kGateMinFrames; }
inline uint32_t gate_eval_begin(uint32_t T) { return std::max(T / 4, T / 2 - std::min(T / 2, kGateWindow)); }

// Training frames of a view of `spp` frames in a guided batch render: the
// first train_spp (0 = all).
inline uint32_t training_frames(uint32_t spp, uint32_t train_spp) {
    return train_spp == 0 ? spp : std::min(train_spp, spp);
}

// Packs the remaining frames of every view (views in order, frames in order)
// into launches of at most `capacity` paths.  Training frames come first,
// round-robin over the views (the next k frames of every view still
// training, k chosen so that one round fills about `train_capacity` paths),
// so that the shared network learns from all views at once, then the other
// frames view by view.  A launch holds either training frames only, at most
// `train_capacity` paths of them (and at least one frame), or none.
template <typename Views>
void batch_plan(const Views &views, uint32_t capacity, bool accumulate, std::vector<LaunchDesc> &plan,
                const PlanState &ps, uint32_t train_capacity = 0, bool gate = false) {
    const size_t nv = views.size();
    // (view, frame) in launch order
    std::vector<std::pair<uint32_t, uint32_t>> order;
    std::vector<uint32_t> cur_f(ps.next);
    for (;;) {
        uint64_t active = 0;
        for (size_t vi = 0; vi < nv; ++vi)
            if (cur_f[vi] < ps.train_end[vi]) active += (uint64_t)views[vi].width * views[vi].height;
        if (active == 0) break;
        const uint32_t k = (uint32_t)std::max<uint64_t>(1, train_capacity / active);
        for (uint32_t vi = 0; vi < (uint32_t)nv; ++vi)
            for (uint32_t j = 0; j < k && cur_f[vi] < ps.train_end[vi]; ++j) order.emplace_back(vi, cur_f[vi]++);
    }
    for (uint32_t vi = 0; vi < (uint32_t)nv; ++vi)
        for (uint32_t f = cur_f[vi]; f < views[vi].spp; ++f) order.emplace_back(vi, f);
    plan.clear();
    LaunchDesc cur;
    std::memset(&cur, 0, sizeof(cur));
    // A view has at most one segment per launch: its frames are folded into
    // its film by one thread per pixel, in order.
    std::vector<uint8_t> in_launch(nv, 0);
    auto flush = [&]() {
        if (cur.n_entries) plan.push_back(cur);
        for (uint32_t k = 0; k < cur.n_segs; ++k) in_launch[cur.g[k].view] = 0;
        std::memset(&cur, 0, sizeof(cur));
    };
    for (const auto &vf : order) {
        const uint32_t vi = vf.first, f = vf.second;
        const auto &v = views[vi];
        const uint32_t npix = v.width * v.height, T = ps.T[vi];
        const bool new_seg = cur.n_segs == 0 || cur.g[cur.n_segs - 1].view != vi;
        const bool train_frame = f < ps.train_end[vi];
        const bool train_launch = cur.train_frames > 0;
        const uint32_t cap = train_frame ? std::min(capacity, train_capacity) : capacity;
        if ((uint64_t)cur.total_paths + npix > cap || cur.n_entries == kMaxLaunchEntries ||
            (new_seg && (cur.n_segs == kMaxLaunchSegs || in_launch[vi])) ||
            (cur.n_entries && train_frame != train_launch))
            flush();
        const bool seg_start = cur.n_segs == 0 || cur.g[cur.n_segs - 1].view != vi;
        LaunchEntry &e = cur.e[cur.n_entries];
        e.view = vi | (train_frame ? kEntryTrain : 0u);
        if (gate && train_frame && gate_eligible(T) && f >= gate_eval_begin(T) && f < T / 2) {
            e.view |= kEntryEval;
            if ((f - gate_eval_begin(T)) & 1u) e.view |= kEntryPlain;
            if (f == T / 2 - 1) e.view |= kEntryDecide;
        }
        if (train_frame) cur.train_frames += 1;
        e.frame_seed = v.random_seed + f;
        e.path_begin = cur.total_paths;
        e.pixels = npix;
        if (seg_start) {
            in_launch[vi] = 1;
            LaunchSeg &g = cur.g[cur.n_segs++];
            g.view = vi;
            g.first_entry = cur.n_entries;
            g.frames = 0;
            g.pix_begin = cur.seg_pixels;
            g.sample_cnt0 = accumulate ? v.sample_cnt + f : 0u;
            g.accumulate = accumulate ? 1u : 0u;
            g.restart = (gate && accumulate && !train_frame && T >= kGateMinFrames && f >= T) ? 1u : 0u;
            g.restart_cnt0 = g.restart ? f - T : 0u;
            cur.seg_pixels += npix;
        }
        cur.g[cur.n_segs - 1].frames += 1;
        cur.n_entries += 1;
        cur.total_paths += npix;
    }
    flush();
}

// Records the frames a launch rendered in the plan state (a view's frames are
// planned in increasing order).
template <typename Views>
void plan_advance(PlanState &ps, const LaunchDesc &l, const Views &views) {
    for (uint32_t en = 0; en < l.n_entries; ++en) {
        const uint32_t v = l.e[en].view & kEntryViewMask;
        ps.next[v] = l.e[en].frame_seed - views[v].random_seed + 1u;
    }
}

// The gate switched `view` off: it trains on no further frames.  True when
// this changes the remaining plan.
inline bool plan_gate_off(PlanState &ps, uint32_t view) {
    if (ps.train_end[view] <= ps.next[view]) return false;
    ps.train_end[view] = ps.next[view];
    return true;
}

}  // namespace pupil
