// How a launch's paths split into concurrent chains of kernels (pt_render and
// batched launches), shared by the host code that plans the chains, the
// raygen kernels that fill each chain's queue part, and the host tests
// (tests/chain_sim.cpp).
#pragma once

#include <cstdint>

#ifdef __HIPCC__
#define PT_CHAIN_HD __host__ __device__
#else
#define PT_CHAIN_HD
#endif

namespace pupil {

// At most this many chains per launch, of at least kChainPaths paths each;
// batched queues hold a multiple of kChainLcm slots, so every chain count
// divides them.
constexpr uint32_t kMaxChains = 4;
constexpr uint32_t kChainPaths = 1000000u;
constexpr uint32_t kChainLcm = 12;

// Chains of a launch of `paths` paths (one when `one` is set).
inline uint32_t chain_count(uint64_t paths, bool one) {
    if (one) return 1u;
    const uint64_t n = paths / kChainPaths;
    return n < 1 ? 1u : (n > kMaxChains ? kMaxChains : (uint32_t)n);
}

// One chain of a pt_render launch: frames [frame0, frame0 + frames) of the
// pixels at places [pix0, pix0 + pixels) of the film's tile order, in queue
// slots [slot0, slot0 + paths).
struct ChainPart {
    uint32_t frame0, frames, pix0, pixels, slot0, paths;
};

// With at least as many frames as chains, a chain takes a share of the
// frames of every pixel (the chains then run through the same pixels at the
// same time and share L2); otherwise a band of pixels in every frame.
inline void chain_parts(uint32_t frames, uint32_t npix, uint32_t nchain, ChainPart *part) {
    const bool by_frames = frames >= nchain;
    uint32_t slot = 0;
    for (uint32_t c = 0; c < nchain; ++c) {
        ChainPart &k = part[c];
        if (by_frames) {
            k.frame0 = frames * c / nchain;
            k.frames = frames * (c + 1) / nchain - k.frame0;
            k.pix0 = 0;
            k.pixels = npix;
        } else {
            k.frame0 = 0;
            k.frames = frames;
            k.pix0 = (uint32_t)((uint64_t)npix * c / nchain);
            k.pixels = (uint32_t)((uint64_t)npix * (c + 1) / nchain) - k.pix0;
        }
        k.slot0 = slot;
        k.paths = k.frames * k.pixels;
        slot += k.paths;
    }
}

// Frame (of the launch) and tile-order pixel place of a chain's queue slot i:
// the chain's frames lie side by side, so a wave carries a few pixels in all
// of them.
PT_CHAIN_HD inline void chain_slot(uint32_t frame0, uint32_t frames, uint32_t pix0, uint32_t i, uint32_t &frame,
                                   uint32_t &place) {
    frame = frame0 + i % frames;
    place = pix0 + i / frames;
}

// Batched launches: chain c of n takes the launch's slots [begin(c), begin(c + 1))
// (path counts come from the device launch descriptor) into a fixed share of
// the queues, batch_share(queue_paths, n) slots.
PT_CHAIN_HD inline uint32_t batch_chain_begin(uint32_t total, uint32_t c, uint32_t n) {
    return (uint32_t)((uint64_t)total * c / n);
}
inline uint32_t batch_queue_slots(uint32_t capacity) { return (capacity + kChainLcm - 1) / kChainLcm * kChainLcm; }
inline uint32_t batch_share(uint32_t queue_slots, uint32_t n) { return queue_slots / n; }

}  // namespace pupil
