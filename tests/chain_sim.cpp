// Host check of how launches split into concurrent chains
// (pupiloptixlab_amd/csrc/pt_chains.h), for the chain tests: every (frame,
// pixel place) of a pt_render launch is traced by exactly one chain slot, the
// chains' queue parts are disjoint and cover the launch, and every batched
// chain's path count fits its fixed queue share.  Prints "ok" or the first
// failure.
#include <cstdio>
#include <vector>

#include "pt_chains.h"

using namespace pupil;

static bool check_render(uint32_t frames, uint32_t npix, bool one) {
    const uint64_t total = (uint64_t)frames * npix;
    const uint32_t n = chain_count(total, one);
    if (n < 1 || n > kMaxChains || (one && n != 1)) return false;
    if (!one && total >= (uint64_t)kMaxChains * kChainPaths && n != kMaxChains) return false;
    ChainPart part[kMaxChains];
    chain_parts(frames, npix, n, part);
    std::vector<uint8_t> seen(total, 0);
    uint64_t slot = 0;
    for (uint32_t c = 0; c < n; ++c) {
        const ChainPart &k = part[c];
        if (k.slot0 != slot || k.paths != k.frames * k.pixels) return false;
        if (k.frame0 + k.frames > frames || k.pix0 + k.pixels > npix) return false;
        for (uint32_t i = 0; i < k.paths; ++i) {
            uint32_t f, place;
            chain_slot(k.frame0, k.frames, k.pix0, i, f, place);
            if (f < k.frame0 || f >= k.frame0 + k.frames || place < k.pix0 || place >= k.pix0 + k.pixels) return false;
            uint8_t &s = seen[(uint64_t)f * npix + place];
            if (s) return false;
            s = 1;
        }
        slot += k.paths;
    }
    if (slot != total) return false;
    for (uint8_t s : seen)
        if (!s) return false;
    return true;
}

static bool check_batch(uint32_t capacity) {
    const uint32_t n = chain_count(capacity, false);
    const uint32_t queue = batch_queue_slots(capacity), share = batch_share(queue, n);
    if (queue < capacity || (uint64_t)share * n > queue) return false;
    const uint32_t totals[] = {0u, 1u, 2u, 3u, 5u, 7u, 11u, capacity / 3, capacity / 2 + 1, capacity - 1, capacity};
    for (uint32_t total : totals) {
        if (total > capacity) continue;
        uint64_t sum = 0;
        for (uint32_t c = 0; c < n; ++c) {
            const uint32_t lo = batch_chain_begin(total, c, n), hi = batch_chain_begin(total, c + 1, n);
            if (hi < lo || hi - lo > share) return false;
            sum += hi - lo;
        }
        if (sum != total || batch_chain_begin(total, n, n) != total) return false;
    }
    return true;
}

int main() {
    const uint32_t films[] = {1u, 63u, 64u, 97u * 61u, 512u * 512u, 1000000u, 1920u * 1080u, 2073599u};
    const uint32_t frame_counts[] = {1u, 2u, 3u, 4u, 5u, 7u, 20u, 64u};
    for (uint32_t npix : films)
        for (uint32_t frames : frame_counts)
            for (int one = 0; one < 2; ++one) {
                if ((uint64_t)frames * npix > 40000000ull) continue;  // keep the host check fast
                if (!check_render(frames, npix, one != 0)) {
                    std::printf("render split failed: frames %u pixels %u one %d\n", frames, npix, one);
                    return 1;
                }
            }
    const uint32_t caps[] = {12u, 1000u, 999999u, 1000000u, 2000001u, 3999999u, 4000000u, 4194304u, 33554432u,
                             33554433u, 67108864u, 100000007u};
    for (uint32_t cap : caps)
        if (!check_batch(cap)) {
            std::printf("batch split failed: capacity %u\n", cap);
            return 1;
        }
    std::printf("ok\n");
    return 0;
}
