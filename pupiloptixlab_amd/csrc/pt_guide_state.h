// Device-resident state of the guiding network and its online optimizer.
//
// Everything the optimizer decides (whether a step has samples, whether its
// gradients are finite, the Adam bias corrections, the dynamic loss scale)
// lives in a device control block, so rendering, sample collection and
// training steps are queued on one stream without host round trips.
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "../../include/pupil_hip.h"
#include "pt_guide_dev.h"

namespace pupil {

constexpr float kGuideGradClip = 100.f;   // max |dLoss/doutput| per sample, cross-entropy part
constexpr float kGuideL2Clip = 1.f;       // ... relative-L2 part
constexpr float kGuideScaleInit = 32.f;   // initial loss scale of the FP16 backward pass
constexpr float kGuideScaleMax = 32768.f;
constexpr uint32_t kGuideScaleGrowth = 1000;  // finite steps before the scale doubles
constexpr uint32_t kGuideDwSlice = 512;   // samples per weight-gradient partial

struct GuideCtl {
    float loss_scale;
    uint32_t step;         // applied Adam steps
    uint32_t good;         // finite steps since the last scale change
    uint32_t bad;          // non-finite gradient seen in the current step
    uint32_t skipped;      // steps dropped for non-finite gradients
    uint32_t active;       // the current step has samples
    uint32_t batch_valid;  // samples in the current step
    uint32_t n_samples;    // samples collected since the last collection reset
    uint32_t n_nonzero;    // completed records of the last launch with a contribution
    uint32_t n_zero;       // ... and without
    unsigned long long total_samples;  // samples kept over all collections
    float c1, c2;          // Adam bias corrections of the current step
    float loss_sum, loss_cnt;
    float last_loss;
    uint32_t last_count;
    float fp_sum;          // weighted sum of F/p over collected samples
    float fp_cnt;          // weighted count of collected samples with F/p > 0
    float fp_norm;         // selection-loss normalization of the current step
    uint32_t ring_head;    // next slot of the sample ring (collections into a ring)
    uint32_t ring_start;   // first slot of the current ring collection
};

// Training records written by the guided shading kernel, one per guided
// vertex of a training path (component-major, 5 float4 per record):
//   0: position.xyz, BSDF pdf of the sampled direction
//   1: sampled direction.xyz, |cos| to the shading normal
//   2: path throughput after the vertex (T_{k+1}).xyz, hash grid of the vertex
//   3: sample weight f |cos| / p.xyz
//   4: radiance sum L_after - T_{k+1} Le of the next vertex.xyz, epoch bits
// Record k of a path is completed by the shading of vertex k + 1 (at once,
// with zero throughput, when the sampled direction carries nothing).
struct GuideRecords {
    float4 *rec;
    uint32_t capacity;   // records
    uint32_t stride;     // one camera path in `stride` is a training path
    uint32_t phase;      // training paths: path % stride == phase
    uint32_t max_depth;  // records per training path
    uint32_t epoch;      // launch stamp
};

class GuideState {
public:
    bool enabled = false;
    pt_guide_config cfg{};
    GuideLevel levels[kGuideLevels]{};
    float bounds_lo[3]{}, bounds_inv[3]{};  // grid 0
    uint32_t grids = 1;                     // hash grids, one per scene; the MLP is shared
    std::vector<float> grid_bounds;         // per grid: lo.xyz, 0, inv_ext.xyz, 0
    float *d_grid_bounds = nullptr;
    size_t grid_entries = 0;                // table entries per grid (levels x T)
    size_t table_entries = 0;               // all grids
    bool for_batch = false;                 // one grid per batch view (pt_guide_enable_batch)

    // parameters and optimizer moments
    uint32_t *table16 = nullptr;  // FP16 feature pairs used by the kernels
    float2 *table32 = nullptr, *table_m = nullptr, *table_v = nullptr, *table_g = nullptr;
    uint32_t *grid_touched = nullptr;  // per grid: a sample of the current step names it
    float *w32 = nullptr, *w_m = nullptr, *w_v = nullptr, *w_grad = nullptr;
    _Float16 *fwd_a = nullptr, *bwd_a = nullptr;
    GuideCtl *ctl = nullptr;

    // per-step scratch
    uint16_t *xt = nullptr, *dyt = nullptr;
    uint32_t batch_cap = 0;

    // training samples (3 float4 each) and render-side records
    float4 *samples = nullptr;
    uint32_t sample_cap = 0;   // allocated
    uint32_t sample_fill = 0;  // used by the next collection and steps (set by ensure_samples)
    float4 *rec = nullptr;
    uint32_t rec_cap = 0;
    uint32_t epoch = 0;
    uint32_t phase = 0;

    uint64_t steps_launched = 0;
    hipStream_t stream = nullptr;

    ~GuideState() { release(); }
    void release();
    // lo / hi: 3 floats per grid (the scene boxes the grids cover).
    bool init(const pt_guide_config &c, const float *lo, const float *hi, uint32_t n_grids, hipStream_t s,
              std::string &err);
    GuideView view() const;
    // Allocates room for n samples and makes n the collection size: a
    // collection keeps a uniform subsample of at most n samples of its launch.
    bool ensure_samples(uint32_t n, std::string &err);
    bool ensure_records(uint32_t n, std::string &err);
    // Sample collection from the finished paths of one launch: packed at the
    // start of the buffer, or (ring_n > 0) at most ring_cap samples appended
    // to a ring of ring_n slots, which then holds the last collections.
    bool collect(const float4 *out_l, uint32_t total_paths, const GuideRecords &r, std::string &err,
                 uint32_t ring_n = 0, uint32_t ring_cap = 0);
    // `steps` optimizer steps over the collected samples, batch_size each;
    // steps without samples are skipped on the device.
    bool train(uint32_t steps, float fp_norm, std::string &err);
    // ... when the samples only name grids [grid_lo, grid_hi): Adam visits those.
    bool train(uint32_t steps, float fp_norm, uint32_t grid_lo, uint32_t grid_hi, std::string &err);
    // Steps over the sample ring (ring_n = m B slots): step k takes every m-th
    // slot from offset (first_step + k) mod m, so each batch spans the whole
    // ring; Adam visits the grids the batch's samples name.  Ring samples are normalized per
    // collection, so the selection-loss normalization is 1 (fp_norm unused).
    bool train_ring(uint32_t steps, float fp_norm, uint32_t first_step, uint32_t ring_n, std::string &err);
    // Host-sample entry points (tests / external trainers).
    bool upload_samples(const float *host, uint32_t n, std::string &err);
    bool gradients(uint32_t n, float fp_norm, float *table_grad, float *weight_grad, float *loss, std::string &err);
    bool get_params(float *table, float *weights, std::string &err);
    bool set_params(const float *table, const float *weights, std::string &err);
    bool query(const float *pos, uint32_t n, float *out, uint32_t grid, std::string &err);
    bool info(pt_guide_info &out, std::string &err);
    bool grid_box(uint32_t grid, float lo[3], float inv[3], std::string &err);
    bool read_samples(float *out, uint32_t max_samples, uint32_t &count, std::string &err);

private:
    bool ensure_batch(uint32_t batch, std::string &err);
    bool launch_step(uint32_t first, uint32_t batch, float fp_norm, bool apply, uint32_t grid_lo, uint32_t grid_hi,
                     std::string &err, uint32_t ring_n = 0);
    bool pack(std::string &err);
};

}  // namespace pupil
