// Host interface of the denoising kernels (pt_denoise.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace pupil {

constexpr float kDenoiseAlbedoEps = 1e-3f;  // albedo floor for demodulation

struct DenoiseParams {
    uint32_t iterations;    // a-trous levels (strides 1, 2, 4, ...)
    float sigma_luminance;  // luminance edge stopping, in units of the local standard deviation
    float sigma_normal;     // exponent of the normal weight
    float sigma_albedo;     // albedo distance scale
};

// Filters `in` (w x h float4, xyz = color) into `out` (float4, w = 1).
// albedo / normal: 3 floats per pixel or null.  scratch: 2 w h float4.
// All pointers are device memory; work is enqueued on `stream`.
hipError_t denoise_launch(hipStream_t stream, uint32_t w, uint32_t h, const float4 *in, const float *albedo,
                          const float *normal, float4 *out, float4 *scratch, const DenoiseParams &prm);

}  // namespace pupil
