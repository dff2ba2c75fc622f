// Path-replay adjoint of the integrator (host interface).
#pragma once

#include <hip/hip_runtime.h>

#include <vector>

#include "dev_scene.h"

namespace ptd {

// Gradient buffers (device, accumulated with +=, any may be null).
struct AdjointGrads {
    float* base;   // num_materials x 3: the material's baseColor factor
    float* rough;  // num_materials
    float* metal;  // num_materials
    float* tex;    // texel pool x 4 (base-colour textures)
    float* env;    // env_w x env_h x 4
};

// One view of a scene: camera and batch parameters (params.frame_count,
// params.spf), img_grad = d loss / d (batch mean image), W*H float4, and an
// optional replay image (W*H float4) that receives the batch mean
// re-computed by the replay pass.
struct AdjointView {
    RenderParams params;
    const float4* img_grad;
    float4* replay;
};

// Replays every view's batch (the forward's random numbers) in one launch
// and accumulates d loss / d parameter into g.  Synchronous.
hipError_t launch_adjoint(const DevScene& s, const std::vector<AdjointView>& views, const AdjointGrads& g,
                          DevCounters* ctr, bool count, hipStream_t stream);

}  // namespace ptd
