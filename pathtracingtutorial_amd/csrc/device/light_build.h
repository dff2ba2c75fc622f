// Light-selection tables built on the GPU.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>

#include "pt_amd.h"

namespace ptd {

// Alias table + pmf of n non-negative weights in device memory: pmf[i] =
// w[i] / sum(w); alias[i] = {prob, index bits} in the layout sample_alias
// reads (u2 < prob keeps bucket i, else alias).  The table gives each i the
// probability pmf[i], like the reference's Vose construction
// (lights/light_sampler.cpp:7-55); its layout differs (Vose is sequential;
// this one matches light buckets to heavy buckets along prefix sums of their
// deficits and surpluses, in parallel).  Synchronous on `stream`.
hipError_t build_alias_device(const float* weights, uint32_t n, float* pmf, float2* alias, hipStream_t stream,
                              std::string* err);

// The same build with its working buffers in a device array of this library
// (one per GPU, PT_ALIAS_NOALLOC_MAX_WEIGHTS weights at most; more gives
// hipErrorInvalidValue): no device memory is allocated or freed.  Builds on
// one GPU share the array, so builds that overlap on one GPU race.
hipError_t build_alias_device_noalloc(const float* weights, uint32_t n, float* pmf, float2* alias, hipStream_t stream,
                                      std::string* err);

// The environment's tables (light_sampler.cpp:112-149 weights: luminance *
// sin(theta) of the texel centre) from its device texels.
hipError_t build_env_tables_device(const float4* env, int w, int h, float* pmf, float2* alias, hipStream_t stream,
                                   std::string* err);

// The emissive-triangle tables (light_sampler.cpp:57-110) from the device
// geometry: faces {v0, v1, v2, material} in BVH order over vpos, and the
// materials.  Weight = area * max(luminance(emissive * strength), 0) for
// faces whose material emits; the emissive list holds {face, pmf bits, cdf
// bits, 0} in face order like the host build, tri_pmf (nf floats, zeroed
// here) each face's selection pmf.  *emissive and *alias (*count entries)
// are allocated here and owned by the caller; both stay null, and *count 0,
// when no face emits.  Synchronous on `stream`.
hipError_t build_emissive_tables_device(const float4* vpos, const uint4* faces, uint32_t nf, const pt_material* mats,
                                        uint32_t nmat, float* tri_pmf, uint32_t* count, uint4** emissive,
                                        float2** alias, hipStream_t stream, std::string* err);

}  // namespace ptd
