// Per-path samplers (integer arithmetic, bit-identical to the reference).
//
//  * Owen-scrambled "padded" Sobol (Part_2/src/samplers/sobol.cuh:14-117):
//    get1D hashes (px, py, dim) into a seed, shuffles the sample index with an
//    Owen scramble, takes the Van der Corput point (bit reversal), scrambles
//    it again with a derived seed, and converts with min(u * 2^-32, 1-2^-24).
//    The first two hash rounds of dimensionSeed depend only on the pixel, so
//    they are computed once per path (same values, fewer instructions).
//  * PCG-style hash RNG (samplers/pcg.cuh:11-27): state seeded from
//    (x, y, frame, sample), LCG step, RXS-M output, float = (w>>22 ^ w)/2^32.
//
// A global seed extends both: seed 0 reproduces the reference sequences.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ptd {

__device__ __forceinline__ uint32_t brev32(uint32_t x) { return __builtin_bitreverse32(x); }

// sobol.cuh:48-57
__device__ __forceinline__ uint32_t hash_mix(uint32_t x) {
    x ^= x >> 17;
    x ^= x * 0xed5ad4bbu;
    x ^= x >> 11;
    x ^= x * 0xac4c1b51u;
    x ^= x >> 15;
    x ^= x * 0x31848babu;
    x ^= x >> 14;
    return x;
}

// sobol.cuh:59-66
__device__ __forceinline__ uint32_t owen_hash(uint32_t x, uint32_t seed) {
    x += seed;
    x ^= x * 0x6c50b47cu;
    x ^= x * 0xb82f1e52u;
    x ^= x * 0xc7afe638u;
    x ^= x * 0x8d22f6e6u;
    return x;
}

__device__ __forceinline__ uint32_t owen_scramble(uint32_t v, uint32_t seed) {
    return brev32(owen_hash(brev32(v), seed));
}

constexpr float kOneMinusEps = 0.99999994f;
constexpr float kInv2Pow32 = 1.0f / 4294967296.0f;

struct Sampler {
    uint32_t type;       // 0 PCG, 1 Sobol
    uint32_t state;      // PCG state
    uint32_t index;      // Sobol sample index
    uint32_t dim;        // Sobol dimension
    uint32_t pix_hash;   // hashMix(hashMix(px + c) ^ py-mix), sobol.cuh:78-81

    __device__ __forceinline__ void init_sobol(int px, int py, uint32_t sample_index, uint32_t seed) {
        type = 1;
        index = sample_index;
        dim = 0;
        uint32_t h = static_cast<uint32_t>(px);
        h = hash_mix(h + 0x9e3779b9u);
        h ^= static_cast<uint32_t>(py) + 0x9e3779b9u + (h << 6) + (h >> 2);
        h = hash_mix(h);
        if (seed != 0u) h ^= hash_mix(seed ^ 0x85ebca6bu);
        pix_hash = h;
    }

    // pcg.cuh:18-27 (x, y from the launch; sample = index within the batch)
    __device__ __forceinline__ void init_pcg(int x, int y, uint32_t frame, uint32_t sample, uint32_t seed) {
        type = 0;
        uint32_t s = static_cast<uint32_t>(x) * 1973u + static_cast<uint32_t>(y) * 9277u + frame * 26699u +
                     sample * 12149u + seed * 0x9E3779B9u;
        s = s * 747796405u + 2891336453u;
        state = s | 1u;
    }

    // Raw 32-bit value of the next dimension (before float conversion).
    __device__ __forceinline__ uint32_t next_u32() {
        if (type == 1) {
            uint32_t h = pix_hash;
            h ^= dim + 0x9e3779b9u + (h << 6) + (h >> 2);
            h = hash_mix(h);  // dimensionSeed(px, py, dim)
            uint32_t shuffled = owen_scramble(index, h);
            uint32_t raw = brev32(shuffled);
            dim++;
            return owen_scramble(raw, hash_mix(h));
        }
        uint32_t old = state;
        state = old * 747796405u + 2891336453u;
        uint32_t word = ((old >> ((old >> 28u) + 4u)) ^ old) * 277803737u;
        return (word >> 22u) ^ word;
    }

    // Position in the sequence: the next dimension (Sobol) or the state (PCG).
    __device__ __forceinline__ uint32_t key() const { return type == 1 ? dim : state; }

    // Advances by k values without computing them.
    __device__ __forceinline__ void skip(int k) {
        if (type == 1) {
            dim += static_cast<uint32_t>(k);
        } else {
            for (int i = 0; i < k; ++i) state = state * 747796405u + 2891336453u;
        }
    }

    __device__ __forceinline__ float next() {
        uint32_t u = next_u32();
        if (type == 1) return fminf(static_cast<float>(u) * kInv2Pow32, kOneMinusEps);
        return static_cast<float>(u) / 4294967296.0f;  // pcg.cuh:15 (can round to 1.0)
    }
};

}  // namespace ptd
