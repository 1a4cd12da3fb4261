// Region profile of the transport kernels: a counter build only, compiled
// with -DG4GO_REGION_PROFILE (g4go_amd/build.py build_profile); the product
// build defines the markers as nothing.
//
// For each region the profile keeps, summed over every wave of a launch:
//   [0] wave cycles spent between the region's begin and end markers
//       (s_memtime, the shader clock), including cycles in which the wave
//       waited for the SIMD while other waves issued;
//   [1] wave entries (executions of the region by a wave);
//   [2] active lanes at the end marker, summed over the entries.
// With several waves sharing each SIMD and the kernel VALU-issue bound, a
// region's share of the wave cycles approximates its share of the issued
// VALU instructions, and [2] / (64 [1]) is the lane utilisation there.
// Markers sit at single-entry points; an end marker inside divergent code
// runs under the lanes that reach it, so lanes that leave the region early
// (a rare return) are not counted in [2].
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace g4go {

enum Region : int {
    kRegionRefill = 0,     // refill block of the persistent loop (queue pops, item lookup, generation)
    kRegionGenerate,       // generation of the photons of idle lanes (inside the refill block)
    kRegionStep,           // one transport step of the alive lanes
    kRegionCompact,        // hit staging, histogram, termination counters
    kRegionGeometry,       // nearest boundary (box exit, daughters) before the step's Philox block
    kRegionRayleigh,       // Rayleigh scattering
    kRegionSurface,        // boundary: face snap, neighbour, surface physics
    kRegionAbsorbDetect,   // surface absorption and detection (efficiency)
    kRegionCount
};

#if defined(G4GO_REGION_PROFILE)
// one row per wave of the largest transport block (512 threads)
constexpr int kRegionWaves = 8;
static __shared__ unsigned long long g4go_region_lds[kRegionWaves][kRegionCount][3];
static __device__ unsigned long long g4go_region_global[kRegionCount][3];
#endif

#if defined(G4GO_REGION_PROFILE) && defined(__HIP_DEVICE_COMPILE__)

__device__ __forceinline__ unsigned long long region_clock() { return __builtin_readcyclecounter(); }

__device__ __forceinline__ void region_end(int r, unsigned long long t0) {
    const unsigned long long t1 = region_clock();
    const uint64_t exec = __builtin_amdgcn_read_exec();
    if (__lane_id() == static_cast<unsigned>(__builtin_ctzll(exec))) {
        unsigned long long* s = g4go_region_lds[threadIdx.x >> 6][r];
        s[0] += t1 - t0;
        s[1] += 1ull;
        s[2] += static_cast<unsigned long long>(__builtin_popcountll(exec));
    }
}

__device__ __forceinline__ void region_zero() {
    unsigned long long* s = &g4go_region_lds[0][0][0];
    for (uint32_t i = threadIdx.x; i < kRegionWaves * kRegionCount * 3u; i += blockDim.x) {
        s[i] = 0ull;
    }
}

__device__ __forceinline__ void region_flush() {
    for (uint32_t i = threadIdx.x; i < kRegionCount * 3u; i += blockDim.x) {
        unsigned long long v = 0;
        for (int w = 0; w < kRegionWaves; ++w) {
            v += (&g4go_region_lds[w][0][0])[i];
        }
        if (v != 0ull) {
            atomicAdd(&g4go_region_global[0][0] + i, v);
        }
    }
}

#define G4GO_REGION_BEGIN(r) const unsigned long long g4go_region_t0_##r = ::g4go::region_clock()
#define G4GO_REGION_END(r) ::g4go::region_end((r), g4go_region_t0_##r)
#define G4GO_REGION_ZERO() ::g4go::region_zero()
#define G4GO_REGION_FLUSH() ::g4go::region_flush()
#else
#define G4GO_REGION_BEGIN(r) ((void)0)
#define G4GO_REGION_END(r) ((void)0)
#define G4GO_REGION_ZERO() ((void)0)
#define G4GO_REGION_FLUSH() ((void)0)
#endif

}  // namespace g4go
