// Launch-side declarations shared by the kernels and the C-ABI host code.
#pragma once

#include <hip/hip_runtime.h>

#include "transport.h"

namespace g4go {

// Smallest dynamic LDS budget per block for the staged scene (BVH nodes,
// daughter boxes, property pool, scene records): with the histogram (1 KB) and
// the waves' hit stages (3.25 KB per wave), four 256-thread blocks (16 waves,
// the register-limited occupancy) still fit in 160 KB.  Larger blocks get
// more (lds_scene_budget).
constexpr uint32_t kLdsSceneBytes = 26112;
constexpr int kHistBins = 256;
constexpr int kHistSlots = kHistBins + 2;  // + underflow (256) + overflow (257)
// The arrival-time histogram: kHistBins bins of 200 ns / kHistBins from 0 ns.
constexpr double kHistLoNs = 0.0;
constexpr double kHistBinsPerNs = kHistBins / 200.0;

// Work is cut into items of `chunk` consecutive photons of one source record
// (one event of a synthetic batch, or one emission record).  Items are not
// stored: item c belongs to source s with item_offsets[s] <= c <
// item_offsets[s + 1] (exclusive prefix of ceil(count / chunk)), found by a
// wave-uniform binary search when the item is claimed.  This is the
// reference's emissionOffsets layout (OptiXTransportHost.cpp:643-666) at item
// rather than photon granularity.
struct ItemRef {
    uint32_t event_index;   // index into the batch's event list
    uint32_t source_index;  // emission index (emission source) or event index (synthetic)
    uint32_t first;         // first local photon index within the source record
    uint32_t count;
};

// Per-event counters (src/optical/optix/OptiXDeviceData.cuh:159-167).
struct EventStats {
    unsigned long long detected, absorbed, escaped, truncated, max_bounce, invalid, zero_step;
};

// The transport kernels accumulate per-event counters into kEventReplicas
// copies of the per-event array (block b writes copy b % kEventReplicas), so
// that the waves of a batch dominated by a few large events do not all
// serialise on the atomics of one record; reduce_event_stats sums the copies
// (the maximum for max_bounce) afterwards.  Copy r of event e is
// event_stats[r * event_stride + e].
constexpr uint32_t kEventReplicas = 32;

enum Counter : int {
    kCounterSteps = 0,      // transport steps executed (the reference's totalBounceCount)
    kCounterPhotons = 1,    // photons generated
    kCounterDropped = 2,    // detections beyond the hit capacity
    kCounterQueueHead = 3,  // work-queue cursor (low 32 bits used)
    kCounterCount = 4,
};

enum SourceKind : int { kSourceSynthetic = 0, kSourceEmission = 1 };

struct KernelParams {
    SceneView scene;
    Rng rng;
    StepConfig cfg;
    uint32_t max_steps;
    uint32_t refill_threshold;  // persistent form: refill when at least this many lanes idle
    SyntheticSource synth;
    const Emission* emissions;
    const uint32_t* header;          // device-resident batch header: {n_sources, chunk, n_events, 0}
    const uint32_t* event_ids;       // per event index
    const uint32_t* item_offsets;    // n_sources + 1 (item count = item_offsets[n_sources])
    const uint32_t* source_counts;   // photons per source record
    const uint32_t* source_event;    // event index per source record (emission source)
    const uint32_t* source_photon_offset;  // n_sources + 1: first launch index of each source's photons
    uint32_t item_begin, item_end;   // item range of this launch; item_end == 0: the whole batch
    Hit* hits;
    uint32_t hit_capacity;
    uint32_t* hit_count;
    uint32_t* hit_launch;  // launch index of each hit's photon (source_photon_offset[source] + local index)
    unsigned long long* histogram;
    EventStats* event_stats;   // kEventReplicas copies, event_stride records apart
    uint32_t event_stride;
    uint32_t pad_es;
    unsigned long long* counters;
    // Scene parts staged into LDS by every block, in this order (each only
    // when it fits; otherwise read in global memory): BVH nodes + daughter
    // boxes (lds_entries 32-byte entries), the property pool (lds_pool_floats
    // floats, padded to 16 bytes), then the volume, surface, mesh plane and
    // triangle, and material records selected by lds_records (kLdsVolumes | kLdsSurfaces |
    // kLdsPlanes | kLdsMaterials).  transport_lds_bytes gives the dynamic LDS
    // size.
    uint32_t lds_entries;
    uint32_t lds_pool_floats;
    uint32_t lds_records;
    uint32_t lds_stage_quads;  // 16-byte word of dynamic LDS where the hit stages begin (set at launch)
    // Persistent form: ready-photon ring entries per wave, in dynamic LDS
    // after the hit stages (set at launch from the LDS left over, at most
    // kRingMaxSlots; 0: idle lanes only get photons from generation passes)
    uint32_t ring_slots;
    // Persistent form: the start states of the batch's photons, generated by
    // generate_kernel before the transport launch, three float4 per photon in
    // launch order (StartState); null: idle lanes generate their photons at
    // refill.
    const float4* starts;
};

// Start state of one photon as generate_kernel writes it: position and
// energy (NaN when start_photon rejects the photon), direction and the low
// word of the FP64 time, polarisation and the high word.
constexpr uint32_t kStartQuads = 3;

// Parameterised property tables for the adjoint: theta offset of each
// material's absorption / Rayleigh table and each surface's reflectivity /
// efficiency table (kInvalid = not a parameter).  A table of n values
// occupies theta[offset .. offset + n).
struct ParamMap {
    const uint32_t* material_abs;
    const uint32_t* material_rayleigh;
    const uint32_t* surface_reflectivity;
    const uint32_t* surface_efficiency;
    uint32_t n_params;
};

struct AdjointParams {
    KernelParams fwd;          // the forward batch being replayed
    const Hit* hits;           // its detections (device order)
    const uint32_t* hit_launch;  // launch index of each detection's photon
    uint32_t n_hits;
    const double* slot_weight;  // dLoss/dN + dLoss/dh_b per histogram slot (kHistSlots)
    double* grad;               // n_params, accumulated
    ParamMap map;
    // [0] replayed, [1] path mismatches, [2] replays scored step by step on a
    // second pass, [3] replay steps, [4] hit-queue head (zeroed before launch)
    unsigned long long* counters;
};

constexpr uint32_t kLdsVolumes = 1u, kLdsSurfaces = 2u, kLdsMaterials = 4u, kLdsPlanes = 8u;

// Threads per block of the transport kernels: 64 to kMaxBlock, whole waves.
// Each wave of a block has its own hit stage in dynamic LDS, ahead of the
// staged scene: 64 hit records of 48 bytes and their 64 launch indices.
constexpr uint32_t kMaxBlock = 512;
constexpr uint32_t kStageBytesPerWave = 64u * 48u + 64u * 4u;

// Ready-photon ring of the persistent form: 16-byte words per entry, entries
// per wave at most.  A launch gives the ring what its blocks' LDS share leaves
// after the static LDS (at most kLdsStaticReserve), the staged scene and the hit
// stages, with kResidentWaves waves resident per CU (the register-limited
// occupancy of the persistent kernels).
constexpr uint32_t kRingQuads = 6;
constexpr uint32_t kRingMaxSlots = 64;
constexpr uint32_t kResidentWaves = 16;
constexpr uint32_t kLdsStaticReserve = 4096;

// Ready-photon ring entries per wave that a block's scene budget leaves free.
constexpr uint32_t kRingMinSlots = 8;

// Bytes of dynamic LDS the staged scene may take in a block of `block`
// threads on a device with lds_per_cu bytes of LDS per CU: the block's share
// at kResidentWaves resident waves, less the static LDS, the hit stages and
// kRingMinSlots ring entries per wave; at least kLdsSceneBytes.  512-thread
// blocks on a 160 KB CU get 44 KB, enough for a ring of 48 mesh crystals
// (volumes, face planes, daughter boxes and BVH).
inline size_t lds_scene_budget(uint32_t block, size_t lds_per_cu) {
    const size_t waves = block < 64u ? 1u : block / 64u;
    const size_t resident = waves >= kResidentWaves ? 1u : kResidentWaves / waves;
    const size_t share = lds_per_cu / resident;
    const size_t taken = kLdsStaticReserve + waves * (kStageBytesPerWave + kRingMinSlots * kRingQuads * 16u);
    return share > taken + kLdsSceneBytes ? share - taken : kLdsSceneBytes;
}

// Dynamic LDS bytes of a transport launch (the staged scene).
__host__ __device__ inline size_t transport_lds_bytes(const KernelParams& p) {
    size_t bytes = static_cast<size_t>(p.lds_entries) * 32u + (static_cast<size_t>(p.lds_pool_floats) + 3u) / 4u * 16u;
    if (p.lds_records & kLdsVolumes) bytes += static_cast<size_t>(p.scene.n_volumes) * sizeof(Volume);
    if (p.lds_records & kLdsSurfaces) bytes += static_cast<size_t>(p.scene.n_surfaces) * sizeof(Surface);
    if (p.lds_records & kLdsPlanes) bytes += mesh_record_bytes(p.scene);
    if (p.lds_records & kLdsMaterials) bytes += static_cast<size_t>(p.scene.n_materials) * sizeof(Material);
    return bytes;
}
int launch_transport(const KernelParams& p, int source, bool persistent, int block, int grid, hipStream_t stream);
// Writes the start states of all photons of the batch (items of p, launch
// order) to starts[kStartQuads * n_photons]: the generation pre-pass of the
// persistent form.
int launch_generate(const KernelParams& p, int source, float4* starts, int grid, hipStream_t stream);
int launch_zero(void* ptr, size_t bytes, hipStream_t stream);
// Sums the kEventReplicas copies of the first header[2] events into out.
int launch_reduce_event_stats(const EventStats* replicas, uint32_t stride, const uint32_t* header, EventStats* out,
                              hipStream_t stream);
int launch_adjoint(const AdjointParams& p, int source, int block, int grid, hipStream_t stream);

// Launch-order output: reorders the hits of a batch into the order of their
// photons' launch indices (the order of a stable compaction over the launch,
// OptiXHitCompaction.cu:8-30).  Scratch: mask and word_prefix hold
// ceil(n_photons / 32) words, block_prefix ceil(words / kOrderWordsPerBlock).
constexpr uint32_t kOrderWordsPerBlock = 1024;
struct OrderParams {
    const Hit* hits;              // unordered, hit_count[0] of them (clamped to capacity)
    const uint32_t* launch;       // their launch indices
    const uint32_t* hit_count;
    uint32_t capacity;
    uint32_t n_photons;
    uint32_t* mask;               // one bit per launch index
    uint32_t* word_prefix;        // exclusive popcount prefix inside each block of words
    uint32_t* block_prefix;       // exclusive prefix over blocks of words
    Hit* out_hits;                // launch order
    uint32_t* out_launch;
};
int launch_order_hits(const OrderParams& p, hipStream_t stream);
// Work-item and photon offsets of a synthetic batch computed on the device
// (the batch's photon counts are in device memory): one block scans the
// n <= kPrefixMaxSources counts into item_offsets[n + 1] (exclusive prefix of
// ceil(count / chunk)) and photon_offsets[n + 1], and writes the batch header
// {n, chunk, n, 0}.
constexpr uint32_t kPrefixMaxSources = 1024;
int launch_synthetic_prefix(const uint32_t* counts, uint32_t n, uint32_t chunk, uint32_t* header,
                            uint32_t* item_offsets, uint32_t* photon_offsets, hipStream_t stream);
// out[0..2] = counters[kCounterSteps], counters[kCounterPhotons], counters[kCounterDropped].
int launch_export_counters(const unsigned long long* counters, unsigned long long* out, hipStream_t stream);
int launch_trace(const KernelParams& p, uint32_t event_id, uint32_t photon_id, uint32_t max_steps, float* out,
                 uint32_t* n_out, hipStream_t stream);
int launch_eval_math(const float* in, float* out, int n, int which, hipStream_t stream);
int launch_eval_fresnel(const float* in, float* out, int n, hipStream_t stream);

}  // namespace g4go
