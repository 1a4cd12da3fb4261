// gfx950 transport kernels.
//
// One lane transports one photon at a time (wave64).  Two launch forms share
// the same step code:
//   * plain: each wave owns one work item of at most 64 photons, launched as
//     one grid over the batch (the reference's one-thread-per-photon launch,
//     OptiXTransportHost.cpp:746-750);
//   * persistent: a fixed grid whose waves pop work items from a device
//     queue and refill idle lanes in place once at least `refill_threshold`
//     of the 64 lanes are idle, so long photon histories do not hold a whole
//     wave.
// Detections are compacted in place: a ballot and a lane rank place each
// wave's hits in an LDS stage, which is written out with one global atomic per
// group of at least 32 hits (replacing the flag array and
// cub::DeviceSelect::Flagged pass of OptiXHitCompaction.cu:8-30); they are
// binned into an LDS arrival-time histogram flushed once per block.  Every
// block first copies the scene into LDS as far as it fits (stage_scene).
#include "kernels.h"

namespace g4go {

namespace {

__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }

// Number of lanes below this one whose bit is set in `mask`.
__device__ __forceinline__ uint32_t rank_below(uint64_t mask) {
    return __builtin_amdgcn_mbcnt_hi(static_cast<uint32_t>(mask >> 32),
                                     __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(mask), 0u));
}

__device__ __forceinline__ uint32_t uniform(uint32_t v) {
    return static_cast<uint32_t>(__builtin_amdgcn_readfirstlane(static_cast<int>(v)));
}

__device__ __forceinline__ uint32_t wave_sum(uint32_t v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        v += __shfl_xor(v, off);
    }
    return v;
}

__device__ __forceinline__ uint32_t wave_max(uint32_t v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        v = max(v, __shfl_xor(v, off));
    }
    return v;
}

__device__ __forceinline__ float lane_float(float v, int lane) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), lane));
}

// Minimum of v over the wave's 64 lanes (every lane active), by DPP moves:
// butterflies within each row of 16 lanes, then the row minima carried up by
// row broadcasts; lane 63 ends with the minimum.
__device__ __forceinline__ uint32_t wave_min_u32(uint32_t v) {
    const auto step = [](uint32_t x, uint32_t y) { return y < x ? y : x; };
    int x = static_cast<int>(v);
    x = static_cast<int>(step(x, __builtin_amdgcn_update_dpp(x, x, 0xB1, 0xF, 0xF, false)));   // quad_perm 1,0,3,2
    x = static_cast<int>(step(x, __builtin_amdgcn_update_dpp(x, x, 0x4E, 0xF, 0xF, false)));   // quad_perm 2,3,0,1
    x = static_cast<int>(step(x, __builtin_amdgcn_update_dpp(x, x, 0x141, 0xF, 0xF, false)));  // row_half_mirror
    x = static_cast<int>(step(x, __builtin_amdgcn_update_dpp(x, x, 0x140, 0xF, 0xF, false)));  // row_mirror
    x = static_cast<int>(step(x, __builtin_amdgcn_update_dpp(x, x, 0x142, 0xA, 0xF, false)));  // row_bcast:15
    x = static_cast<int>(step(x, __builtin_amdgcn_update_dpp(x, x, 0x143, 0xC, 0xF, false)));  // row_bcast:31
    return static_cast<uint32_t>(__builtin_amdgcn_readlane(x, 63));
}

// Lanes of a wave whose daughter search the whole wave makes, at most.
constexpr int kWaveSearchLanes = 16;

// Daughter searches of this step for the alive lanes in a volume of more than
// eight and at most kScanDaughters daughters (scenes with convex mesh solids,
// where such a volume's daughters are all tested: scans_daughters).  A photon in such a
// volume (the air around a ring of crystals) makes a long serial search on a
// few lanes while the rest of the wave waits.  When at most kWaveSearchLanes
// lanes need one, the wave serves them in turn: for each, lane k tests
// daughters k, k + 64, ... of its volume with daughter_entry and keeps its
// smallest entry (ties to the lower volume index), and a wave minimum of
// (distance bits, volume index) picks the answer, the same one as a single
// lane testing every daughter.  Every lane of the wave must call this.
// Returns the searched lanes' answers; `which` is kDaughterUnsearched for the
// others (transport_step then searches itself).
template <int kGeom>
__device__ __forceinline__ DaughterHit wave_daughter_search(const SceneView& sc, const Photon& ph, bool alive,
                                                            float eps) {
    DaughterHit mine{kInf, kDaughterUnsearched, 0, kInvalid};
    if (sc.max_daughters <= 8u) {
        return mine;
    }
    bool want = false;
    if (alive && ph.volume < sc.n_volumes) {
        const uint32_t n = sc.volumes[ph.volume].daughter_count;
        want = n > 8u && n <= kScanDaughters;
    }
    uint64_t pending = __ballot(want);
    if (pending == 0ull || __popcll(pending) > kWaveSearchLanes) {
        return mine;
    }
    const uint32_t lane = lane_id();
    const V3 own_inv = {1.0f / ph.dir.x, 1.0f / ph.dir.y, 1.0f / ph.dir.z};
    while (pending != 0ull) {
        const int l = __builtin_ctzll(pending);
        pending &= pending - 1ull;
        const V3 p = {lane_float(ph.pos.x, l), lane_float(ph.pos.y, l), lane_float(ph.pos.z, l)};
        const V3 d = {lane_float(ph.dir.x, l), lane_float(ph.dir.y, l), lane_float(ph.dir.z, l)};
        const V3 inv = {lane_float(own_inv.x, l), lane_float(own_inv.y, l), lane_float(own_inv.z, l)};
        const Volume& v = sc.volumes[static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(ph.volume), l))];
        const uint32_t begin = uniform(v.daughter_begin), count = uniform(v.daughter_count);
        float bt = kInf;
        uint32_t bj = kInvalid, bf = kInvalid;
        int ba = 0;
        for (uint32_t k = lane; k < count; k += 64u) {
            const DaughterBox b = load_dbox(sc.dbox[begin + k]);
            int a = 0;
            uint32_t f = kInvalid;
            const float t = daughter_entry<kGeom, true>(sc, b, p, d, inv, eps, a, f);
            const uint32_t j = b.volume;
            if (t < bt || (t == bt && t < kInf && j < bj)) {
                bt = t;
                bj = j;
                ba = a;
                bf = f;
            }
        }
        // a finite entry is >= eps > 0 (+ 0.0f turns a -0 into +0), so its
        // bits order like the distance; equal distances go to the lower
        // volume index
        const uint32_t tkey = bt < kInf ? __float_as_uint(bt + 0.0f) : 0xffffffffu;
        const uint32_t tmin = wave_min_u32(tkey);
        DaughterHit h{kInf, kInvalid, 0, kInvalid};
        if (tmin != 0xffffffffu) {
            uint64_t at = __ballot(tkey == tmin);
            if ((at & (at - 1ull)) != 0ull) {
                const uint32_t jmin = wave_min_u32(tkey == tmin ? bj : 0xffffffffu);
                at = __ballot(tkey == tmin && bj == jmin);
            }
            const int w = __builtin_ctzll(at);
            h.t = lane_float(bt, w);
            h.which = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(bj), w));
            h.axis = __builtin_amdgcn_readlane(ba, w);
            h.plane = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(bf), w));
        }
        if (lane == static_cast<uint32_t>(l)) {
            mine = h;
        }
    }
    return mine;
}

// The block's copy of the per-event counters (computed where a flush needs
// it rather than held through the transport loop).
__device__ __forceinline__ EventStats* event_copy(const KernelParams& p) {
    return p.event_stats + static_cast<size_t>(blockIdx.x % kEventReplicas) * p.event_stride;
}

// The launch's parameters, re-derived opaquely (an empty asm statement on the
// kernel-argument pointer, no instruction): loads through the result are
// scalar loads of the kernel-argument segment where the transport loop uses
// them, not values the compiler hoists into SGPRs for the whole loop and then
// spills into VGPR lanes (every reload of a spilled value is a VALU
// instruction; the queue, hit-list and source fields are used only at
// refills and flushes).
__device__ __forceinline__ const KernelParams& fresh_params() {
    typedef const KernelParams __attribute__((address_space(4))) ArgParams;
    ArgParams* kp = (ArgParams*)__builtin_amdgcn_kernarg_segment_ptr();
    asm volatile("" : "+s"(kp));
    return *(const KernelParams*)kp;
}

struct LaneStats {
    uint32_t detected, absorbed, escaped, truncated, invalid, zero_step, max_bounce;
};

__device__ __forceinline__ void flush_event(EventStats* es, uint32_t event, LaneStats& s) {
    if (event == kInvalid) {
        return;
    }
    EventStats& e = es[event];
    if (s.detected) atomicAdd(&e.detected, static_cast<unsigned long long>(s.detected));
    if (s.absorbed) atomicAdd(&e.absorbed, static_cast<unsigned long long>(s.absorbed));
    if (s.escaped) atomicAdd(&e.escaped, static_cast<unsigned long long>(s.escaped));
    if (s.truncated) atomicAdd(&e.truncated, static_cast<unsigned long long>(s.truncated));
    if (s.invalid) atomicAdd(&e.invalid, static_cast<unsigned long long>(s.invalid));
    if (s.zero_step) atomicAdd(&e.zero_step, static_cast<unsigned long long>(s.zero_step));
    if (s.max_bounce) atomicMax(&e.max_bounce, static_cast<unsigned long long>(s.max_bounce));
    s = LaneStats{0, 0, 0, 0, 0, 0, 0};
}

// Flushes the counters of the lanes in `needs` to their events with one set
// of atomics per distinct event among them (wave-uniform control flow).
__device__ __forceinline__ void flush_event_group(EventStats* es, uint32_t event, LaneStats& s, bool needs) {
    uint64_t pending = __ballot(needs);
    while (pending != 0ull) {
        const uint32_t leader = static_cast<uint32_t>(__ffsll(static_cast<long long>(pending)) - 1);
        const uint32_t e = uniform(static_cast<uint32_t>(__shfl(static_cast<int>(event), static_cast<int>(leader))));
        const bool member = needs && event == e;
        LaneStats r;
        r.detected = wave_sum(member ? s.detected : 0u);
        r.absorbed = wave_sum(member ? s.absorbed : 0u);
        r.escaped = wave_sum(member ? s.escaped : 0u);
        r.truncated = wave_sum(member ? s.truncated : 0u);
        r.invalid = wave_sum(member ? s.invalid : 0u);
        r.zero_step = wave_sum(member ? s.zero_step : 0u);
        r.max_bounce = wave_max(member ? s.max_bounce : 0u);
        if (lane_id() == leader) {
            flush_event(es, e, r);
        }
        if (member) {
            s = LaneStats{0, 0, 0, 0, 0, 0, 0};
        }
        pending &= ~__ballot(member);
    }
}

// Histogram slot of arrival time t: bin floor((t - kHistLoNs) *
// kHistBinsPerNs), then underflow (and NaN) and overflow slots.  A product
// rather than a division by the bin width keeps an FP64 divide out of every
// detection.
__device__ __forceinline__ int hist_slot(double t) {
    const double x = (t - kHistLoNs) * kHistBinsPerNs;
    if (!(x >= 0.0)) {
        return kHistBins;  // underflow (and NaN)
    }
    if (x >= static_cast<double>(kHistBins)) {
        return kHistBins + 1;
    }
    return static_cast<int>(x);
}

// Locates work item c (wave-uniform): binary search of the item offsets.
template <int kSource>
__device__ __forceinline__ ItemRef locate_item(const KernelParams& p, uint32_t c, uint32_t n_sources,
                                               uint32_t chunk) {
    uint32_t lo = 0, hi = n_sources;  // last s with item_offsets[s] <= c
    while (hi - lo > 1u) {
        const uint32_t mid = uniform((lo + hi) >> 1);
        if (p.item_offsets[mid] <= c) {
            lo = mid;
        } else {
            hi = mid;
        }
    }
    ItemRef r;
    r.source_index = uniform(lo);
    r.first = uniform((c - p.item_offsets[lo]) * chunk);
    r.count = uniform(min(chunk, p.source_counts[lo] - r.first));
    r.event_index = kSource == kSourceSynthetic ? r.source_index : uniform(p.source_event[lo]);
    return r;
}

template <int kSource>
__device__ __forceinline__ void generate(Photon& ph, const KernelParams& p, const SceneView& sv, uint32_t event_index,
                                         uint32_t source, uint32_t local) {
    if (kSource == kSourceSynthetic) {
        generate_synthetic(ph, p.synth, p.rng, p.event_ids[event_index], local);
    } else {
        generate_emission(ph, p.emissions[source], sv, p.rng, local);
    }
}

// Detected photons are staged per wave in LDS and moved to the global hit
// list in groups: one atomic reservation per group instead of one per wave
// step (a single global counter serialises at the L2 otherwise).
constexpr uint32_t kStageCap = 64;    // staged records per wave (a step adds at most 64)
constexpr uint32_t kStageFlush = 32;  // flush once this many are staged
static_assert(kStageBytesPerWave == 3u * kStageCap * 16u + kStageCap * 4u, "stage layout");

__device__ __forceinline__ void flush_stage(const KernelParams& p, const float4* stage, const uint32_t* src,
                                            uint32_t staged, uint32_t lane) {
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
    uint32_t base = 0;
    if (lane == 0u) {
        base = atomicAdd(p.hit_count, staged);
    }
    base = uniform(static_cast<uint32_t>(__shfl(static_cast<int>(base), 0)));
    float4* dst = reinterpret_cast<float4*>(p.hits);
#pragma unroll 1
    for (uint32_t i = lane; i < 3u * staged; i += 64u) {
        if (base + i / 3u < p.hit_capacity) {
            dst[3ull * base + i] = stage[i];
        }
    }
    if (p.hit_launch != nullptr) {
#pragma unroll 1
        for (uint32_t i = lane; i < staged; i += 64u) {
            if (base + i < p.hit_capacity) {
                p.hit_launch[base + i] = src[i];
            }
        }
    }
    if (lane == 0u && base + staged > p.hit_capacity) {
        const uint32_t kept = base >= p.hit_capacity ? 0u : p.hit_capacity - base;
        atomicAdd(p.counters + kCounterDropped, static_cast<unsigned long long>(staged - kept));
    }
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

__device__ __forceinline__ void store_hit(float4* d, const Photon& ph, uint32_t sensor) {
    d[0] = make_float4(ph.pos.x, ph.pos.y, ph.pos.z, ph.dir.x);
    d[1] = make_float4(ph.dir.y, ph.dir.z, ph.energy, __uint_as_float(sensor));
    const uint64_t tb = __double_as_longlong(ph.time);
    d[2] = make_float4(__uint_as_float(ph.event_id), __uint_as_float(ph.photon_id),
                       __uint_as_float(static_cast<uint32_t>(tb)), __uint_as_float(static_cast<uint32_t>(tb >> 32)));
}

// Start state of a generated photon (kStartQuads float4 at out), checked by
// start_photon; a rejected start is marked by a NaN energy.
__device__ __forceinline__ void write_start(float4* out, Photon& g) {
    const bool ok = start_photon(g);
    const uint64_t tb = static_cast<uint64_t>(__double_as_longlong(g.time));
    out[0] = make_float4(g.pos.x, g.pos.y, g.pos.z, ok ? g.energy : u2f(0x7fc00000u));
    out[1] = make_float4(g.dir.x, g.dir.y, g.dir.z, __uint_as_float(static_cast<uint32_t>(tb)));
    out[2] = make_float4(g.pol.x, g.pol.y, g.pol.z, __uint_as_float(static_cast<uint32_t>(tb >> 32)));
}

// Photon `local` of source `source` (event index event_index) from its start
// state at launch index `launch`: the identifiers as generate() sets them.
// Returns false when the start was rejected.
template <int kSource>
__device__ __forceinline__ bool load_start(Photon& ph, const KernelParams& p, uint32_t event_index, uint32_t source,
                                           uint32_t local, uint32_t launch) {
    const float4* r = p.starts + static_cast<size_t>(kStartQuads) * launch;
    const float4 a = r[0], b = r[1], c = r[2];
    ph.pos = V3{a.x, a.y, a.z};
    ph.energy = a.w;
    ph.dir = V3{b.x, b.y, b.z};
    ph.pol = V3{c.x, c.y, c.z};
    ph.time = __longlong_as_double(
        static_cast<long long>((static_cast<uint64_t>(__float_as_uint(c.w)) << 32) | __float_as_uint(b.w)));
    if (kSource == kSourceSynthetic) {
        ph.photon_id = local;
        ph.event_id = p.event_ids[event_index];
        ph.volume = p.synth.volume;
    } else {
        const Emission& em = p.emissions[source];
        ph.photon_id = em.first_photon_id + local;
        ph.event_id = em.event_id;
        ph.volume = em.volume;
    }
    ph.step = 0;
    ph.material = kInvalid;
    ph.grid_i = kInvalid;
    return a.w == a.w;
}

// Ring entry of a ready photon (kRingQuads 16-byte words): position and
// energy (NaN: start rejected); direction and the time's low word;
// polarisation and its high word; photon id, event id, volume, material; the
// cached bulk properties; hot-grid node and weight, launch index (a parked
// photon: its step) and event index.
__device__ __forceinline__ void ring_put(float4* s, const Photon& ph, float energy, uint32_t word, uint32_t event) {
    const uint64_t tb = static_cast<uint64_t>(__double_as_longlong(ph.time));
    s[0] = make_float4(ph.pos.x, ph.pos.y, ph.pos.z, energy);
    s[1] = make_float4(ph.dir.x, ph.dir.y, ph.dir.z, __uint_as_float(static_cast<uint32_t>(tb)));
    s[2] = make_float4(ph.pol.x, ph.pol.y, ph.pol.z, __uint_as_float(static_cast<uint32_t>(tb >> 32)));
    s[3] = make_float4(__uint_as_float(ph.photon_id), __uint_as_float(ph.event_id), __uint_as_float(ph.volume),
                       __uint_as_float(ph.material));
    s[4] = make_float4(ph.rindex, ph.inv_velocity, ph.interaction_length, ph.absorb_fraction);
    s[5] = make_float4(__uint_as_float(ph.grid_i), ph.grid_f, __uint_as_float(word), __uint_as_float(event));
}

__device__ __forceinline__ void ring_get(const float4* s, Photon& ph, uint32_t& word, uint32_t& event) {
    const float4 a = s[0], b = s[1], c = s[2], d = s[3], e = s[4], f = s[5];
    ph.pos = V3{a.x, a.y, a.z};
    ph.energy = a.w;
    ph.dir = V3{b.x, b.y, b.z};
    ph.pol = V3{c.x, c.y, c.z};
    ph.time = __longlong_as_double(
        static_cast<long long>((static_cast<uint64_t>(__float_as_uint(c.w)) << 32) | __float_as_uint(b.w)));
    ph.photon_id = __float_as_uint(d.x);
    ph.event_id = __float_as_uint(d.y);
    ph.volume = __float_as_uint(d.z);
    ph.material = __float_as_uint(d.w);
    ph.rindex = e.x;
    ph.inv_velocity = e.y;
    ph.interaction_length = e.z;
    ph.absorb_fraction = e.w;
    ph.grid_i = __float_as_uint(f.x);
    ph.grid_f = f.y;
    word = __float_as_uint(f.z);
    event = __float_as_uint(f.w);
}

// A lane's transported photon, parked in its ring slot during a generation pass.
__device__ __forceinline__ void ring_park(float4* s, const Photon& ph) { ring_put(s, ph, ph.energy, ph.step, kInvalid); }

// Swaps the freshly generated photon (launch index, event index; ok: its
// start was accepted) into the slot for the photon parked there.
__device__ __forceinline__ void ring_swap(float4* s, Photon& ph, bool ok, uint32_t launch, uint32_t event) {
    Photon parked;
    uint32_t step, unused;
    ring_get(s, parked, step, unused);
    ring_put(s, ph, ok ? ph.energy : u2f(0x7fc00000u), launch, event);
    ph = parked;
    ph.step = step;
}

// Takes a ready photon from a ring slot; false for a rejected start.
__device__ __forceinline__ bool ring_take(const float4* s, Photon& ph, uint32_t& launch, uint32_t& event) {
    ring_get(s, ph, launch, event);
    ph.step = 0;
    return ph.energy == ph.energy;
}

}  // namespace

// kLdsMode: which scene parts a launch guarantees are staged in LDS
// (lds_mode): 0 none for certain, 1 the volume and surface records (and the
// mesh planes of a scene with mesh solids), 2 those and the whole property
// pool and the material records.  Their pointers then come from the LDS array
// on every path, so the compiler reads them with LDS instructions instead of
// flat ones.
enum LdsMode : int { kLdsNone = 0, kLdsHot = 1, kLdsAll = 2 };

inline int lds_mode(const KernelParams& p) {
    const uint32_t hot = kLdsVolumes | kLdsSurfaces | (has_meshes(p.scene) ? kLdsPlanes : 0u);
    const uint32_t all = hot | kLdsMaterials;
    if ((p.lds_records & all) == all && p.scene.n_pool != 0 && p.lds_pool_floats == p.scene.n_pool) {
        return kLdsAll;
    }
    return (p.lds_records & hot) == hot ? kLdsHot : kLdsNone;
}

// Copies the scene parts that p selects (lds_entries, lds_pool_floats,
// lds_records; the kLdsMode parts unconditionally) into `lds`, in the order
// BVH nodes + daughter boxes, pool, volumes, surfaces, mesh planes,
// materials, and points sv at the copies.  The caller synchronises the block
// before using them.
template <int kLdsMode, int kGeom = kGeomBoxes>
__device__ __forceinline__ void stage_scene(const KernelParams& p, float4* lds, SceneView& sv) {
    if (p.lds_entries != 0u) {
        const uint32_t node_words = 2u * p.scene.n_bvh_nodes;
        const uint32_t words = 2u * p.lds_entries;
        const float4* nodes = reinterpret_cast<const float4*>(p.scene.bvh);
        const float4* boxes = reinterpret_cast<const float4*>(p.scene.dbox);
        for (uint32_t i = threadIdx.x; i < words; i += blockDim.x) {
            lds[i] = i < node_words ? nodes[i] : boxes[i - node_words];
        }
        sv.bvh = reinterpret_cast<const BvhNode*>(lds);
        sv.dbox = reinterpret_cast<const DaughterBox*>(lds + node_words);
    }
    float4* next_quad = lds + 2u * p.lds_entries;
    // copies `quads` 16-byte words from global memory to the next LDS slot
    const auto stage_quads = [&](const void* src, uint32_t quads) {
        float4* const dst = next_quad;
        const float4* s = reinterpret_cast<const float4*>(src);
        for (uint32_t i = threadIdx.x; i < quads; i += blockDim.x) {
            dst[i] = s[i];
        }
        next_quad += quads;
        return static_cast<const void*>(dst);
    };
    if (kLdsMode >= kLdsAll || p.lds_pool_floats != 0u) {
        // the pool region of the scene upload is padded to 256 bytes, so whole
        // float4 reads stay inside it
        sv.pool = static_cast<const float*>(stage_quads(p.scene.pool, (p.lds_pool_floats + 3u) / 4u));
    }
    if (kLdsMode >= kLdsHot || (p.lds_records & kLdsVolumes)) {
        sv.volumes = static_cast<const Volume*>(stage_quads(p.scene.volumes, p.scene.n_volumes * (sizeof(Volume) / 16u)));
    }
    if (kLdsMode >= kLdsHot || (p.lds_records & kLdsSurfaces)) {
        sv.surfaces =
            static_cast<const Surface*>(stage_quads(p.scene.surfaces, p.scene.n_surfaces * (sizeof(Surface) / 16u)));
    }
    if (kGeom != kGeomBoxes && (kLdsMode >= kLdsHot || (p.lds_records & kLdsPlanes))) {  // box scenes have none
        // planes, triangles and face sibling lists, whole 16-byte words (a
        // scene with triangle solids runs kGeomTriangles or kGeomOpen; below
        // them n_tris is 0)
        const uint32_t quads = p.scene.n_planes * (sizeof(Plane) / 16u) +
                               (kGeom >= kGeomTriangles ? p.scene.n_tris * (sizeof(Tri) / 16u) : 0u) +
                               p.scene.n_sib_quads;
        sv.planes = static_cast<const Plane*>(stage_quads(p.scene.planes, quads));
    }
    if (kLdsMode >= kLdsAll || (p.lds_records & kLdsMaterials)) {
        sv.materials = static_cast<const Material*>(
            stage_quads(p.scene.materials, p.scene.n_materials * (sizeof(Material) / 16u)));
    }
}

// kPregen (persistent form): idle lanes load the start states that
// generate_kernel wrote for their photons (KernelParams::starts) instead of
// generating them at refill.
template <int kSource, bool kPersistent, int kLdsMode, int kGeom, bool kPregen = false>
__device__ __forceinline__ void transport_body(const KernelParams& p) {
    __shared__ uint32_t s_hist[kHistSlots];
    // dynamic LDS: the staged scene (transport_lds_bytes), then from word
    // lds_stage_quads each wave's hit stage (kStageCap records of three
    // 16-byte words, then their kStageCap launch indices)
    extern __shared__ float4 s_dynamic[];
    for (int i = threadIdx.x; i < kHistSlots; i += blockDim.x) {
        s_hist[i] = 0u;
    }
    SceneView sv = p.scene;
    stage_scene<kLdsMode, kGeom>(p, s_dynamic, sv);
    G4GO_REGION_ZERO();
    __syncthreads();
    float4* const stage = s_dynamic + p.lds_stage_quads + uniform(threadIdx.x >> 6) * (kStageBytesPerWave / 16u);
    uint32_t* const stage_src = reinterpret_cast<uint32_t*>(stage + 3u * kStageCap);
    uint32_t staged = 0;  // wave-uniform

    const uint32_t lane = lane_id();
    const uint32_t n_sources = uniform(p.header[0]);
    const uint32_t chunk = uniform(p.header[1]);
    const uint32_t first_item = p.item_end != 0u ? p.item_begin : 0u;
    const uint32_t end_item = p.item_end != 0u ? p.item_end : uniform(p.item_offsets[n_sources]);


    Photon ph;
    ph.material = kInvalid;
    bool alive = false;
    uint32_t my_event = kInvalid;
    uint32_t my_launch = kInvalid;
    LaneStats st{0, 0, 0, 0, 0, 0, 0};
    unsigned long long wave_steps = 0, wave_photons = 0;

    // wave-uniform queue cursor (persistent form)
    uint32_t it_event = 0, it_source = 0, it_first = 0, it_next = 0, it_end = 0;
    bool done = false;
    if constexpr (!kPersistent) {
        // plain form: the wave's one work item, a photon per lane, generated
        // before the transport loop
        const uint32_t wave = first_item + blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
        if (wave < end_item) {
            const ItemRef item = locate_item<kSource>(p, uniform(wave), n_sources, chunk);
            wave_photons = item.count;
            if (lane < item.count) {
                const uint32_t local = item.first + lane;
                my_event = item.event_index;
                my_launch = p.source_photon_offset[item.source_index] + local;
                bool ok;
                if (kPregen) {
                    ok = load_start<kSource>(ph, p, item.event_index, item.source_index, local, my_launch);
                } else {
                    generate<kSource>(ph, p, sv, item.event_index, item.source_index, local);
                    ph.material = kInvalid;
                    ok = start_photon(ph);
                }
                if (ok) {
                    prepare_start(ph, sv);
                    alive = true;
                } else {
                    ++st.invalid;
                }
            }
        }
        done = true;
    }

    // Persistent form: the wave's ring of ready photons in dynamic LDS, after
    // the block's hit stages (ring_slots entries of kRingQuads words).  A
    // generation pass fills it through the alive lanes; lanes that go idle
    // later take its entries [0, ring_count) from the top.
    constexpr bool kInLaneRefill = kPersistent && kGeom >= kGeomTriangles;
    const uint32_t ring_slots = kPersistent && !kInLaneRefill ? uniform(p.ring_slots) : 0u;
    float4* const ring = s_dynamic + p.lds_stage_quads + (blockDim.x >> 6) * (kStageBytesPerWave / 16u) +
                         uniform(threadIdx.x >> 6) * ring_slots * kRingQuads;
    uint32_t ring_count = 0;  // wave-uniform

    const KernelParams& p_arg = p;
    for (;;) {
        // (the kernel's only argument is the KernelParams that p refers to)
        const KernelParams& p = kPersistent ? fresh_params() : p_arg;
        // ---- refill idle lanes (persistent form) ----
        uint64_t idle = __ballot(!alive);
        uint32_t n_idle = static_cast<uint32_t>(__popcll(idle));
        if (kInLaneRefill && (n_idle >= p.refill_threshold || n_idle == 64u)) {
            // triangle-solid instances: idle lanes generate their photons in
            // place, item by item (the ring's LDS traffic per photon costs more
            // than it saves at their few steps per photon)
            while (n_idle > 0u) {
                if (it_next >= it_end) {
                    if (done) {
                        break;
                    }
                    uint32_t c = 0;
                    if (lane == 0u) {
                        c = atomicAdd(reinterpret_cast<uint32_t*>(p.counters + kCounterQueueHead), 1u);
                    }
                    c = first_item + uniform(static_cast<uint32_t>(__shfl(static_cast<int>(c), 0)));
                    if (c >= end_item) {
                        done = true;
                        break;
                    }
                    const ItemRef item = locate_item<kSource>(p, c, n_sources, chunk);
                    it_event = item.event_index;
                    it_source = item.source_index;
                    it_first = item.first;
                    it_next = 0;
                    it_end = item.count;
                    continue;
                }
                const uint32_t take = min(it_end - it_next, n_idle);
                const uint32_t rank = rank_below(idle);
                const bool fill = !alive && rank < take;
                flush_event_group(event_copy(p), my_event, st, fill && my_event != it_event && my_event != kInvalid);
                if (fill) {
                    my_event = it_event;
                    my_launch = p.source_photon_offset[it_source] + it_first + it_next + rank;
                    generate<kSource>(ph, p, sv, it_event, it_source, it_first + it_next + rank);
                    ph.material = kInvalid;
                    if (start_photon(ph)) {
                        alive = true;
                    } else {
                        ++st.invalid;
                    }
                }
                it_next += take;
                wave_photons += take;
                idle = __ballot(!alive);
                n_idle = static_cast<uint32_t>(__popcll(idle));
            }
        } else if (kPersistent && !kInLaneRefill && (n_idle >= p.refill_threshold || n_idle == 64u)) {
            G4GO_REGION_BEGIN(kRegionRefill);
            // (1) idle lanes take ready photons from the top of the ring
            if (ring_count != 0u) {
                const uint32_t take = min(ring_count, n_idle);
                const uint32_t rank = rank_below(idle);
                const bool fill = !alive && rank < take;
                uint32_t event = kInvalid, launch = kInvalid;
                bool ok = false;
                if (fill) {
                    ok = ring_take(ring + (ring_count - 1u - rank) * kRingQuads, ph, launch, event);
                }
                flush_event_group(event_copy(p), my_event, st, fill && my_event != event && my_event != kInvalid);
                if (fill) {
                    my_event = event;
                    my_launch = launch;
                    if (ok) {
                        alive = true;
                    } else {
                        ++st.invalid;
                    }
                }
                ring_count -= take;
                idle = __ballot(!alive);
                n_idle = static_cast<uint32_t>(__popcll(idle));
            }
            // (2) generation pass, once the ring is empty: the photons of the
            // idle lanes and, through up to ring_slots alive lanes, the ring's
            // next photons are generated together on the wave
            if (n_idle != 0u && ring_count == 0u && !(done && it_next >= it_end)) {
                G4GO_REGION_BEGIN(kRegionGenerate);
                const uint32_t to_ring = min(64u - n_idle, ring_slots);
                const uint32_t n_want = n_idle + to_ring;
                const uint32_t rank = alive ? rank_below(~idle) : rank_below(idle);
                const bool ring_lane = alive && rank < to_ring;
                const uint32_t want = alive ? n_idle + rank : rank;  // the lane's place in the pass
                // (a) claim photons from the queue in pass order
                uint32_t g_event = kInvalid, g_source = 0, g_local = 0;
                bool got = false;
                uint32_t claimed = 0;  // wave-uniform
                uint32_t items_used = 0;  // wave-uniform: work items the pass draws from
                while (claimed < n_want) {
                    if (it_next >= it_end) {
                        if (done) {
                            break;
                        }
                        uint32_t c = 0;
                        if (lane == 0u) {
                            c = atomicAdd(reinterpret_cast<uint32_t*>(p.counters + kCounterQueueHead), 1u);
                        }
                        c = first_item + uniform(static_cast<uint32_t>(__shfl(static_cast<int>(c), 0)));
                        if (c >= end_item) {
                            done = true;
                            break;
                        }
                        const ItemRef item = locate_item<kSource>(p, c, n_sources, chunk);
                        it_event = item.event_index;
                        it_source = item.source_index;
                        it_first = item.first;
                        it_next = 0;
                        it_end = item.count;
                        continue;
                    }
                    const uint32_t take = min(it_end - it_next, n_want - claimed);
                    ++items_used;
                    if ((!alive || ring_lane) && want >= claimed && want - claimed < take) {
                        got = true;
                        g_event = it_event;
                        g_source = it_source;
                        g_local = it_first + it_next + (want - claimed);
                    }
                    it_next += take;
                    claimed += take;
                }
                wave_photons += claimed;
                // (b) a ring lane parks its own photon in its slot while it
                // generates the new one
                float4* const slot = ring + rank * kRingQuads;
                const bool swap = ring_lane && got;
                if (swap) {
                    ring_park(slot, ph);
                }
                // (c) generation and the first step's property setup
                bool ok = false;
                uint32_t g_launch = kInvalid;
                if (items_used == 1u) {
                    // one work item: its event and source are wave-uniform
                    if (got) {
                        g_launch = p.source_photon_offset[it_source] + g_local;
                        if (kPregen) {
                            ok = load_start<kSource>(ph, p, it_event, it_source, g_local, g_launch);
                        } else {
                            generate<kSource>(ph, p, sv, it_event, it_source, g_local);
                            ph.material = kInvalid;
                            ok = start_photon(ph);
                        }
                    }
                } else if (got) {
                    g_launch = p.source_photon_offset[g_source] + g_local;
                    if (kPregen) {
                        ok = load_start<kSource>(ph, p, g_event, g_source, g_local, g_launch);
                    } else {
                        generate<kSource>(ph, p, sv, g_event, g_source, g_local);
                        ph.material = kInvalid;
                        ok = start_photon(ph);
                    }
                }
                if (got && ok) {
                    prepare_start(ph, sv);
                }
                // (d) the new photon into the slot, the parked one back
                if (swap) {
                    ring_swap(slot, ph, ok, g_launch, g_event);
                }
                const bool fresh = !alive && got;
                flush_event_group(event_copy(p), my_event, st, fresh && my_event != g_event && my_event != kInvalid);
                if (fresh) {
                    my_event = g_event;
                    my_launch = g_launch;
                    if (ok) {
                        alive = true;
                    } else {
                        ++st.invalid;
                    }
                }
                ring_count = claimed > n_idle ? claimed - n_idle : 0u;
                G4GO_REGION_END(kRegionGenerate);
            }
            G4GO_REGION_END(kRegionRefill);
        }
        const uint64_t alive_lanes = __ballot(alive);
        if (alive_lanes == 0ull) {
            if (done && it_next >= it_end && ring_count == 0u) {
                break;
            }
            continue;
        }
        wave_steps += static_cast<unsigned long long>(__popcll(alive_lanes));

        // ---- one transport step ----
        int status = kAlive;
        uint32_t sensor = kInvalid;
        G4GO_REGION_BEGIN(kRegionStep);
        DaughterHit searched{kInf, kDaughterUnsearched, 0, kInvalid};
        if constexpr (kGeom == kGeomConvex) {
            searched = wave_daughter_search<kGeom>(sv, ph, alive, p.cfg.eps);
        }
        if (alive) {
            bool zero = false;
            status = transport_step<false, kGeom>(ph, sv, p.rng, p.cfg, zero, sensor, nullptr, searched);
            ++ph.step;
            st.zero_step += zero ? 1u : 0u;
            if (status == kAlive && ph.step >= p.max_steps) {
                status = -1;  // truncated
            }
        }
        G4GO_REGION_END(kRegionStep);

        // ---- fused hit compaction: ballot + prefix into the wave's LDS stage ----
        G4GO_REGION_BEGIN(kRegionCompact);
        const uint64_t det = __ballot(status == kDetected);
        if (det != 0ull) {
            const uint32_t n_det = static_cast<uint32_t>(__popcll(det));
            if (staged + n_det > kStageCap) {  // rare: more than 64 - 31 detections in one step
                flush_stage(p, stage, stage_src, staged, lane);
                staged = 0;
            }
            if (status == kDetected) {
                const uint32_t slot = staged + rank_below(det);
                store_hit(stage + 3u * slot, ph, sensor);
                stage_src[slot] = my_launch;
                atomicAdd(&s_hist[hist_slot(ph.time)], 1u);
            }
            staged += n_det;
            if (staged >= kStageFlush) {
                flush_stage(p, stage, stage_src, staged, lane);
                staged = 0;
            }
        }
        if (alive && status != kAlive) {
            alive = false;
            st.max_bounce = max(st.max_bounce, ph.step);
            if (status == kDetected) {
                ++st.detected;
            } else if (status == kAbsorbedBulk || status == kAbsorbedSurface) {
                ++st.absorbed;
            } else if (status == kEscaped) {
                ++st.escaped;
            } else if (status == kInvalidState) {
                ++st.invalid;
            } else {
                ++st.truncated;
            }
        }
        G4GO_REGION_END(kRegionCompact);
    }

    if (staged != 0u) {
        flush_stage(p, stage, stage_src, staged, lane);
    }
    flush_event_group(event_copy(p), my_event, st, my_event != kInvalid);
    if (lane == 0u) {
        if (wave_steps) atomicAdd(p.counters + kCounterSteps, static_cast<unsigned long long>(wave_steps));
        if (wave_photons) atomicAdd(p.counters + kCounterPhotons, static_cast<unsigned long long>(wave_photons));
    }
    __syncthreads();
    for (int i = threadIdx.x; i < kHistSlots; i += blockDim.x) {
        if (s_hist[i] != 0u) {
            atomicAdd(p.histogram + i, static_cast<unsigned long long>(s_hist[i]));
        }
    }
    G4GO_REGION_FLUSH();
}

// Waves per SIMD of the transport kernels.  Box and convex-mesh instances:
// four, i.e. at most 128 VGPRs.  Left alone the compiler schedules the body
// into ~140 registers (three waves); the fourth wave hides more of the
// dependent scene-load latency of a step than the few spills it forces cost.
// Triangle-solid instances: three (168 VGPRs), whose longer steps spilled at
// 128 (576 bytes of scratch per lane); the L-shaped guide runs 8% faster.
template <int kGeom>
constexpr int kTransportWavesPerEu = kGeom >= kGeomTriangles ? 3 : 4;

// Persistent form.
template <int kSource, int kLdsMode, int kGeom, bool kPregen>
__global__ void __launch_bounds__(kMaxBlock) __attribute__((amdgpu_waves_per_eu(kTransportWavesPerEu<kGeom>)))
transport_kernel_persistent(KernelParams p) {
    transport_body<kSource, true, kLdsMode, kGeom, kPregen>(p);
}

// Generation pre-pass of the persistent form: wave w generates the photons of
// work items w, w + waves, ... (64 photons at a time, one per lane) and
// writes their start states in launch order.  Generation reads only the
// source and, for emission records, the material tables in global memory.
template <int kSource>
__global__ void __launch_bounds__(256) generate_kernel(KernelParams p, float4* starts) {
    const SceneView sv = p.scene;
    const uint32_t lane = lane_id();
    const uint32_t n_sources = uniform(p.header[0]);
    const uint32_t chunk = uniform(p.header[1]);
    const uint32_t n_items = uniform(p.item_offsets[n_sources]);
    const uint32_t waves = gridDim.x * (blockDim.x >> 6);
    for (uint32_t c = uniform(blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6)); c < n_items; c += waves) {
        const ItemRef item = locate_item<kSource>(p, c, n_sources, chunk);
        const uint32_t base = uniform(p.source_photon_offset[item.source_index]);
        for (uint32_t k = lane; k < item.count; k += 64u) {
            Photon g;
            generate<kSource>(g, p, sv, item.event_index, item.source_index, item.first + k);
            write_start(starts + static_cast<size_t>(kStartQuads) * (base + item.first + k), g);
        }
    }
}

// The plain form gets the persistent form's register cap, so that the two
// forms differ only in the work queue and refill.
template <int kSource, int kLdsMode, int kGeom>
__global__ void __launch_bounds__(kMaxBlock) __attribute__((amdgpu_waves_per_eu(kTransportWavesPerEu<kGeom>)))
transport_kernel_plain(KernelParams p) {
    transport_body<kSource, false, kLdsMode, kGeom>(p);
}

template <int kSource, int kLdsMode, int kGeom = kGeomBoxes>
void launch_transport_form(const KernelParams& p, bool persistent, int block, int grid, size_t lds, hipStream_t stream) {
    // the pre-generated instances exist for box and convex-mesh scenes
    if (persistent && p.starts != nullptr && kGeom < kGeomTriangles) {
        hipLaunchKernelGGL((transport_kernel_persistent<kSource, kLdsMode, kGeom, (kGeom < kGeomTriangles)>), dim3(grid),
                           dim3(block), lds, stream, p);
    } else if (persistent) {
        hipLaunchKernelGGL((transport_kernel_persistent<kSource, kLdsMode, kGeom, false>), dim3(grid), dim3(block), lds,
                           stream, p);
    } else {
        hipLaunchKernelGGL((transport_kernel_plain<kSource, kLdsMode, kGeom>), dim3(grid), dim3(block), lds, stream,
                           p);
    }
}

// Scenes run the instances of the lowest geometry level that holds them, so
// box scenes carry no mesh code and convex-mesh scenes no triangle code.
// Triangle scenes have two LDS modes (their triangles rarely fit).
__host__ __device__ inline int geom_level(const SceneView& sc) {
    return sc.n_tris != 0u ? (sc.open_solids != 0u ? kGeomOpen : kGeomTriangles)
                           : (sc.n_planes != 0u ? kGeomConvex : kGeomBoxes);
}

template <int kSource>
void launch_transport_source(const KernelParams& p, bool persistent, int block, int grid, size_t lds,
                             hipStream_t stream) {
    const int geom = geom_level(p.scene);
    if (geom == kGeomOpen) {
        if (lds_mode(p) >= kLdsHot) {
            launch_transport_form<kSource, kLdsHot, kGeomOpen>(p, persistent, block, grid, lds, stream);
        } else {
            launch_transport_form<kSource, kLdsNone, kGeomOpen>(p, persistent, block, grid, lds, stream);
        }
        return;
    }
    if (geom == kGeomTriangles) {
        if (lds_mode(p) >= kLdsHot) {
            launch_transport_form<kSource, kLdsHot, kGeomTriangles>(p, persistent, block, grid, lds, stream);
        } else {
            launch_transport_form<kSource, kLdsNone, kGeomTriangles>(p, persistent, block, grid, lds, stream);
        }
        return;
    }
    if (geom == kGeomConvex) {
        if (lds_mode(p) == kLdsAll) {
            launch_transport_form<kSource, kLdsAll, kGeomConvex>(p, persistent, block, grid, lds, stream);
        } else if (lds_mode(p) == kLdsHot) {
            launch_transport_form<kSource, kLdsHot, kGeomConvex>(p, persistent, block, grid, lds, stream);
        } else {
            launch_transport_form<kSource, kLdsNone, kGeomConvex>(p, persistent, block, grid, lds, stream);
        }
        return;
    }
    switch (lds_mode(p)) {
        case kLdsAll:
            launch_transport_form<kSource, kLdsAll>(p, persistent, block, grid, lds, stream);
            break;
        case kLdsHot:
            launch_transport_form<kSource, kLdsHot>(p, persistent, block, grid, lds, stream);
            break;
        default:
            launch_transport_form<kSource, kLdsNone>(p, persistent, block, grid, lds, stream);
            break;
    }
}

int launch_generate(const KernelParams& p, int source, float4* starts, int grid, hipStream_t stream) {
    if (source == kSourceSynthetic) {
        hipLaunchKernelGGL(generate_kernel<kSourceSynthetic>, dim3(grid), dim3(256), 0, stream, p, starts);
    } else {
        hipLaunchKernelGGL(generate_kernel<kSourceEmission>, dim3(grid), dim3(256), 0, stream, p, starts);
    }
    return hipGetLastError() == hipSuccess ? 0 : 1;
}

__global__ void zero_kernel(uint32_t* ptr, size_t words) {
    for (size_t i = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x; i < words;
         i += static_cast<size_t>(gridDim.x) * blockDim.x) {
        ptr[i] = 0u;
    }
}

__global__ void reduce_event_stats_kernel(const EventStats* rep, uint32_t stride, const uint32_t* header,
                                          EventStats* out) {
    const uint32_t n_events = header[2];
    for (uint32_t e = blockIdx.x * blockDim.x + threadIdx.x; e < n_events; e += gridDim.x * blockDim.x) {
        EventStats s{0, 0, 0, 0, 0, 0, 0};
        for (uint32_t r = 0; r < kEventReplicas; ++r) {
            const EventStats& c = rep[static_cast<size_t>(r) * stride + e];
            s.detected += c.detected;
            s.absorbed += c.absorbed;
            s.escaped += c.escaped;
            s.truncated += c.truncated;
            s.max_bounce = max(s.max_bounce, c.max_bounce);
            s.invalid += c.invalid;
            s.zero_step += c.zero_step;
        }
        out[e] = s;
    }
}

int launch_reduce_event_stats(const EventStats* replicas, uint32_t stride, const uint32_t* header, EventStats* out,
                              hipStream_t stream) {
    const int grid = static_cast<int>(std::min<uint32_t>((stride + 255) / 256, 1024));
    hipLaunchKernelGGL(reduce_event_stats_kernel, dim3(grid), dim3(256), 0, stream, replicas, stride, header, out);
    return hipGetLastError() == hipSuccess ? 0 : 1;
}

int launch_zero(void* ptr, size_t bytes, hipStream_t stream) {
    const size_t words = bytes / 4;
    if (words == 0) {
        return 0;
    }
    const int grid = static_cast<int>(std::min<size_t>((words + 255) / 256, 1024));
    hipLaunchKernelGGL(zero_kernel, dim3(grid), dim3(256), 0, stream, static_cast<uint32_t*>(ptr), words);
    return hipGetLastError() == hipSuccess ? 0 : 1;
}

int launch_transport(const KernelParams& p, int source, bool persistent, int block, int grid, hipStream_t stream) {
    const size_t scene_lds = transport_lds_bytes(p);
    if ((p.starts != nullptr && geom_level(p.scene) >= kGeomTriangles) || block < 64 ||
        block > static_cast<int>(kMaxBlock) || block % 64 != 0) {
        return 1;
    }
    int device = 0, per_cu = 0, per_block = 0;
    if (hipGetDevice(&device) != hipSuccess ||
        hipDeviceGetAttribute(&per_cu, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor, device) != hipSuccess ||
        hipDeviceGetAttribute(&per_block, hipDeviceAttributeMaxSharedMemoryPerBlock, device) != hipSuccess) {
        return 1;
    }
    KernelParams q = p;
    q.lds_stage_quads = static_cast<uint32_t>((scene_lds + 15u) / 16u);
    const size_t waves = static_cast<size_t>(block / 64);
    size_t lds = static_cast<size_t>(q.lds_stage_quads) * 16u + waves * kStageBytesPerWave;
    if (lds + kLdsStaticReserve > static_cast<size_t>(per_block)) {
        return 1;  // the staged scene was planned for a larger block
    }
    q.ring_slots = 0;
    if (persistent && geom_level(p.scene) < kGeomTriangles) {
        // the ring takes what the block's share of the CU's LDS leaves (the
        // triangle-solid instances refill in place and have no ring)
        const size_t resident = std::max<size_t>(1, kResidentWaves / waves);
        const size_t share = std::min(static_cast<size_t>(per_cu) / resident, static_cast<size_t>(per_block));
        if (share > lds + kLdsStaticReserve) {
            const size_t slots = (share - lds - kLdsStaticReserve) / (waves * kRingQuads * 16u);
            q.ring_slots = static_cast<uint32_t>(std::min<size_t>(kRingMaxSlots, slots));
        }
        lds += waves * q.ring_slots * kRingQuads * 16u;
    }
    if (source == kSourceSynthetic) {
        launch_transport_source<kSourceSynthetic>(q, persistent, block, grid, lds, stream);
    } else {
        launch_transport_source<kSourceEmission>(q, persistent, block, grid, lds, stream);
    }
    return hipGetLastError() == hipSuccess ? 0 : 1;
}

// ---- path-replay adjoint ----------------------------------------------------
// Each detected photon of the forward batch is regenerated from its Philox
// counters and transported again with the same arithmetic, so it follows the
// recorded path bit for bit (checked against the recorded end point and
// time).  Along the way the likelihood-ratio (score-function) derivatives of
// the path density are collected:
//   bulk segment of length s in material m, ended at a boundary or by a
//     Rayleigh scatter: d log p / dL_abs = s / L_abs^2,
//     d log p / dL_ray = s / L_ray^2 - [scatter] / L_ray;
//   surface classified as interaction (reflection): d log p / dR = 1 / R;
//   detection: absorption 1 - R - T then efficiency: d/dR = -1 / (1 - R - T),
//     d/d eff = 1 / eff;
// each chained to the two interpolation nodes of the table at the photon
// energy and weighted by dLoss/dN + dLoss/dh_b of the photon's arrival-time
// slot b, read from the recorded hit before the replay.
//
// A replay first sums its path into two material and two surface slots and
// scores them once it has matched its recorded end.  A path that visits more
// materials or parameterised surfaces than that is replayed a second time
// (it is known to match by then) and scored segment by segment as it goes.
//
// The kernel is persistent like the forward one: each wave takes chunks of
// the hit list from a device counter and refills lanes whose replay ended, so
// a wave's lanes stay busy whatever the spread of path lengths.
constexpr uint32_t kAdjointLdsParams = 4096;
constexpr uint32_t kAdjointChunk = 64;   // hits a wave takes from the queue at a time
constexpr uint32_t kAdjointRefill = 24;  // idle lanes that trigger a refill

// 16-byte words of LDS that the adjoint's gradient accumulator takes (0: it
// accumulates in global memory); the staged scene follows it.
__host__ __device__ inline uint32_t adjoint_grad_quads(uint32_t n_params) {
    return n_params <= kAdjointLdsParams ? (n_params + 1u) / 2u : 0u;
}

namespace {

struct MatSlot {
    uint32_t material;
    uint32_t rayleigh;
    double length;
};

struct SurfSlot {
    uint32_t surface;
    uint32_t count;
};

__device__ __forceinline__ void add_table_grad(double* grad, const float* pool, const Prop& prop, uint32_t offset,
                                               float e, double g) {
    uint32_t i;
    float w;
    prop_node(pool, prop, e, i, w);
    atomicAdd(grad + offset + i, g * (1.0 - static_cast<double>(w)));
    if (w != 0.0f) {
        atomicAdd(grad + offset + i + 1, g * static_cast<double>(w));
    }
}

__device__ __forceinline__ bool add_segment(MatSlot& a, MatSlot& b, uint32_t material, float length, bool rayleigh) {
    MatSlot* slot = a.material == material ? &a : (b.material == material ? &b : nullptr);
    if (slot == nullptr) {
        slot = a.material == kInvalid ? &a : (b.material == kInvalid ? &b : nullptr);
        if (slot == nullptr) {
            return false;
        }
        slot->material = material;
    }
    slot->length += static_cast<double>(length);
    slot->rayleigh += rayleigh ? 1u : 0u;
    return true;
}

__device__ __forceinline__ bool add_reflection(SurfSlot& a, SurfSlot& b, uint32_t surface) {
    SurfSlot* slot = a.surface == surface ? &a : (b.surface == surface ? &b : nullptr);
    if (slot == nullptr) {
        slot = a.surface == kInvalid ? &a : (b.surface == kInvalid ? &b : nullptr);
        if (slot == nullptr) {
            return false;
        }
        slot->surface = surface;
    }
    ++slot->count;
    return true;
}

// Source of launch index g: last s with source_photon_offset[s] <= g.
__device__ __forceinline__ uint32_t launch_source(const KernelParams& p, uint32_t g, uint32_t n_sources) {
    uint32_t lo = 0, hi = n_sources;
    while (hi - lo > 1u) {
        const uint32_t mid = (lo + hi) >> 1;
        if (p.source_photon_offset[mid] <= g) {
            lo = mid;
        } else {
            hi = mid;
        }
    }
    return lo;
}

__device__ __forceinline__ void add_material_grads(double* grad, const SceneView& sc, const ParamMap& map,
                                                   const MatSlot& m, float e, double weight) {
    if (m.material == kInvalid || m.material >= sc.n_materials) {
        return;
    }
    const Material& mat = sc.materials[m.material];
    const uint32_t oa = map.material_abs[m.material];
    if (oa != kInvalid && mat.abs_length.count != 0) {
        const float la = prop_eval(sc.pool, mat.abs_length, e, kInf);
        if (finitef(la) && la > 0.0f) {
            const double l = la;
            add_table_grad(grad, sc.pool, mat.abs_length, oa, e, weight * m.length / (l * l));
        }
    }
    const uint32_t orr = map.material_rayleigh[m.material];
    if (orr != kInvalid && mat.rayleigh_length.count != 0) {
        const float lr = prop_eval(sc.pool, mat.rayleigh_length, e, kInf);
        if (finitef(lr) && lr > 0.0f) {
            const double l = lr;
            add_table_grad(grad, sc.pool, mat.rayleigh_length, orr, e,
                           weight * (m.length / (l * l) - static_cast<double>(m.rayleigh) / l));
        }
    }
}

__device__ __forceinline__ void add_reflection_grads(double* grad, const SceneView& sc, const ParamMap& map,
                                                     const SurfSlot& r, float e, double weight) {
    if (r.surface == kInvalid) {
        return;
    }
    const Surface& sf = sc.surfaces[r.surface];
    const float refl = prop_eval(sc.pool, sf.reflectivity, e, 0.0f);
    if (refl > 0.0f && refl < 1.0f) {
        add_table_grad(grad, sc.pool, sf.reflectivity, map.surface_reflectivity[r.surface], e,
                       weight * static_cast<double>(r.count) / static_cast<double>(refl));
    }
}

__device__ __forceinline__ void add_detection_grads(double* grad, const SceneView& sc, const ParamMap& map,
                                                    uint32_t surface, float e, double weight) {
    if (surface >= sc.n_surfaces) {
        return;
    }
    const Surface& sf = sc.surfaces[surface];
    const uint32_t oe = map.surface_efficiency[surface];
    if (oe != kInvalid && sf.efficiency.count != 0) {
        const float eff = prop_eval(sc.pool, sf.efficiency, e, 0.0f);
        if (eff > 0.0f && eff < 1.0f) {
            add_table_grad(grad, sc.pool, sf.efficiency, oe, e, weight / static_cast<double>(eff));
        }
    }
    const uint32_t orf = map.surface_reflectivity[surface];
    if (orf != kInvalid && sf.reflectivity.count != 0) {
        const float r = prop_eval(sc.pool, sf.reflectivity, e, 0.0f);
        const float t = sf.transmittance.count != 0 ? clamp01(prop_eval(sc.pool, sf.transmittance, e, 0.0f)) : 0.0f;
        const double absorb = 1.0 - static_cast<double>(r) - static_cast<double>(t);
        if (r > 0.0f && r < 1.0f && absorb > 0.0) {
            add_table_grad(grad, sc.pool, sf.reflectivity, orf, e, -weight / absorb);
        }
    }
}

}  // namespace

// Four waves per SIMD (at most 128 VGPRs), as for the persistent transport
// kernel: the replay runs the same step code.
template <int kSource, int kLdsMode, int kGeom>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) adjoint_kernel(AdjointParams a) {
    extern __shared__ double s_grad[];  // gradient accumulator, then the staged scene
    const uint32_t n_params = a.map.n_params;
    const bool in_lds = n_params <= kAdjointLdsParams;
    double* grad = in_lds ? s_grad : a.grad;
    if (in_lds) {
        for (uint32_t i = threadIdx.x; i < n_params; i += blockDim.x) {
            s_grad[i] = 0.0;
        }
    }
    const KernelParams& p = a.fwd;
    SceneView sc = p.scene;
    stage_scene<kLdsMode, kGeom>(p, reinterpret_cast<float4*>(s_grad) + adjoint_grad_quads(n_params), sc);
    __syncthreads();
    const uint32_t n_sources = p.header != nullptr ? p.header[0] : 0u;
    const uint32_t lane = lane_id();
    uint32_t replayed = 0, mismatch = 0, rescored = 0, steps = 0;

    // per-lane replay: hit index, pass (direct: scored step by step), path sums
    Photon ph;
    ph.material = kInvalid;
    bool alive = false, pending = false, direct = false, ok = true;
    uint32_t h = 0, det_surface = kInvalid;
    double weight = 0.0;
    MatSlot m0{kInvalid, 0u, 0.0}, m1{kInvalid, 0u, 0.0};
    SurfSlot r0{kInvalid, 0u}, r1{kInvalid, 0u};
    // wave-uniform cursor over the hit list
    uint32_t it_next = 0, it_end = 0;
    bool done = false;

    for (;;) {
        // ---- idle lanes take the next hits of the wave's chunk ----
        uint64_t idle = __ballot(!alive && !pending);
        uint32_t n_idle = static_cast<uint32_t>(__popcll(idle));
        if (n_idle >= kAdjointRefill || n_idle == 64u) {
            while (n_idle > 0u) {
                if (it_next >= it_end) {
                    if (done) {
                        break;
                    }
                    uint32_t c = 0;
                    if (lane == 0u) {
                        c = atomicAdd(reinterpret_cast<uint32_t*>(a.counters + 4), kAdjointChunk);
                    }
                    c = uniform(static_cast<uint32_t>(__shfl(static_cast<int>(c), 0)));
                    if (c >= a.n_hits) {
                        done = true;
                        break;
                    }
                    it_next = c;
                    it_end = min(a.n_hits, c + kAdjointChunk);
                    continue;
                }
                const uint32_t take = min(it_end - it_next, n_idle);
                const uint32_t rank = rank_below(idle);
                if (!alive && !pending && rank < take) {
                    h = it_next + rank;
                    pending = true;
                    direct = false;
                }
                it_next += take;
                idle = __ballot(!alive && !pending);
                n_idle = static_cast<uint32_t>(__popcll(idle));
            }
        }
        // ---- start the replays of new hits and second passes ----
        if (pending) {
            pending = false;
            const Hit& hit = a.hits[h];
            if (kSource == kSourceSynthetic) {
                generate_synthetic(ph, p.synth, p.rng, hit.event_id, hit.photon_id);
            } else {
                const uint32_t src = launch_source(p, a.hit_launch[h], n_sources);
                const Emission& em = p.emissions[src];
                generate_emission(ph, em, sc, p.rng, hit.photon_id - em.first_photon_id);
            }
            ph.material = kInvalid;
            if (start_photon(ph)) {
                alive = true;
                ok = true;
                det_surface = kInvalid;
                m0 = MatSlot{kInvalid, 0u, 0.0};
                m1 = MatSlot{kInvalid, 0u, 0.0};
                r0 = SurfSlot{kInvalid, 0u};
                r1 = SurfSlot{kInvalid, 0u};
                weight = a.slot_weight[hist_slot(hit.time)];
            } else {
                ++mismatch;
            }
        }
        if (__ballot(alive) == 0ull) {
            if (done) {
                break;
            }
            continue;
        }

        // ---- one replayed step ----
        if (alive) {
            bool zero = false;
            uint32_t sensor = kInvalid;
            StepTrace tr;
            const int status = transport_step<true, kGeom>(ph, sc, p.rng, p.cfg, zero, sensor, &tr);
            ++ph.step;
            ++steps;
            const float e = ph.energy;
            if ((tr.segment > 0.0f || tr.event == kTraceRayleigh) && tr.material < sc.n_materials) {
                if (direct) {
                    add_material_grads(grad, sc, a.map,
                                       MatSlot{tr.material, tr.event == kTraceRayleigh ? 1u : 0u,
                                               static_cast<double>(tr.segment)},
                                       e, weight);
                } else {
                    ok = add_segment(m0, m1, tr.material, tr.segment, tr.event == kTraceRayleigh) && ok;
                }
            }
            if (tr.event == kTraceInteraction && tr.surface < sc.n_surfaces &&
                a.map.surface_reflectivity[tr.surface] != kInvalid) {
                if (direct) {
                    add_reflection_grads(grad, sc, a.map, SurfSlot{tr.surface, 1u}, e, weight);
                } else {
                    ok = add_reflection(r0, r1, tr.surface) && ok;
                }
            }
            if (status == kDetected) {
                det_surface = tr.surface;
            }
            if (status != kAlive || ph.step >= p.max_steps) {
                alive = false;
                const Hit& hit = a.hits[h];
                if (status != kDetected || ph.time != hit.time || ph.pos.x != hit.pos[0] || ph.pos.y != hit.pos[1] ||
                    ph.pos.z != hit.pos[2]) {
                    ++mismatch;
                } else if (!ok) {
                    pending = true;  // replay again, scoring each segment as it goes
                    direct = true;
                } else {
                    if (!direct) {
                        add_material_grads(grad, sc, a.map, m0, e, weight);
                        add_material_grads(grad, sc, a.map, m1, e, weight);
                        add_reflection_grads(grad, sc, a.map, r0, e, weight);
                        add_reflection_grads(grad, sc, a.map, r1, e, weight);
                    }
                    add_detection_grads(grad, sc, a.map, det_surface, e, weight);
                    ++replayed;
                    rescored += direct ? 1u : 0u;
                }
            }
        }
    }
    const uint32_t wr = wave_sum(replayed), wm = wave_sum(mismatch), wu = wave_sum(rescored), ws = wave_sum(steps);
    if (lane == 0u) {
        if (wr) atomicAdd(a.counters + 0, static_cast<unsigned long long>(wr));
        if (wm) atomicAdd(a.counters + 1, static_cast<unsigned long long>(wm));
        if (wu) atomicAdd(a.counters + 2, static_cast<unsigned long long>(wu));
        if (ws) atomicAdd(a.counters + 3, static_cast<unsigned long long>(ws));
    }
    if (in_lds) {
        __syncthreads();
        for (uint32_t i = threadIdx.x; i < n_params; i += blockDim.x) {
            if (s_grad[i] != 0.0) {
                atomicAdd(a.grad + i, s_grad[i]);
            }
        }
    }
}

template <int kSource>
void launch_adjoint_source(const AdjointParams& p, int block, int grid, size_t lds, hipStream_t stream) {
    const int geom = geom_level(p.fwd.scene);
    if (geom == kGeomOpen) {
        if (lds_mode(p.fwd) >= kLdsHot) {
            hipLaunchKernelGGL((adjoint_kernel<kSource, kLdsHot, kGeomOpen>), dim3(grid), dim3(block), lds, stream, p);
        } else {
            hipLaunchKernelGGL((adjoint_kernel<kSource, kLdsNone, kGeomOpen>), dim3(grid), dim3(block), lds, stream, p);
        }
        return;
    }
    if (geom == kGeomTriangles) {
        if (lds_mode(p.fwd) >= kLdsHot) {
            hipLaunchKernelGGL((adjoint_kernel<kSource, kLdsHot, kGeomTriangles>), dim3(grid), dim3(block), lds, stream,
                               p);
        } else {
            hipLaunchKernelGGL((adjoint_kernel<kSource, kLdsNone, kGeomTriangles>), dim3(grid), dim3(block), lds,
                               stream, p);
        }
        return;
    }
    if (geom == kGeomConvex) {
        if (lds_mode(p.fwd) == kLdsAll) {
            hipLaunchKernelGGL((adjoint_kernel<kSource, kLdsAll, kGeomConvex>), dim3(grid), dim3(block), lds, stream, p);
        } else if (lds_mode(p.fwd) == kLdsHot) {
            hipLaunchKernelGGL((adjoint_kernel<kSource, kLdsHot, kGeomConvex>), dim3(grid), dim3(block), lds, stream, p);
        } else {
            hipLaunchKernelGGL((adjoint_kernel<kSource, kLdsNone, kGeomConvex>), dim3(grid), dim3(block), lds, stream,
                               p);
        }
        return;
    }
    switch (lds_mode(p.fwd)) {
        case kLdsAll:
            hipLaunchKernelGGL((adjoint_kernel<kSource, kLdsAll, kGeomBoxes>), dim3(grid), dim3(block), lds, stream, p);
            break;
        case kLdsHot:
            hipLaunchKernelGGL((adjoint_kernel<kSource, kLdsHot, kGeomBoxes>), dim3(grid), dim3(block), lds, stream, p);
            break;
        default:
            hipLaunchKernelGGL((adjoint_kernel<kSource, kLdsNone, kGeomBoxes>), dim3(grid), dim3(block), lds, stream, p);
            break;
    }
}

int launch_adjoint(const AdjointParams& p, int source, int block, int grid, hipStream_t stream) {
    // the gradient accumulator (when it fits), then the scene the forward launch staged
    const size_t lds = static_cast<size_t>(adjoint_grad_quads(p.map.n_params)) * 16u + transport_lds_bytes(p.fwd);
    if (source == kSourceSynthetic) {
        launch_adjoint_source<kSourceSynthetic>(p, block, grid, lds, stream);
    } else {
        launch_adjoint_source<kSourceEmission>(p, block, grid, lds, stream);
    }
    return hipGetLastError() == hipSuccess ? 0 : 1;
}

// ---- launch-order output ----------------------------------------------------
// (1) mark: one bit per detected launch index; (2) per block of
// kOrderWordsPerBlock mask words, the exclusive popcount prefix of each word
// and the block total; (3) exclusive prefix over the block totals (one block);
// (4) scatter: the rank of a hit among the detected launch indices is
// block_prefix + word_prefix + popcount of the lower bits of its word.
__global__ void order_mark_kernel(OrderParams p) {
    const uint32_t n = min(*p.hit_count, p.capacity);
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const uint32_t g = p.launch[i];
        if (g < p.n_photons) {
            atomicOr(p.mask + (g >> 5), 1u << (g & 31u));
        }
    }
}

__global__ void __launch_bounds__(256) order_word_prefix_kernel(OrderParams p, uint32_t n_words) {
    __shared__ uint32_t s_sum[256];
    const uint32_t base = blockIdx.x * kOrderWordsPerBlock + threadIdx.x * 4u;  // 4 words per thread
    uint32_t c[4];
    uint32_t total = 0;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const uint32_t w = base + k;
        c[k] = w < n_words ? static_cast<uint32_t>(__popc(p.mask[w])) : 0u;
        total += c[k];
    }
    s_sum[threadIdx.x] = total;
    __syncthreads();
    // inclusive Hillis-Steele scan over the 256 thread totals
    for (uint32_t off = 1; off < 256u; off <<= 1) {
        const uint32_t v = threadIdx.x >= off ? s_sum[threadIdx.x - off] : 0u;
        __syncthreads();
        s_sum[threadIdx.x] += v;
        __syncthreads();
    }
    uint32_t run = s_sum[threadIdx.x] - total;  // exclusive prefix of this thread's words
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const uint32_t w = base + k;
        if (w < n_words) {
            p.word_prefix[w] = run;
        }
        run += c[k];
    }
    if (threadIdx.x == 255u) {
        p.block_prefix[blockIdx.x] = s_sum[255];  // block total, turned into a prefix next
    }
}

__global__ void __launch_bounds__(1024) order_block_prefix_kernel(OrderParams p, uint32_t n_blocks) {
    __shared__ uint32_t s_sum[1024];
    uint32_t carry = 0;
    for (uint32_t start = 0; start < n_blocks; start += 1024u) {
        const uint32_t i = start + threadIdx.x;
        const uint32_t v = i < n_blocks ? p.block_prefix[i] : 0u;
        s_sum[threadIdx.x] = v;
        __syncthreads();
        for (uint32_t off = 1; off < 1024u; off <<= 1) {
            const uint32_t u = threadIdx.x >= off ? s_sum[threadIdx.x - off] : 0u;
            __syncthreads();
            s_sum[threadIdx.x] += u;
            __syncthreads();
        }
        if (i < n_blocks) {
            p.block_prefix[i] = carry + s_sum[threadIdx.x] - v;
        }
        const uint32_t chunk_total = s_sum[1023];
        __syncthreads();
        carry += chunk_total;
    }
}

__global__ void order_scatter_kernel(OrderParams p) {
    const uint32_t n = min(*p.hit_count, p.capacity);
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const uint32_t g = p.launch[i];
        if (g >= p.n_photons) {
            continue;
        }
        const uint32_t w = g >> 5;
        const uint32_t below = p.mask[w] & ((1u << (g & 31u)) - 1u);
        const uint32_t rank = p.block_prefix[w / kOrderWordsPerBlock] + p.word_prefix[w] +
                              static_cast<uint32_t>(__popc(below));
        if (rank < n) {
            p.out_hits[rank] = p.hits[i];
            p.out_launch[rank] = g;
        }
    }
}

__global__ void __launch_bounds__(kPrefixMaxSources) synthetic_prefix_kernel(const uint32_t* counts, uint32_t n,
                                                                               uint32_t chunk, uint32_t* header,
                                                                               uint32_t* item_offsets,
                                                                               uint32_t* photon_offsets) {
    __shared__ uint32_t s_items[kPrefixMaxSources];
    __shared__ uint32_t s_photons[kPrefixMaxSources];
    const uint32_t t = threadIdx.x;
    const uint32_t c = t < n ? counts[t] : 0u;
    s_items[t] = static_cast<uint32_t>((static_cast<uint64_t>(c) + chunk - 1u) / chunk);
    s_photons[t] = c;
    __syncthreads();
    // inclusive Hillis-Steele scan over the block
    for (uint32_t off = 1; off < kPrefixMaxSources; off <<= 1) {
        const uint32_t a = t >= off ? s_items[t - off] : 0u;
        const uint32_t b = t >= off ? s_photons[t - off] : 0u;
        __syncthreads();
        s_items[t] += a;
        s_photons[t] += b;
        __syncthreads();
    }
    if (t < n) {
        item_offsets[t + 1] = s_items[t];
        photon_offsets[t + 1] = s_photons[t];
    }
    if (t == 0) {
        item_offsets[0] = 0u;
        photon_offsets[0] = 0u;
        header[0] = n;
        header[1] = chunk;
        header[2] = n;
        header[3] = 0u;
    }
}

int launch_synthetic_prefix(const uint32_t* counts, uint32_t n, uint32_t chunk, uint32_t* header,
                            uint32_t* item_offsets, uint32_t* photon_offsets, hipStream_t stream) {
    if (n > kPrefixMaxSources || chunk == 0u) {
        return 1;
    }
    hipLaunchKernelGGL(synthetic_prefix_kernel, dim3(1), dim3(kPrefixMaxSources), 0, stream, counts, n, chunk, header,
                       item_offsets, photon_offsets);
    return hipGetLastError() == hipSuccess ? 0 : 1;
}

__global__ void export_counters_kernel(const unsigned long long* counters, unsigned long long* out) {
    if (threadIdx.x == 0) {
        out[0] = counters[kCounterSteps];
        out[1] = counters[kCounterPhotons];
        out[2] = counters[kCounterDropped];
    }
}

int launch_export_counters(const unsigned long long* counters, unsigned long long* out, hipStream_t stream) {
    hipLaunchKernelGGL(export_counters_kernel, dim3(1), dim3(64), 0, stream, counters, out);
    return hipGetLastError() == hipSuccess ? 0 : 1;
}

#if defined(G4GO_REGION_PROFILE)
// Counter build only: read and clear the region profile (kRegionCount x 3 u64).
extern "C" int g4go_region_profile_read(unsigned long long* out) {
    return hipMemcpyFromSymbol(out, HIP_SYMBOL(g4go_region_global), sizeof(g4go_region_global)) == hipSuccess ? 0 : 1;
}
extern "C" int g4go_region_profile_clear() {
    static const unsigned long long zeros[kRegionCount][3] = {};
    return hipMemcpyToSymbol(HIP_SYMBOL(g4go_region_global), zeros, sizeof(zeros)) == hipSuccess ? 0 : 1;
}
#endif

int launch_order_hits(const OrderParams& p, hipStream_t stream) {
    const uint32_t n_words = (p.n_photons + 31u) / 32u;
    const uint32_t n_blocks = (n_words + kOrderWordsPerBlock - 1) / kOrderWordsPerBlock;
    if (n_words == 0) {
        return 0;
    }
    const int grid = static_cast<int>(std::min<uint64_t>((static_cast<uint64_t>(p.capacity) + 255) / 256, 4096));
    hipLaunchKernelGGL(zero_kernel, dim3(std::max<uint32_t>(1, std::min<uint32_t>((n_words + 255) / 256, 4096))),
                       dim3(256), 0, stream, p.mask, static_cast<size_t>(n_words));
    hipLaunchKernelGGL(order_mark_kernel, dim3(std::max(grid, 1)), dim3(256), 0, stream, p);
    hipLaunchKernelGGL(order_word_prefix_kernel, dim3(n_blocks), dim3(256), 0, stream, p, n_words);
    hipLaunchKernelGGL(order_block_prefix_kernel, dim3(1), dim3(1024), 0, stream, p, n_blocks);
    hipLaunchKernelGGL(order_scatter_kernel, dim3(std::max(grid, 1)), dim3(256), 0, stream, p);
    return hipGetLastError() == hipSuccess ? 0 : 1;
}

// ---- step-by-step trace of one synthetic photon (tests and diagnostics) ----
// Per step, 16 floats: pos[3], dir[3], pol[3], time (float), status, zero,
// step, volume, energy, 0.
__global__ void trace_kernel(KernelParams p, uint32_t event_id, uint32_t photon_id, uint32_t max_steps, float* out,
                             uint32_t* n_out) {
    if (blockIdx.x != 0 || threadIdx.x != 0) {
        return;
    }
    Photon ph;
    generate_synthetic(ph, p.synth, p.rng, event_id, photon_id);
    ph.material = kInvalid;
    uint32_t n = 0;
    if (start_photon(ph)) {
        for (; n < max_steps;) {
            bool zero = false;
            uint32_t sensor = kInvalid;
            const int geom = geom_level(p.scene);
            const int status =
                geom == kGeomOpen
                    ? transport_step<false, kGeomOpen>(ph, p.scene, p.rng, p.cfg, zero, sensor, nullptr)
                : geom == kGeomTriangles
                    ? transport_step<false, kGeomTriangles>(ph, p.scene, p.rng, p.cfg, zero, sensor, nullptr)
                    : (geom == kGeomConvex
                           ? transport_step<false, kGeomConvex>(ph, p.scene, p.rng, p.cfg, zero, sensor, nullptr)
                           : transport_step<false, kGeomBoxes>(ph, p.scene, p.rng, p.cfg, zero, sensor, nullptr));
            ++ph.step;
            float* o = out + 16 * n;
            o[0] = ph.pos.x;
            o[1] = ph.pos.y;
            o[2] = ph.pos.z;
            o[3] = ph.dir.x;
            o[4] = ph.dir.y;
            o[5] = ph.dir.z;
            o[6] = ph.pol.x;
            o[7] = ph.pol.y;
            o[8] = ph.pol.z;
            o[9] = static_cast<float>(ph.time);
            o[10] = static_cast<float>(status);
            o[11] = zero ? 1.0f : 0.0f;
            o[12] = static_cast<float>(ph.step);
            o[13] = static_cast<float>(ph.volume);
            o[14] = ph.energy;
            o[15] = 0.0f;
            ++n;
            if (status != kAlive) {
                break;
            }
        }
    }
    *n_out = n;
}

int launch_trace(const KernelParams& p, uint32_t event_id, uint32_t photon_id, uint32_t max_steps, float* out,
                 uint32_t* n_out, hipStream_t stream) {
    hipLaunchKernelGGL(trace_kernel, dim3(1), dim3(64), 0, stream, p, event_id, photon_id, max_steps, out, n_out);
    return hipGetLastError() == hipSuccess ? 0 : 1;
}

// ---- element-wise evaluation kernels used by the parity tests ----
__global__ void eval_math_kernel(const float* in, float* out, int n, int which) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) {
        return;
    }
    const float x = in[i];
    if (which == 0) {
        out[i] = g4_logf(x);
    } else if (which == 1) {
        out[i] = g4_expf(x);
    } else {
        float s, c;
        g4_sincos_turn(x, s, c);
        out[2 * i] = s;
        out[2 * i + 1] = c;
    }
}

int launch_eval_math(const float* in, float* out, int n, int which, hipStream_t stream) {
    if (n <= 0) {
        return 0;
    }
    hipLaunchKernelGGL(eval_math_kernel, dim3((n + 255) / 256), dim3(256), 0, stream, in, out, n, which);
    return hipGetLastError() == hipSuccess ? 0 : 1;
}

// in: 11 floats per case (dir[3], pol[3], normal[3], n1, n2)
// out: 14 floats per case (T, tir, reflected dir[3], transmitted dir[3], reflected pol[3], transmitted pol[3])
__global__ void eval_fresnel_kernel(const float* in, float* out, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) {
        return;
    }
    const float* a = in + 11 * i;
    const Fresnel f = compute_fresnel(V3{a[0], a[1], a[2]}, V3{a[3], a[4], a[5]}, V3{a[6], a[7], a[8]}, a[9], a[10]);
    float* o = out + 14 * i;
    o[0] = static_cast<float>(f.transmittance);
    o[1] = f.tir ? 1.0f : 0.0f;
    const V3 vs[4] = {f.reflected_dir, f.transmitted_dir, f.reflected_pol, f.transmitted_pol};
    for (int k = 0; k < 4; ++k) {
        o[2 + 3 * k] = vs[k].x;
        o[3 + 3 * k] = vs[k].y;
        o[4 + 3 * k] = vs[k].z;
    }
}

int launch_eval_fresnel(const float* in, float* out, int n, hipStream_t stream) {
    if (n <= 0) {
        return 0;
    }
    hipLaunchKernelGGL(eval_fresnel_kernel, dim3((n + 255) / 256), dim3(256), 0, stream, in, out, n);
    return hipGetLastError() == hipSuccess ? 0 : 1;
}

}  // namespace g4go
