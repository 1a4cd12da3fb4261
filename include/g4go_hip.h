/*
 * g4go_hip.h -- C ABI of the MI355X (gfx950) optical-photon transport backend.
 *
 * This is the drop-in boundary for G4GO's GPU optical backend.  The reference
 * binds its backend through the C++ class
 *     G4GO::Optical::OptiXTransportHost
 * (include/g4go/optical/optix/OptiXTransportHost.hpp:10-26) with the methods
 *     OptiXTransportHost(PhotonTransportConfig)      -> g4go_create
 *     PrepareScene(const Scene&)                      -> g4go_prepare_scene
 *     EnqueueEmissions(const Scene&, const Batch&)    -> g4go_enqueue_emissions
 *     CompleteOldestBatch() -> PhotonTransportOutput  -> g4go_complete
 * and the CUB compaction helpers QueryOptiXHitCompactionBytes/CompactOptiXHits
 * (include/g4go/optical/optix/OptiXHitCompaction.hpp:10-14), which this
 * backend fuses into the transport kernel (no separate entry point).
 *
 * Every function returns a g4go_status; on failure g4go_last_error() holds a
 * message.  No C++ exception crosses this boundary.  Plain pointers and sizes
 * only; all host buffers passed in are borrowed for the duration of the call.
 */
#ifndef G4GO_HIP_H
#define G4GO_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes, mapped by the C++ shim onto the reference's exception types
 * (OptiXTransportHost.cpp:47-52, 574-630, 514-525, 780-781). */
typedef enum g4go_status {
    G4GO_OK = 0,
    G4GO_ERROR_RUNTIME = 1,          /* std::runtime_error: HIP failure, device fault */
    G4GO_ERROR_INVALID_ARGUMENT = 2, /* std::invalid_argument: bad scene or batch */
    G4GO_ERROR_OVERFLOW = 3,         /* std::overflow_error: > 2^32-1 photons/emissions/events */
    G4GO_ERROR_LOGIC = 4             /* std::logic_error: queue full/empty, no scene */
} g4go_status;

#define G4GO_INVALID_ID 0xffffffffu

typedef struct g4go_transport g4go_transport;

/* PhotonTransportConfig (include/g4go/optical/PhotonTransportConfig.hpp:26-40),
 * plus the launch form of this backend. */
typedef struct g4go_config {
    uint64_t seed;
    uint32_t max_photons_per_event;
    uint32_t max_bounces_per_photon;
    float boundary_tolerance_mm;
    uint32_t target_photons_per_batch;
    uint32_t max_events_per_batch;
    uint32_t enable_performance_diagnostics;
    int32_t device;        /* HIP device ordinal, -1 = current */
    uint32_t sort_hits;    /* 1: detections in launch order -- the order of their photons in the batch
                              (records in order, photons in order), as the reference's stable
                              compaction returns them (OptiXHitCompaction.cu:8-30); 0: unordered */
    uint32_t max_in_flight; /* batches enqueued before the oldest completes: 0 or 1 = the reference's
                              single slot (OptiXTransportHost.cpp:523-525), up to 4: a ring of slots,
                              each with its own stream and buffers, completed in enqueue order */
} g4go_config;

/* Launch form.
 *   persistent=0: plain form, one wave per work item of <= 64 photons over a
 *     grid covering the batch; with per_event=1, one such launch per event
 *     (the per-event launches a caller without batching would make).
 *   persistent=1: blocks_per_cu x CU-count blocks pop work items of `chunk`
 *     photons from a device queue and refill idle lanes once at least
 *     refill_threshold of the 64 lanes are idle: from the wave's ring of
 *     ready photons in LDS while it holds any, else by a generation pass in
 *     which the whole wave generates the idle lanes' photons and the ring's
 *     next ones (scenes with triangle solids: the idle lanes generate their
 *     own); with graph=1 the launch is captured once in a hipGraph and
 *     replayed for every batch whose buffers still fit. */
typedef struct g4go_launch {
    uint32_t persistent;
    uint32_t block_size;       /* 64, 128, 256, 512 */
    uint32_t chunk;            /* photons per work item (persistent form) */
    uint32_t refill_threshold; /* 1..64 */
    uint32_t blocks_per_cu;    /* persistent grid = blocks_per_cu * CU count */
    uint32_t per_event;        /* plain form: one launch per event */
    uint32_t graph;            /* persistent form: replay a captured hipGraph */
    uint32_t pregenerate;      /* persistent form: 1 = a generation pass on full waves writes every
                                  photon's start state (48 B in HBM) before the transport, whose idle
                                  lanes load them at refill; 0 = idle lanes generate their photons at
                                  refill (always so in scenes with triangle solids) */
} g4go_launch;

/* A property table: `count` (energy eV, value) pairs with ascending energies;
 * count = 0 means the property is absent.  Mirrors PropertyTable
 * (include/g4go/optical/Scene.hpp:61-67). */
typedef struct g4go_property {
    uint32_t count;
    const float* energy_ev;
    const float* values;
} g4go_property;

/* Material (Scene.hpp:69-76) plus a Rayleigh scattering length, which the
 * reference exporter rejects (Geant4SceneExporter.cpp:501-503). */
typedef struct g4go_material {
    g4go_property rindex;
    g4go_property group_velocity_mm_per_ns;
    g4go_property abs_length_mm;
    g4go_property rayleigh_length_mm;
    g4go_property scintillation_spectrum[3]; /* raw spectra; the CDF is built here */
} g4go_material;

/* Surface (Scene.hpp:78-92). type: 0 dielectric_dielectric, 1 dielectric_metal;
 * model: 0 glisur, 1 unified; finish: 0 polished, 1 ground. */
typedef struct g4go_surface {
    uint32_t type, model, finish;
    float model_value;
    g4go_property reflectivity, efficiency, transmittance;
    g4go_property specular_lobe, specular_spike, backscatter, surface_roughness;
} g4go_surface;

/* Volume (Scene.hpp:107-120).  Its solid is an axis-aligned box in world
 * coordinates (centre and half lengths, mm), unless a g4go_volume_mesh of
 * the scene names the volume: then the mesh is its solid and centre and half
 * lengths are ignored. */
typedef struct g4go_volume {
    uint32_t volume_id;
    uint32_t material_id;
    uint32_t parent_volume_id; /* G4GO_INVALID_ID for the world */
    uint32_t skin_surface_id;  /* G4GO_INVALID_ID = none */
    uint32_t sensor_id;        /* G4GO_INVALID_ID = not a sensor */
    float center_mm[3];
    float half_mm[3];
} g4go_volume;

typedef struct g4go_surface_binding { /* SurfaceBinding (Scene.hpp:122-126) */
    uint32_t from_volume_id, to_volume_id, surface_id;
} g4go_surface_binding;

/* Closed triangle mesh of one volume, in world coordinates: the volume's
 * MeshGeometry (Scene.hpp:94-100) mapped through its Transform
 * (Scene.hpp:42-59), as the reference places it with an instance transform
 * (OptiXTransportHost.cpp:1221-1262).  A convex mesh (no vertex outside the
 * plane of a triangle by more than 1e-6 of the mesh size, at most 1024
 * distinct face planes) is held as its face planes and may be wound either
 * way; any other mesh is held as its triangles and must be wound
 * consistently (each edge traversed once each way; inwards or outwards).
 * The boundary normal is turned to face the photon, as in the reference
 * (OptiXTransportDevice.cu:789-791).  Rejected with
 * G4GO_ERROR_INVALID_ARGUMENT: fewer than four triangles, an index out of
 * range, a degenerate triangle, an edge not shared by exactly two triangles
 * (an open mesh; vertices are welded by equal coordinates), an inconsistent
 * winding of a non-convex mesh, no enclosed volume, more than 65536
 * triangles in a non-convex mesh. */
typedef struct g4go_volume_mesh {
    uint32_t volume_id;
    uint32_t n_vertices;
    const double* vertices_mm;      /* [n_vertices][3] */
    const uint32_t* indices;        /* [n_triangles][3] */
    uint32_t n_triangles;
    uint32_t reserved;              /* 0 */
    const double* triangle_normals; /* [n_triangles][3] boundary normals (triangleNormals,
                                     * CreateTriangleNormals, OptiXTransportHost.cpp:1084-1116),
                                     * or NULL: the normal of each triangle's plane */
} g4go_volume_mesh;

typedef struct g4go_scene {
    const g4go_material* materials;
    uint32_t n_materials;
    const g4go_surface* surfaces;
    uint32_t n_surfaces;
    const g4go_volume* volumes;
    uint32_t n_volumes;
    const g4go_surface_binding* bindings;
    uint32_t n_bindings;
    uint32_t world_volume_id;
    const g4go_volume_mesh* meshes; /* at most one per volume; NULL when n_meshes is 0 */
    uint32_t n_meshes;
} g4go_scene;

/* OpticalEmission, 128 bytes, layout-identical to
 * include/g4go/optical/PhotonTransport.hpp:40-69 (type: 0 Direct,
 * 1 Cerenkov, 2 Scintillation). */
typedef struct g4go_emission {
    float position_mm[3];
    float time_ns;
    float direction[3];
    float step_length_mm;
    float step_delta_mm[3];
    float pre_velocity_mm_per_ns;
    float polarization[3];
    float delta_velocity_mm_per_ns;
    float energy_ev;
    float pre_mean_photon_count;
    float post_mean_photon_count;
    float decay_time_ns;
    float rise_time_ns;
    float charge;
    uint32_t event_id;
    uint32_t first_photon_id;
    uint32_t photon_count;
    uint32_t volume_id;
    uint32_t material_id;
    uint32_t spectrum_id;
    uint8_t type;
    uint8_t reserved[15];
} g4go_emission;

/* Synthetic source: positions uniform in [lo, hi], isotropic directions,
 * random linear polarisation, wavelength N(mean, sigma) nm, emission time
 * Exp(decay) ns; `volume_id` is the volume the photons start in. */
typedef struct g4go_synthetic_source {
    float lo_mm[3];
    float hi_mm[3];
    float wavelength_mean_nm;
    float wavelength_sigma_nm;
    float decay_time_ns;
    uint32_t volume_id;
} g4go_synthetic_source;

/* Detected photon (PhotonDetection, PhotonTransport.hpp:71-81, with the
 * arrival time in FP64). 48 bytes. */
typedef struct g4go_hit {
    float position_mm[3];
    float direction[3];
    float energy_ev;
    uint32_t sensor_id;
    uint32_t event_id;
    uint32_t photon_id;
    double time_ns;
} g4go_hit;

/* Per-event statistics (PhotonTransportEventStatistics, PhotonTransport.hpp:37-39). */
typedef struct g4go_event_stats {
    uint32_t event_id;
    uint32_t reserved;
    uint64_t captured, detected, absorbed, escaped, truncated, max_bounce, invalid, zero_step;
} g4go_event_stats;

#define G4GO_HIST_BINS 256
#define G4GO_HIST_SLOTS 258 /* bins 0..255 over [0, 200) ns, 256 underflow, 257 overflow */

typedef struct g4go_batch_stats {
    uint64_t captured, generated, detected, absorbed, escaped, truncated, max_bounce, invalid, zero_step;
    uint64_t total_steps;      /* transport steps (totalBounceCount) */
    uint64_t dropped_hits;     /* detections beyond the hit buffer (always 0 when sized by photons) */
    double transport_time_ms;  /* enqueue -> completion wall time */
    double kernel_ms;          /* transport kernel, HIP events */
    double host_to_device_ms;
    double device_to_host_ms;
    uint64_t host_to_device_bytes;
    uint64_t device_to_host_bytes;
    /* per-stage timers of PhotonTransportPerformance (OptiXTransportHost.cpp:845-876), HIP events */
    double device_memset_ms;         /* clearing the counters and event accumulators (0 when a graph replays them) */
    double device_order_ms;          /* launch-order pass over the hits (deviceHitCompactionMs) */
    double metadata_to_host_ms;      /* counters, histogram and event statistics to the host */
    uint64_t metadata_to_host_bytes;
    double hits_to_host_ms;          /* the hit records to the host (0 for g4go_complete_device) */
    uint64_t hits_to_host_bytes;
} g4go_batch_stats;

/* Results of one completed batch.  Arrays are owned by the handle and stay
 * valid until the next g4go_complete or g4go_destroy. */
typedef struct g4go_batch_result {
    const g4go_hit* hits;
    uint64_t n_hits;
    const g4go_event_stats* events;
    uint64_t n_events;
    const uint64_t* histogram; /* G4GO_HIST_SLOTS counts */
    g4go_batch_stats stats;
} g4go_batch_result;

/* ---- lifecycle (OptiXTransportHost ctor/dtor) ---- */
g4go_status g4go_create(const g4go_config* config, g4go_transport** out);
void g4go_destroy(g4go_transport* handle);
/* Message of the last failure on this handle (or of the last g4go_create
 * failure when handle is NULL). */
const char* g4go_last_error(const g4go_transport* handle);
/* Launch form of the handle's batches.  A new handle uses the persistent form
 * {persistent 1, block 256, chunk 256, refill 24, 8 blocks per CU, per_event
 * 0, graph 0, pregenerate 0}, the form tuned on MI355X for crystal batches. */
g4go_status g4go_set_launch(g4go_transport* handle, const g4go_launch* launch);
g4go_status g4go_get_launch(const g4go_transport* handle, g4go_launch* launch);

/* ---- PrepareScene (OptiXTransportHost.cpp:511-519) ---- */
g4go_status g4go_prepare_scene(g4go_transport* handle, const g4go_scene* scene);

/* ---- EnqueueEmissions (OptiXTransportHost.cpp:521-777) ----
 * emissions[n_emissions], event_ids[n_events], emission_event_indices[n_emissions]. */
g4go_status g4go_enqueue_emissions(g4go_transport* handle, const g4go_emission* emissions, uint64_t n_emissions,
                                   const uint32_t* event_ids, uint64_t n_events,
                                   const uint32_t* emission_event_indices);

/* Synthetic batch: event_ids[n_events] with photon_counts[n_events] photons
 * each, generated on the device from (seed, eventID, photonID). */
g4go_status g4go_enqueue_synthetic(g4go_transport* handle, const g4go_synthetic_source* source,
                                   const uint32_t* event_ids, const uint32_t* photon_counts, uint64_t n_events);

/* ---- device-memory interop (torch and other stream-ordered callers) ----
 * No counterpart in the reference's interface; the reference's caller hands
 * over host spans (OpticalBatchScheduler.cpp:521, 551).
 *
 * g4go_set_stream: every batch enqueued afterwards first waits for the work
 * queued on `stream` (a hipStream_t, e.g. torch.cuda.current_stream()) when
 * it was enqueued; NULL stops waiting.  The backend's own streams keep the
 * default priority.
 *
 * g4go_enqueue_emissions_device: g4go_enqueue_emissions with the records and
 * the emission event indices in device memory (read in order after the work
 * on the stream set by g4go_set_stream, else on the backend's stream); the
 * event IDs are a host array.  The records are validated as on the host
 * path. */
g4go_status g4go_set_stream(g4go_transport* handle, void* stream);
g4go_status g4go_enqueue_emissions_device(g4go_transport* handle, const g4go_emission* emissions_device,
                                          uint64_t n_emissions, const uint32_t* event_ids, uint64_t n_events,
                                          const uint32_t* emission_event_indices_device);

/* ---- CompleteOldestBatch (OptiXTransportHost.cpp:779-883) ---- */
g4go_status g4go_complete(g4go_transport* handle, g4go_batch_result* out);

/* Same as g4go_complete, but the detections stay in device memory: out->hits
 * is NULL, out->n_hits is their count and *hits_device points at n_hits
 * g4go_hit records in HBM (in launch order when config.sort_hits is set,
 * otherwise in completion order), valid until the next enqueue. */
g4go_status g4go_complete_device(g4go_transport* handle, g4go_batch_result* out, const void** hits_device);

/* ---- allocation-free synthetic batch ----
 * No counterpart in the reference: for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  The variant of g4go_enqueue_synthetic +
 * g4go_complete_device that allocates nothing: the batch description is in
 * device memory, the caller passes the output buffers, and the work queue,
 * batch header, item offsets and per-event accumulation copies live in the
 * library's fixed device scratch array (G4GO_NOALLOC_SCRATCH_BYTES per GPU,
 * held until the process ends).  At most G4GO_NOALLOC_MAX_EVENTS events per
 * call; larger batches are rejected with G4GO_ERROR_INVALID_ARGUMENT.
 *
 * The call only enqueues device work on `stream` (a hipStream_t; NULL: the
 * handle's own stream) and returns: no allocation, no host synchronisation,
 * no hipMemset (fill kernels clear what starts at zero), so it can be
 * captured into a graph.  The outputs are complete when the stream reaches
 * that point.  The handle's scene must already be resident on the device
 * (prepared and transported at least once by the ordinary entry points), and
 * must not change while such work is in flight.  Every call on one GPU shares
 * that GPU's scratch array: calls whose work overlaps on one GPU (on
 * different streams, or from different handles) race and give wrong results.
 * The transport is the handle's persistent form (block size, chunk, refill
 * threshold, blocks per CU of g4go_get_launch), with the handle's seed; the
 * hits are the same records as g4go_complete_device's, in completion order.
 * The caller guarantees distinct event IDs and fewer than 2^32 photons. */
#define G4GO_NOALLOC_MAX_EVENTS 1024
#define G4GO_NOALLOC_SCRATCH_BYTES 1844224 /* header, counters, offsets, 32 x 1024 event counters */

typedef struct g4go_device_outputs {
    g4go_hit* hits;              /* hit_capacity records; detections beyond it are counted as dropped */
    uint64_t hit_capacity;       /* at most 2^32 - 1 */
    uint32_t* hit_count;         /* 1: detections (may exceed hit_capacity) */
    uint64_t* histogram;         /* G4GO_HIST_SLOTS arrival-time counts */
    uint64_t* event_counters;    /* 7 per event, in batch order: detected, absorbed, escaped,
                                    truncated, max_bounce, invalid, zero_step */
    uint64_t* batch_counters;    /* 3: transport steps, generated photons, dropped hits */
} g4go_device_outputs;

/* event_ids_device[n_events], photon_counts_device[n_events]: device arrays.
 * All pointers of `outputs` are device pointers. */
g4go_status g4go_transport_synthetic_noalloc(g4go_transport* handle, const g4go_synthetic_source* source,
                                             const uint32_t* event_ids_device, const uint32_t* photon_counts_device,
                                             uint32_t n_events, const g4go_device_outputs* outputs, void* stream);

/* ---- differentiation (path-replay adjoint) ----
 * A parameter layout binds property tables of the prepared scene to a flat
 * parameter vector theta, table after table in the order given:
 *   kind 0: absorption length of material `index` (all its values),
 *   kind 1: Rayleigh length of material `index`,
 *   kind 2: reflectivity of surface `index`,
 *   kind 3: efficiency of surface `index`. */
typedef struct g4go_param_table {
    uint32_t kind;
    uint32_t index;
} g4go_param_table;

typedef struct g4go_adjoint_stats {
    uint64_t replayed;     /* detected photons replayed and scored */
    uint64_t mismatched;   /* replays that did not end in their recorded detection (0 expected) */
    uint64_t rescored;     /* of those, paths through more than two materials or two parameterised
                              surfaces, scored step by step on a second replay */
    uint64_t steps;        /* transport steps replayed */
    double kernel_ms;
} g4go_adjoint_stats;

/* Binds tables to theta; *n_params receives the length of theta. */
g4go_status g4go_set_parameter_layout(g4go_transport* handle, const g4go_param_table* tables, uint32_t n_tables,
                                      uint32_t* n_params);
/* Current values of theta from the scene. */
g4go_status g4go_get_parameters(g4go_transport* handle, double* theta, uint64_t n_params);
/* Writes theta into the bound tables (rounded to FP32) for the next batches. */
g4go_status g4go_set_parameters(g4go_transport* handle, const double* theta, uint64_t n_params);
/* Gradient of a loss of the last completed batch's detected count N and
 * histogram h: slot_weight[G4GO_HIST_SLOTS] = dLoss/dN + dLoss/dh_b per
 * arrival-time slot (bins 0..255, underflow, overflow).  Writes
 * grad[n_params] = dLoss/dtheta (likelihood-ratio estimator over the batch's
 * detected photons, replayed on the device from their Philox counters).
 * Must follow g4go_complete or g4go_complete_device of that batch. */
g4go_status g4go_adjoint(g4go_transport* handle, const double* slot_weight, double* grad, uint64_t n_params,
                         g4go_adjoint_stats* stats);

/* ---- test entry points (device evaluation of the shared physics) ----
 * which: 0 log(x), 1 exp(x), 2 sin/cos(2*pi*x) (two outputs per input). */
g4go_status g4go_eval_math(g4go_transport* handle, int which, const float* in, float* out, uint64_t n);
/* 11 inputs per case (dir[3], pol[3], normal[3], n1, n2) -> 14 outputs
 * (transmittance, tir, reflected dir[3], transmitted dir[3], reflected pol[3], transmitted pol[3]). */
g4go_status g4go_eval_fresnel(g4go_transport* handle, const float* in, float* out, uint64_t n);

/* Step-by-step trace of one synthetic photon (tests and diagnostics): up to
 * max_steps records of 16 floats (pos[3], dir[3], pol[3], time, status,
 * zero-step flag, step count, volume index, energy, 0). */
g4go_status g4go_trace_synthetic(g4go_transport* handle, const g4go_synthetic_source* source, uint32_t event_id,
                                 uint32_t photon_id, uint32_t max_steps, float* out, uint32_t* n_steps);

/* Name (gcnArchName and marketing name) and compute-unit count of the
 * handle's device; initialises HIP on the device. */
g4go_status g4go_device_info(g4go_transport* handle, char* name, uint64_t name_len, int32_t* cu_count);

/* Library version string. */
const char* g4go_version(void);

#ifdef __cplusplus
}
#endif

#endif /* G4GO_HIP_H */
