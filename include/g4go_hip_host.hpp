// g4go_hip_host.hpp -- C++ host shim over the C ABI in g4go_hip.h.
//
// BasicHipTransportHost keeps the three-method interface of the reference's
// GPU backend, G4GO::Optical::OptiXTransportHost
// (include/g4go/optical/optix/OptiXTransportHost.hpp:10-26):
//     explicit ctor(PhotonTransportConfig)
//     PrepareScene(const Scene&)
//     EnqueueEmissions(const Scene&, const PhotonTransportBatch&)
//     CompleteOldestBatch() -> PhotonTransportOutput
// and turns status codes back into the exception types that class throws
// (runtime_error, invalid_argument, overflow_error, logic_error;
// src/optical/optix/OptiXTransportHost.cpp:47-52, 574-630, 779-781).
//
// The shim is a template over a type bundle `Api` that names the caller's
// data types (Scene, PhotonTransportConfig, PhotonTransportBatch,
// PhotonTransportOutput, ...).  g4go_hip::api below restates those types with
// the reference's member names and layouts, so the shim builds and is tested
// without Geant4; inside the reference tree the bundle names the reference's
// own types instead.  The shim includes no HIP header: it reaches the device
// only through g4go_hip.h.
//
// Geometry: a volume's triangle mesh (Scene.hpp:94-105) is mapped to world
// space with its transform (Scene.hpp:48-59).  A mesh that is an axis-aligned
// box there (all eight corners of its bounding box present, every vertex on a
// corner) is passed as a box solid, the backend's fastest geometry; any other
// mesh is passed as a world-space triangle mesh (g4go_volume_mesh, triangle
// normals turned with it), which the backend accepts when it is closed and
// convex (PrepareScene throws std::invalid_argument otherwise).
#pragma once

#include "g4go_hip.h"

#include <algorithm>
#include <array>
#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstddef>
#include <cstdint>
#include <cstring>
#include <deque>
#include <limits>
#include <mutex>
#include <span>
#include <stdexcept>
#include <string>
#include <thread>
#include <type_traits>
#include <utility>
#include <vector>

namespace g4go_hip {

// ---------------------------------------------------------------------------
// Restatement of the reference's data types (include/g4go/optical/Scene.hpp,
// PhotonTransport.hpp, PhotonTransportConfig.hpp), member names kept.
namespace api {

constexpr std::uint32_t InvalidID = std::numeric_limits<std::uint32_t>::max();

struct ThreeVector {  // the members of G4ThreeVector the shim reads
    double px{}, py{}, pz{};
    double x() const { return px; }
    double y() const { return py; }
    double z() const { return pz; }
};

enum class SurfaceType : std::uint8_t { DielectricDielectric, DielectricMetal };
enum class SurfaceModel : std::uint8_t { Glisur, Unified };
enum class SurfaceFinish : std::uint8_t { Polished, Ground };

struct Rotation {
    float xx{1.0F}, xy{}, xz{};
    float yx{}, yy{1.0F}, yz{};
    float zx{}, zy{}, zz{1.0F};
};

struct Transform {
    Rotation rotation{};
    ThreeVector translationMm{};
};

struct PropertyTable {
    std::vector<float> energyEv{};
    std::vector<float> values{};
    bool Empty() const { return energyEv.empty(); }
};

struct Material {
    std::string name{};
    PropertyTable rindex{};
    float rindexMax{1.0F};
    PropertyTable groupVelocityMmPerNs{};
    PropertyTable absLengthMm{};
    std::array<PropertyTable, 3> scintillationSpectrum{};
    PropertyTable rayleighLengthMm{};  // not in the reference's Material; read when present
};

struct Surface {
    std::string name{};
    SurfaceType type{SurfaceType::DielectricDielectric};
    SurfaceModel model{SurfaceModel::Unified};
    SurfaceFinish finish{SurfaceFinish::Polished};
    float modelValue{1.0F};
    PropertyTable reflectivity{};
    PropertyTable efficiency{};
    PropertyTable transmittance{};
    PropertyTable rindex{};
    PropertyTable specularLobe{};
    PropertyTable specularSpike{};
    PropertyTable backscatter{};
    PropertyTable surfaceRoughness{};
};

struct MeshGeometry {
    std::string name{};
    std::vector<ThreeVector> verticesMm{};
    std::vector<std::uint32_t> indices{};
    std::vector<ThreeVector> triangleNormals{};
    std::vector<std::uint8_t> triangleFlags{};
};

struct Geometry {
    std::string name{};
    MeshGeometry mesh{};
};

struct Volume {
    std::string name{};
    std::uint32_t volumeID{InvalidID};
    std::uint32_t physicalVolumeID{InvalidID};
    std::uint32_t copyNo{};
    std::uint32_t geometryID{InvalidID};
    std::uint32_t materialID{InvalidID};
    std::uint32_t parentVolumeID{InvalidID};
    std::uint32_t skinSurfaceID{InvalidID};
    std::uint32_t sensorID{InvalidID};
    std::uint32_t depth{};
    bool mayHaveCoincidentBoundary{};
    Transform transform{};
};

struct SurfaceBinding {
    std::uint32_t fromVolumeID{InvalidID};
    std::uint32_t toVolumeID{InvalidID};
    std::uint32_t surfaceID{InvalidID};
};

// Scene with the reference's accessors (Scene.hpp:128-166); IDs are indices.
class Scene {
public:
    std::uint32_t AddMaterial(Material m) { return push(fMaterials, std::move(m)); }
    std::uint32_t AddSurface(Surface s) { return push(fSurfaces, std::move(s)); }
    std::uint32_t AddGeometry(Geometry g) { return push(fGeometries, std::move(g)); }
    std::uint32_t AddVolume(Volume v) {
        v.volumeID = static_cast<std::uint32_t>(fVolumes.size());
        return push(fVolumes, std::move(v));
    }
    void AddSurfaceBinding(SurfaceBinding b) { fSurfaceBindings.push_back(b); }
    const Geometry* FindGeometry(std::uint32_t id) const { return id < fGeometries.size() ? &fGeometries[id] : nullptr; }
    const std::vector<Material>& Materials() const { return fMaterials; }
    const std::vector<Surface>& Surfaces() const { return fSurfaces; }
    const std::vector<Geometry>& Geometries() const { return fGeometries; }
    const std::vector<Volume>& Volumes() const { return fVolumes; }
    const std::vector<SurfaceBinding>& SurfaceBindings() const { return fSurfaceBindings; }
    std::uint32_t WorldVolumeID() const { return fWorldVolumeID; }
    void WorldVolumeID(std::uint32_t id) { fWorldVolumeID = id; }

private:
    template <class T>
    static std::uint32_t push(std::vector<T>& v, T x) {
        v.push_back(std::move(x));
        return static_cast<std::uint32_t>(v.size() - 1);
    }
    std::vector<Material> fMaterials;
    std::vector<Surface> fSurfaces;
    std::vector<Geometry> fGeometries;
    std::vector<Volume> fVolumes;
    std::vector<SurfaceBinding> fSurfaceBindings;
    std::uint32_t fWorldVolumeID{InvalidID};
};

enum class OpticalEmissionType : std::uint8_t { Direct, Cerenkov, Scintillation };

enum class PhotonTransportStatisticField : std::uint8_t {
    Generated = 0, Captured, Detected, Absorbed, Escaped, Truncated, MaxBounce, InvalidState, ZeroStep, TransportTime,
};

constexpr std::uint32_t StatisticFieldBit(PhotonTransportStatisticField f) {
    return std::uint32_t{1} << static_cast<std::uint32_t>(f);
}

struct alignas(16) OpticalEmission {
    std::array<float, 3> positionMm{};
    float timeNs{};
    std::array<float, 3> direction{};
    float stepLengthMm{};
    std::array<float, 3> stepDeltaMm{};
    float preVelocityMmPerNs{};
    std::array<float, 3> polarization{};
    float deltaVelocityMmPerNs{};
    float energyEv{};
    float preMeanPhotonCount{};
    float postMeanPhotonCount{};
    float decayTimeNs{};
    float riseTimeNs{};
    float charge{};
    std::uint32_t eventID{};
    std::uint32_t firstPhotonID{};
    std::uint32_t photonCount{};
    std::uint32_t volumeID{InvalidID};
    std::uint32_t materialID{InvalidID};
    std::uint32_t spectrumID{};
    OpticalEmissionType type{OpticalEmissionType::Direct};
};

struct alignas(16) PhotonDetection {
    std::array<float, 3> positionMm{};
    float timeNs{};
    std::array<float, 3> direction{};
    float energyEv{};
    std::uint32_t eventID{};
    std::uint32_t photonID{};
    std::uint32_t sensorID{};
};

struct PhotonTransportStatistics {
    std::uint32_t validFields{};
    std::uint64_t generatedCount{};
    std::uint64_t capturedCount{};
    std::uint64_t detectedCount{};
    std::uint64_t absorbedCount{};
    std::uint64_t escapedCount{};
    std::uint64_t truncatedCount{};
    std::uint64_t maxBounceCount{};
    std::uint64_t invalidStateCount{};
    std::uint64_t zeroStepCount{};
    double transportTimeMs{};
};

struct PhotonTransportEventStatistics {
    std::uint32_t eventID{};
    PhotonTransportStatistics statistics{};
};

struct PhotonTransportPerformance {  // the members the backend fills (OptiXTransportHost.cpp:845-876)
    double hostToDeviceMs{};
    std::uint64_t hostToDeviceBytes{};
    double deviceMemsetMs{};
    double optiXKernelMs{};
    double deviceHitCompactionMs{};
    double deviceToHostMs{};
    double deviceMetadataToHostMs{};
    std::uint64_t deviceMetadataToHostBytes{};
    double deviceHitsToHostMs{};
    std::uint64_t deviceHitsToHostBytes{};
    double hostHitCompactionMs{};
    std::uint64_t totalBounceCount{};
    std::uint64_t coincidentCandidateTraceCount{};
    std::uint64_t coincidentCandidateHitCount{};
};

struct PhotonTransportOutput {
    std::vector<PhotonDetection> detections{};
    PhotonTransportStatistics statistics{};
    PhotonTransportPerformance performance{};
    std::vector<PhotonTransportEventStatistics> eventStatistics{};
};

struct PhotonTransportBatch {
    std::span<const OpticalEmission> emissions{};
    std::span<const std::uint32_t> eventIDs{};
    std::span<const std::uint32_t> emissionEventIndices{};
};

struct PhotonTransportConfig {
    std::uint64_t seed{};
    std::uint32_t maxPhotonsPerEvent{5'000'000};
    std::uint32_t maxBouncesPerPhoton{4096};
    float boundaryToleranceMm{1.0e-4F};
    std::uint32_t targetPhotonsPerBatch{1'000'000};
    std::uint32_t maxEventsPerBatch{10'000};
    bool enablePerformanceDiagnostics{};
};

// The type bundle BasicHipTransportHost is instantiated with.
struct Types {
    using Scene = api::Scene;
    using PhotonTransportConfig = api::PhotonTransportConfig;
    using PhotonTransportBatch = api::PhotonTransportBatch;
    using PhotonTransportOutput = api::PhotonTransportOutput;
    using PhotonDetection = api::PhotonDetection;
    using OpticalEmission = api::OpticalEmission;
    using PhotonTransportEventStatistics = api::PhotonTransportEventStatistics;
    using PhotonTransportStatisticField = api::PhotonTransportStatisticField;
};

static_assert(sizeof(OpticalEmission) == 128);
static_assert(sizeof(PhotonDetection) == 48);

}  // namespace api

// ---------------------------------------------------------------------------
// Status -> exception (the reference's exception types per failure class).
[[noreturn]] inline void throw_status(g4go_status status, const std::string& message) {
    switch (status) {
    case G4GO_ERROR_INVALID_ARGUMENT:
        throw std::invalid_argument(message);
    case G4GO_ERROR_OVERFLOW:
        throw std::overflow_error(message);
    case G4GO_ERROR_LOGIC:
        throw std::logic_error(message);
    default:
        throw std::runtime_error(message);
    }
}

inline void check(g4go_status status, const g4go_transport* handle) {
    if (status != G4GO_OK) {
        throw_status(status, g4go_last_error(handle));
    }
}

// ---------------------------------------------------------------------------
// Scene flattening: C-ABI arrays over a caller's Scene.  Property arrays are
// borrowed from the scene (g4go_prepare_scene copies them during the call).
template <class Scene>
class FlatScene {
public:
    FlatScene(const Scene& scene, float tolerance_mm) {
        for (const auto& m : scene.Materials()) {
            g4go_material out{};
            out.rindex = table(m.rindex, "rindex");
            out.group_velocity_mm_per_ns = table(m.groupVelocityMmPerNs, "group velocity");
            out.abs_length_mm = table(m.absLengthMm, "absorption length");
            if constexpr (requires { m.rayleighLengthMm; }) {
                out.rayleigh_length_mm = table(m.rayleighLengthMm, "Rayleigh length");
            }
            for (int k = 0; k < 3; ++k) {
                out.scintillation_spectrum[k] = table(m.scintillationSpectrum[k], "scintillation spectrum");
            }
            materials_.push_back(out);
        }
        for (const auto& s : scene.Surfaces()) {
            g4go_surface out{};
            out.type = static_cast<std::uint32_t>(s.type);
            out.model = static_cast<std::uint32_t>(s.model);
            out.finish = static_cast<std::uint32_t>(s.finish);
            out.model_value = s.modelValue;
            out.reflectivity = table(s.reflectivity, "reflectivity");
            out.efficiency = table(s.efficiency, "efficiency");
            out.transmittance = table(s.transmittance, "transmittance");
            out.specular_lobe = table(s.specularLobe, "specular lobe");
            out.specular_spike = table(s.specularSpike, "specular spike");
            out.backscatter = table(s.backscatter, "backscatter");
            out.surface_roughness = table(s.surfaceRoughness, "surface roughness");
            surfaces_.push_back(out);
        }
        for (const auto& v : scene.Volumes()) {
            g4go_volume out{};
            out.volume_id = v.volumeID;
            out.material_id = v.materialID;
            out.parent_volume_id = v.parentVolumeID;
            out.skin_surface_id = v.skinSurfaceID;
            out.sensor_id = v.sensorID;
            world_solid(scene, v, tolerance_mm, out.center_mm, out.half_mm);
            volumes_.push_back(out);
        }
        for (const auto& b : scene.SurfaceBindings()) {
            bindings_.push_back(g4go_surface_binding{b.fromVolumeID, b.toVolumeID, b.surfaceID});
        }
        for (const MeshStore& m : mesh_store_) {  // pointers taken once the stores stop growing
            g4go_volume_mesh out{};
            out.volume_id = m.volume_id;
            out.n_vertices = static_cast<std::uint32_t>(m.vertices.size() / 3);
            out.vertices_mm = m.vertices.data();
            out.indices = m.indices.data();
            out.n_triangles = static_cast<std::uint32_t>(m.indices.size() / 3);
            out.triangle_normals = m.normals.empty() ? nullptr : m.normals.data();
            meshes_.push_back(out);
        }
        view_.materials = materials_.data();
        view_.n_materials = static_cast<std::uint32_t>(materials_.size());
        view_.surfaces = surfaces_.data();
        view_.n_surfaces = static_cast<std::uint32_t>(surfaces_.size());
        view_.volumes = volumes_.data();
        view_.n_volumes = static_cast<std::uint32_t>(volumes_.size());
        view_.bindings = bindings_.data();
        view_.n_bindings = static_cast<std::uint32_t>(bindings_.size());
        view_.world_volume_id = scene.WorldVolumeID();
        view_.meshes = meshes_.empty() ? nullptr : meshes_.data();
        view_.n_meshes = static_cast<std::uint32_t>(meshes_.size());
    }
    FlatScene(const FlatScene&) = delete;  // the view points into this object
    FlatScene& operator=(const FlatScene&) = delete;

    const g4go_scene& view() const { return view_; }
    const std::vector<g4go_volume>& volumes() const { return volumes_; }
    const std::vector<g4go_volume_mesh>& meshes() const { return meshes_; }

private:
    template <class Table>
    static g4go_property table(const Table& t, const char* what) {
        if (t.energyEv.size() != t.values.size()) {
            throw std::invalid_argument(std::string(what) + " table: energy and value counts differ");
        }
        if (t.energyEv.size() > std::numeric_limits<std::uint32_t>::max()) {
            throw std::overflow_error(std::string(what) + " table too long");
        }
        return g4go_property{static_cast<std::uint32_t>(t.energyEv.size()), t.energyEv.data(), t.values.data()};
    }

    // A volume's mesh in world space: as a box (centre, half lengths) when it
    // is an axis-aligned box there, else as a mesh store.
    template <class Volume>
    void world_solid(const Scene& scene, const Volume& v, float tolerance_mm, float* center, float* half) {
        const auto* geometry = scene.FindGeometry(v.geometryID);
        const std::string who = "volume " + std::to_string(v.volumeID);
        if (geometry == nullptr || geometry->mesh.verticesMm.empty()) {
            throw std::invalid_argument(who + ": no mesh geometry");
        }
        const auto& r = v.transform.rotation;
        const auto& t = v.transform.translationMm;
        MeshStore m;
        m.volume_id = v.volumeID;
        m.vertices.reserve(3 * geometry->mesh.verticesMm.size());
        double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
        for (const auto& p : geometry->mesh.verticesMm) {
            const double x = p.x(), y = p.y(), z = p.z();
            const double w[3] = {r.xx * x + r.xy * y + r.xz * z + t.x(), r.yx * x + r.yy * y + r.yz * z + t.y(),
                                 r.zx * x + r.zy * y + r.zz * z + t.z()};
            for (int a = 0; a < 3; ++a) {
                lo[a] = std::fmin(lo[a], w[a]);
                hi[a] = std::fmax(hi[a], w[a]);
                m.vertices.push_back(w[a]);
            }
        }
        const double tol = std::max(static_cast<double>(tolerance_mm), 1e-6);
        unsigned corners = 0;
        bool box = true;
        for (size_t i = 0; box && i < m.vertices.size(); i += 3) {
            unsigned corner = 0;
            for (int a = 0; a < 3; ++a) {
                if (std::fabs(m.vertices[i + a] - lo[a]) <= tol) {
                    continue;
                }
                if (std::fabs(m.vertices[i + a] - hi[a]) <= tol) {
                    corner |= 1u << a;
                    continue;
                }
                box = false;
            }
            corners |= 1u << corner;
        }
        if (box && corners == 0xffu) {
            for (int a = 0; a < 3; ++a) {
                center[a] = static_cast<float>(0.5 * (lo[a] + hi[a]));
                half[a] = static_cast<float>(0.5 * (hi[a] - lo[a]));
            }
            return;
        }
        if (geometry->mesh.indices.size() % 3 != 0 ||
            geometry->mesh.indices.size() / 3 > std::numeric_limits<std::uint32_t>::max()) {
            throw std::invalid_argument(who + ": invalid mesh");
        }
        m.indices.assign(geometry->mesh.indices.begin(), geometry->mesh.indices.end());
        for (const auto& n : geometry->mesh.triangleNormals) {
            const double x = n.x(), y = n.y(), z = n.z();
            m.normals.insert(m.normals.end(), {r.xx * x + r.xy * y + r.xz * z, r.yx * x + r.yy * y + r.yz * z,
                                               r.zx * x + r.zy * y + r.zz * z});
        }
        if (!m.normals.empty() && m.normals.size() != m.indices.size()) {
            throw std::invalid_argument(who + ": triangle normal count mismatch");
        }
        mesh_store_.push_back(std::move(m));
    }

    struct MeshStore {
        std::uint32_t volume_id{};
        std::vector<double> vertices;  // world frame, xyz
        std::vector<std::uint32_t> indices;
        std::vector<double> normals;  // world frame, or empty
    };

    std::vector<g4go_material> materials_;
    std::vector<g4go_surface> surfaces_;
    std::vector<g4go_volume> volumes_;
    std::vector<g4go_surface_binding> bindings_;
    std::vector<MeshStore> mesh_store_;
    std::vector<g4go_volume_mesh> meshes_;
    g4go_scene view_{};
};

// ---------------------------------------------------------------------------
// A fixed set of helper threads that split an index range with the calling
// thread.  CompleteOldestBatch uses it to turn a batch's hit records into the
// caller's detection records: that loop streams the freshly copied hits from
// host memory and is bound by one core's memory bandwidth.
class RangePool {
public:
    explicit RangePool(unsigned helpers) {
        threads_.reserve(helpers);
        for (unsigned k = 0; k < helpers; ++k) {
            threads_.emplace_back([this, k] { work(k + 1); });
        }
    }

    ~RangePool() {
        {
            std::lock_guard<std::mutex> lock(mutex_);
            stop_.store(true);
        }
        wake_.notify_all();
        for (std::thread& t : threads_) {
            t.join();
        }
    }

    RangePool(const RangePool&) = delete;
    RangePool& operator=(const RangePool&) = delete;

    // From arm() to disarm() the helpers wait for a job by spinning (with
    // yields) instead of sleeping, so run() starts them without a futex
    // wake-up, which costs tens of microseconds; CompleteOldestBatch arms the
    // pool while it waits for the GPU.
    void arm() {
        {
            std::lock_guard<std::mutex> lock(mutex_);
            armed_.store(true);
        }
        wake_.notify_all();
    }
    void disarm() { armed_.store(false); }

    // Calls body(lo, hi) over disjoint slices covering [0, n), one slice per
    // thread, and returns when all are done.  body must not throw.
    template <class F>
    void run(std::uint64_t n, F& body) {
        if (threads_.empty() || n < kMinParallel) {
            body(std::uint64_t{0}, n);
            return;
        }
        job_ = Job{n, &body, [](void* f, std::uint64_t lo, std::uint64_t hi) { (*static_cast<F*>(f))(lo, hi); }};
        outstanding_.store(static_cast<unsigned>(threads_.size()), std::memory_order_relaxed);
        generation_.fetch_add(1, std::memory_order_release);
        {
            std::lock_guard<std::mutex> lock(mutex_);  // a sleeping helper re-checks under the lock
        }
        wake_.notify_all();
        const auto [lo, hi] = slice(n, 0);
        body(lo, hi);
        while (outstanding_.load(std::memory_order_acquire) != 0) {
            std::this_thread::yield();
        }
    }

private:
    static constexpr std::uint64_t kMinParallel = 16384;

    struct Job {
        std::uint64_t n = 0;
        void* body = nullptr;
        void (*call)(void*, std::uint64_t, std::uint64_t) = nullptr;
    };

    std::pair<std::uint64_t, std::uint64_t> slice(std::uint64_t n, unsigned part) const {
        const std::uint64_t parts = threads_.size() + 1;
        return {n * part / parts, n * (part + 1) / parts};
    }

    void work(unsigned part) {
        std::uint64_t seen = 0;
        for (;;) {
            // the next job: spin while armed, else sleep until one is posted
            for (;;) {
                if (stop_.load()) {
                    return;
                }
                const std::uint64_t g = generation_.load(std::memory_order_acquire);
                if (g != seen) {
                    seen = g;
                    break;
                }
                if (armed_.load(std::memory_order_relaxed)) {
                    std::this_thread::yield();
                    continue;
                }
                std::unique_lock<std::mutex> lock(mutex_);
                wake_.wait(lock, [&] {
                    return stop_.load() || armed_.load() || generation_.load(std::memory_order_acquire) != seen;
                });
            }
            const Job job = job_;
            const auto [lo, hi] = slice(job.n, part);
            job.call(job.body, lo, hi);
            outstanding_.fetch_sub(1, std::memory_order_release);
        }
    }

    std::vector<std::thread> threads_;
    std::mutex mutex_;
    std::condition_variable wake_;
    Job job_{};
    std::atomic<std::uint64_t> generation_{0};
    std::atomic<unsigned> outstanding_{0};
    std::atomic<bool> stop_{false}, armed_{false};
};

// Sized output vectors made ahead of need on a thread of their own: the
// zero-filling construction of n detection records (the value-initialisation
// of std::vector::resize, 0.17-0.21 ms for 1.5e5 records on the test host)
// runs while the GPU transports the next batch instead of after its hits have
// arrived.  request(n) starts one; take() returns it (waiting if it is still
// being made) and leaves none prepared.
template <class T>
class SpareVector {
public:
    SpareVector() : thread_([this] { work(); }) {}

    ~SpareVector() {
        {
            std::lock_guard<std::mutex> lock(mutex_);
            stop_ = true;
        }
        wake_.notify_all();
        thread_.join();
    }

    SpareVector(const SpareVector&) = delete;
    SpareVector& operator=(const SpareVector&) = delete;

    void request(std::size_t n) {
        {
            std::lock_guard<std::mutex> lock(mutex_);
            if (requested_ || ready_) {
                return;
            }
            requested_ = true;
            size_ = n;
        }
        wake_.notify_all();
    }

    // The prepared vector (empty when none was requested).
    std::vector<T> take() {
        std::unique_lock<std::mutex> lock(mutex_);
        done_.wait(lock, [this] { return !requested_; });
        ready_ = false;
        return std::move(vector_);
    }

private:
    void work() {
        for (;;) {
            std::size_t n = 0;
            {
                std::unique_lock<std::mutex> lock(mutex_);
                wake_.wait(lock, [this] { return stop_ || requested_; });
                if (stop_) {
                    return;
                }
                n = size_;
            }
            std::vector<T> v;
            try {
                v.resize(n);
            } catch (...) {
                v = std::vector<T>();  // out of memory: the caller sizes its own
            }
            {
                std::lock_guard<std::mutex> lock(mutex_);
                vector_ = std::move(v);
                requested_ = false;
                ready_ = true;
            }
            done_.notify_all();
        }
    }

    std::mutex mutex_;
    std::condition_variable wake_, done_;
    bool stop_ = false, requested_ = false, ready_ = false;
    std::size_t size_ = 0;
    std::vector<T> vector_;
    std::thread thread_;
};

// ---------------------------------------------------------------------------
template <class Api = api::Types>
class BasicHipTransportHost final {
public:
    using Scene = typename Api::Scene;
    using Config = typename Api::PhotonTransportConfig;
    using Batch = typename Api::PhotonTransportBatch;
    using Output = typename Api::PhotonTransportOutput;
    using Detection = typename Api::PhotonDetection;
    using Emission = typename Api::OpticalEmission;
    using EventStatistics = typename Api::PhotonTransportEventStatistics;
    using Field = typename Api::PhotonTransportStatisticField;

    static_assert(sizeof(Emission) == sizeof(g4go_emission) && std::is_trivially_copyable_v<Emission>,
                  "OpticalEmission must keep the 128-byte record layout");
    static_assert(offsetof(Emission, energyEv) == offsetof(g4go_emission, energy_ev) &&
                      offsetof(Emission, eventID) == offsetof(g4go_emission, event_id) &&
                      offsetof(Emission, photonCount) == offsetof(g4go_emission, photon_count) &&
                      offsetof(Emission, type) == offsetof(g4go_emission, type),
                  "OpticalEmission field offsets differ from g4go_emission");

    // OptiXTransportHost(PhotonTransportConfig) (OptiXTransportHost.hpp:12).
    // in_flight > 1 lets that many batches be enqueued before the oldest is
    // completed (the reference keeps one, OptiXTransportHost.cpp:523-525).
    // conversion_threads is the number of host threads, the caller's
    // included, that turn a completed batch's hits into detection records
    // (1: the calling thread alone).
    explicit BasicHipTransportHost(Config configuration, const g4go_launch* launch = nullptr,
                                   std::uint32_t in_flight = 1, std::uint32_t conversion_threads = 8)
        : tolerance_mm_(configuration.boundaryToleranceMm),
          pool_(conversion_threads > 1 ? conversion_threads - 1 : 0) {
        g4go_config c{};
        c.seed = configuration.seed;
        c.max_photons_per_event = configuration.maxPhotonsPerEvent;
        c.max_bounces_per_photon = configuration.maxBouncesPerPhoton;
        c.boundary_tolerance_mm = configuration.boundaryToleranceMm;
        c.target_photons_per_batch = configuration.targetPhotonsPerBatch;
        c.max_events_per_batch = configuration.maxEventsPerBatch;
        c.enable_performance_diagnostics = configuration.enablePerformanceDiagnostics ? 1u : 0u;
        c.device = -1;
        c.sort_hits = 1;  // detections in launch order, as the reference compacts them
        c.max_in_flight = in_flight;
        check(g4go_create(&c, &handle_), nullptr);
        if (launch != nullptr) {
            const g4go_status s = g4go_set_launch(handle_, launch);
            if (s != G4GO_OK) {
                const std::string message = g4go_last_error(handle_);
                g4go_destroy(handle_);
                throw_status(s, message);
            }
        }
    }

    ~BasicHipTransportHost() { g4go_destroy(handle_); }
    BasicHipTransportHost(const BasicHipTransportHost&) = delete;
    BasicHipTransportHost& operator=(const BasicHipTransportHost&) = delete;

    // PrepareScene (OptiXTransportHost.cpp:511-519): flattens and uploads the
    // scene; the scene's address is remembered so that EnqueueEmissions with
    // another scene object prepares that one first.
    void PrepareScene(const Scene& scene) {
        FlatScene<Scene> flat(scene, tolerance_mm_);
        check(g4go_prepare_scene(handle_, &flat.view()), handle_);
        prepared_ = &scene;
    }

    // EnqueueEmissions (OptiXTransportHost.cpp:521-777).  The batch's spans are
    // borrowed for the call only.
    void EnqueueEmissions(const Scene& scene, const Batch& batch) {
        if (prepared_ != &scene) {
            PrepareScene(scene);
        }
        if (batch.emissionEventIndices.size() != batch.emissions.size()) {
            throw std::invalid_argument("emission event-index count does not match emissions");
        }
        check(g4go_enqueue_emissions(handle_, reinterpret_cast<const g4go_emission*>(batch.emissions.data()),
                                     batch.emissions.size(), batch.eventIDs.data(), batch.eventIDs.size(),
                                     batch.emissionEventIndices.data()),
              handle_);
        // which events own an emission record, and whether the batch has photons:
        // they decide the valid-field masks of the output (OptiXTransportHost.cpp:587-632)
        Pending pending;
        pending.has_emission.assign(batch.eventIDs.size(), false);
        for (std::uint32_t i : batch.emissionEventIndices) {
            pending.has_emission[i] = true;
        }
        for (const auto& e : batch.emissions) {
            pending.photons += e.photonCount;
        }
        pending_.push_back(std::move(pending));
    }

    // CompleteOldestBatch (OptiXTransportHost.cpp:779-883).
    Output CompleteOldestBatch() {
        g4go_batch_result r{};
        // the conversion helpers spin while the GPU finishes the batch
        pool_.arm();
        const g4go_status status = g4go_complete(handle_, &r);
        // g4go_complete retires the oldest batch even when it fails (only an
        // empty queue, where nothing is pending here either, leaves it), so
        // its bookkeeping goes before the status can throw
        Pending batch;
        if (!pending_.empty()) {
            batch = std::move(pending_.front());
            pending_.pop_front();
        }
        if (status != G4GO_OK) {
            pool_.disarm();
        }
        check(status, handle_);
        Output out{};
        const std::uint32_t counts = bit(Field::Detected) | bit(Field::Absorbed) | bit(Field::Escaped) |
                                     bit(Field::Truncated) | bit(Field::MaxBounce) | bit(Field::InvalidState) |
                                     bit(Field::ZeroStep);
        const auto conversion0 = std::chrono::steady_clock::now();
        // the vector prepared while this batch ran, trimmed (no work for
        // trivially destructible records) or grown to the hit count
        out.detections = spare_.take();
        out.detections.resize(r.n_hits);
        Detection* detections = out.detections.data();
        const g4go_hit* hits = r.hits;
        auto convert = [detections, hits](std::uint64_t lo, std::uint64_t hi) {
            for (std::uint64_t i = lo; i < hi; ++i) {
                const g4go_hit& h = hits[i];
                Detection& d = detections[i];
                d.positionMm = {h.position_mm[0], h.position_mm[1], h.position_mm[2]};
                d.timeNs = static_cast<float>(h.time_ns);
                d.direction = {h.direction[0], h.direction[1], h.direction[2]};
                d.energyEv = h.energy_ev;
                d.eventID = h.event_id;
                d.photonID = h.photon_id;
                d.sensorID = h.sensor_id;
            }
        };
        pool_.run(r.n_hits, convert);
        pool_.disarm();
        const double conversion_ms =
            std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - conversion0).count();
        auto& st = out.statistics;
        st.capturedCount = r.stats.captured;
        st.detectedCount = r.stats.detected;
        st.absorbedCount = r.stats.absorbed;
        st.escapedCount = r.stats.escaped;
        st.truncatedCount = r.stats.truncated;
        st.maxBounceCount = r.stats.max_bounce;
        st.invalidStateCount = r.stats.invalid;
        st.zeroStepCount = r.stats.zero_step;
        st.transportTimeMs = r.stats.transport_time_ms;
        // an empty batch reports no batch fields; a batch with photons reports
        // captured, the terminal counts and the transport time
        st.validFields = batch.photons != 0 ? bit(Field::Captured) | counts | bit(Field::TransportTime) : 0u;
        out.eventStatistics.reserve(r.n_events);
        for (std::uint64_t i = 0; i < r.n_events; ++i) {
            const g4go_event_stats& e = r.events[i];
            EventStatistics es{};
            es.eventID = e.event_id;
            es.statistics.capturedCount = e.captured;
            es.statistics.detectedCount = e.detected;
            es.statistics.absorbedCount = e.absorbed;
            es.statistics.escapedCount = e.escaped;
            es.statistics.truncatedCount = e.truncated;
            es.statistics.maxBounceCount = e.max_bounce;
            es.statistics.invalidStateCount = e.invalid;
            es.statistics.zeroStepCount = e.zero_step;
            const bool owns = i < batch.has_emission.size() && batch.has_emission[i];
            // captured when the event owns a record; terminal counts whenever the batch has records
            es.statistics.validFields = (owns ? bit(Field::Captured) : 0u) | (batch.any() ? counts : 0u);
            out.eventStatistics.push_back(es);
        }
        // the reference's stage timers; this backend has no coincident-face re-trace, so its
        // coincidentCandidate counts stay 0
        auto& perf = out.performance;
        perf.hostToDeviceMs = r.stats.host_to_device_ms;
        perf.hostToDeviceBytes = r.stats.host_to_device_bytes;
        perf.deviceMemsetMs = r.stats.device_memset_ms;
        perf.optiXKernelMs = r.stats.kernel_ms;
        perf.deviceHitCompactionMs = r.stats.device_order_ms;
        perf.deviceMetadataToHostMs = r.stats.metadata_to_host_ms;
        perf.deviceMetadataToHostBytes = r.stats.metadata_to_host_bytes;
        perf.deviceHitsToHostMs = r.stats.hits_to_host_ms;
        perf.deviceHitsToHostBytes = r.stats.hits_to_host_bytes;
        perf.deviceToHostMs = perf.deviceMetadataToHostMs + perf.deviceHitsToHostMs;
        perf.hostHitCompactionMs = conversion_ms;
        perf.totalBounceCount = r.stats.total_steps;
        // the next batch's records, sized by this one's hit count with a margin
        spare_.request(r.n_hits + r.n_hits / 8 + 1024);
        return out;
    }

    g4go_transport* handle() const { return handle_; }

private:
    static constexpr std::uint32_t bit(Field f) { return std::uint32_t{1} << static_cast<std::uint32_t>(f); }

    // what the output masks need of each batch in flight, in enqueue order
    struct Pending {
        std::vector<bool> has_emission;  // per event: owns an emission record
        std::uint64_t photons = 0;
        bool any() const {
            for (bool b : has_emission) {
                if (b) {
                    return true;
                }
            }
            return false;
        }
    };

    g4go_transport* handle_ = nullptr;
    const Scene* prepared_ = nullptr;
    float tolerance_mm_;
    std::deque<Pending> pending_;
    RangePool pool_;
    SpareVector<Detection> spare_;
};

using HipTransportHost = BasicHipTransportHost<>;

}  // namespace g4go_hip
