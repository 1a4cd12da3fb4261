// Tests of the C++ host shim (include/g4go_hip_host.hpp) through the C ABI.
//
//   shim_test cpu            host-side checks: mesh -> box flattening, the
//                            exception type of every failure class, an
//                            empty batch through all three methods
//   shim_test gpu OUT.bin    transports a crystal batch of scintillation,
//                            Cerenkov and direct emission records through
//                            the three-method interface and writes the
//                            records and the output to OUT.bin (layout below)
//   shim_test bench SLOTS BATCHES [THREADS [BLOCK CHUNK BLOCKS_PER_CU REFILL GRAPH PREGENERATE]]
//                            throughput of the drop-in path (JSON line)
//   shim_test gpu-mesh OUT.bin  the same with the crystal a tapered mesh
//                            (3x3 mm front face, 4x4 mm back face, turned
//                            into place by its transform), passed on as a
//                            world-space mesh; OUT.bin then ends with that
//                            mesh: u64 n_vertices, n_triangles; vertices
//                            (3 f64 each); indices (3 u32 each)
//   shim_test gpu-open OUT.bin  the same with one side face of the tapered
//                            mesh left out (an open solid)
//
// OUT.bin: u64 n_emissions, n_events, n_detections; emissions (128 B each);
// event IDs (u32); emission event indices (u32); detections (48 B each);
// per event: u32 eventID, u32 validFields, u64 captured, detected, absorbed,
// escaped, truncated, maxBounce, invalidState, zeroStep; batch: u32
// validFields, u32 0, u64 captured, detected, absorbed, escaped, truncated,
// maxBounce, invalidState, zeroStep, totalBounceCount.
#include "g4go_hip_host.hpp"

#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <functional>
#include <string>
#include <typeinfo>

namespace {

using namespace g4go_hip::api;

int failures = 0;

void expect(bool ok, const std::string& what) {
    if (!ok) {
        std::printf("FAIL: %s\n", what.c_str());
        ++failures;
    }
}

template <class E>
void expect_throws(const std::function<void()>& f, const std::string& what) {
    try {
        f();
    } catch (const E&) {
        return;
    } catch (const std::exception& e) {
        std::printf("FAIL: %s threw %s (%s), expected %s\n", what.c_str(), typeid(e).name(), e.what(),
                    typeid(E).name());
        ++failures;
        return;
    }
    std::printf("FAIL: %s did not throw\n", what.c_str());
    ++failures;
}

// Triangle mesh of the box [-h, h] (12 triangles).
Geometry box_mesh(double hx, double hy, double hz) {
    Geometry g;
    for (int k = 0; k < 8; ++k) {
        g.mesh.verticesMm.push_back(ThreeVector{(k & 1) ? hx : -hx, (k & 2) ? hy : -hy, (k & 4) ? hz : -hz});
    }
    const std::uint32_t faces[6][4] = {{0, 2, 6, 4}, {1, 5, 7, 3}, {0, 4, 5, 1}, {2, 3, 7, 6}, {0, 1, 3, 2}, {4, 6, 7, 5}};
    for (const auto& f : faces) {
        for (std::uint32_t i : {f[0], f[1], f[2], f[0], f[2], f[3]}) {
            g.mesh.indices.push_back(i);
        }
    }
    return g;
}

// Frustum along local x from the 2*hy x 2*hz face at x = -hx to the
// 2*by x 2*bz face at x = +hx (12 triangles, wound outwards).
Geometry frustum_x_mesh(double hx, double hy, double hz, double by, double bz) {
    Geometry g = box_mesh(hx, hy, hz);
    for (auto& v : g.mesh.verticesMm) {
        if (v.px > 0) {
            v.py = v.py > 0 ? by : -by;
            v.pz = v.pz > 0 ? bz : -bz;
        }
    }
    return g;
}

PropertyTable table(std::initializer_list<float> e, std::initializer_list<float> v) {
    return PropertyTable{std::vector<float>(e), std::vector<float>(v)};
}

// Crystal 3x3x20 mm (a 20x3x3 mesh turned by 90 degrees about y) in an air
// world, wrapped (R = 0.95) on its world border, with a 3x3x0.1 mm SiPM on +z.
struct CrystalScene {
    Scene scene;
    std::uint32_t world, crystal, sipm;
};

CrystalScene crystal_scene(bool tapered = false, bool open = false) {
    CrystalScene c;
    Scene& s = c.scene;
    Material air;
    air.name = "air";
    air.rindex = table({1.5F, 4.5F}, {1.0F, 1.0F});
    Material crystal;
    crystal.name = "crystal";
    crystal.rindex = table({1.5F, 4.5F}, {1.82F, 1.82F});
    crystal.rindexMax = 1.82F;
    crystal.absLengthMm = table({1.5F, 4.5F}, {400.0F, 400.0F});
    crystal.rayleighLengthMm = table({1.5F, 4.5F}, {500.0F, 500.0F});
    crystal.scintillationSpectrum[0] = table({2.6F, 2.95F, 3.3F}, {0.0F, 1.0F, 0.0F});
    Material silicon;
    silicon.name = "silicon";
    silicon.rindex = table({1.5F, 4.5F}, {3.5F, 3.5F});
    const std::uint32_t m_air = s.AddMaterial(air), m_crystal = s.AddMaterial(crystal);
    const std::uint32_t m_si = s.AddMaterial(silicon);
    Surface wrap;
    wrap.name = "Reflector";
    wrap.type = SurfaceType::DielectricMetal;
    wrap.reflectivity = table({2.9F}, {0.95F});
    Surface cathode;
    cathode.name = "Cathode";
    cathode.type = SurfaceType::DielectricMetal;
    cathode.reflectivity = table({2.9F}, {0.0F});
    cathode.efficiency = table({2.5F, 3.5F}, {0.3F, 0.5F});
    const std::uint32_t s_wrap = s.AddSurface(wrap), s_cathode = s.AddSurface(cathode);
    const std::uint32_t g_world = s.AddGeometry(box_mesh(100, 100, 100));
    Geometry crystal_geometry = tapered ? frustum_x_mesh(10, 1.5, 1.5, 2.0, 2.0) : box_mesh(10, 1.5, 1.5);
    if (open) {
        crystal_geometry.mesh.indices.resize(30);  // one side face left out: an open solid
    }
    const std::uint32_t g_crystal = s.AddGeometry(crystal_geometry);
    const std::uint32_t g_sipm = s.AddGeometry(tapered ? box_mesh(2.0, 2.0, 0.05) : box_mesh(1.5, 1.5, 0.05));
    Volume world;
    world.name = "World";
    world.geometryID = g_world;
    world.materialID = m_air;
    c.world = s.AddVolume(world);
    s.WorldVolumeID(c.world);
    Volume xtal;
    xtal.name = "Crystal";
    xtal.geometryID = g_crystal;
    xtal.materialID = m_crystal;
    xtal.parentVolumeID = c.world;
    xtal.transform.rotation = Rotation{0, 0, -1, 0, 1, 0, 1, 0, 0};  // local x -> world z
    c.crystal = s.AddVolume(xtal);
    Volume sipm;
    sipm.name = "SiPM";
    sipm.geometryID = g_sipm;
    sipm.materialID = m_si;
    sipm.parentVolumeID = c.world;
    sipm.skinSurfaceID = s_cathode;
    sipm.sensorID = 0;
    sipm.transform.translationMm = ThreeVector{0, 0, 10.05};
    c.sipm = s.AddVolume(sipm);
    s.AddSurfaceBinding(SurfaceBinding{c.crystal, c.world, s_wrap});
    return c;
}

std::uint32_t lcg_state = 12345u;
float lcg() {  // uniform [0, 1)
    lcg_state = lcg_state * 1664525u + 1013904223u;
    return static_cast<float>(lcg_state >> 8) * (1.0F / 16777216.0F);
}

void make_batch(const CrystalScene& c, std::vector<OpticalEmission>& em, std::vector<std::uint32_t>& events,
                std::vector<std::uint32_t>& index) {
    for (std::uint32_t e = 0; e < 6; ++e) {
        const std::uint32_t event_id = 100 + 7 * e;
        events.push_back(event_id);
        std::uint32_t next_photon = 0;
        for (int k = 0; k < 9; ++k) {
            OpticalEmission r{};
            r.positionMm = {2.0F * lcg() - 1.0F, 2.0F * lcg() - 1.0F, -9.0F + 1.5F * k};
            r.timeNs = 0.1F * k;
            r.direction = {0.1F, -0.05F, 1.0F};
            r.stepLengthMm = 1.2F;
            r.stepDeltaMm = {0.05F, 0.02F, 1.2F};
            r.preVelocityMmPerNs = 280.0F;
            r.deltaVelocityMmPerNs = -20.0F;
            r.polarization = {1.0F, 0.0F, -0.1F};
            r.energyEv = 2.9F;
            r.eventID = event_id;
            r.firstPhotonID = next_photon;
            r.volumeID = c.crystal;
            r.materialID = 1;
            if (k % 3 == 2) {
                r.type = OpticalEmissionType::Cerenkov;
                r.preMeanPhotonCount = 12.0F;
                r.postMeanPhotonCount = 6.0F;
                r.charge = -1.0F;
                r.photonCount = 300;
            } else if (k % 3 == 1) {
                r.type = OpticalEmissionType::Direct;
                r.photonCount = 40;
            } else {
                r.type = OpticalEmissionType::Scintillation;
                r.decayTimeNs = 40.0F;
                r.riseTimeNs = k == 3 ? 0.5F : 0.0F;
                r.charge = -1.0F;
                r.photonCount = 1500 + 100 * e;
            }
            next_photon += r.photonCount;
            em.push_back(r);
            index.push_back(e);
        }
    }
}

int run_cpu() {
    // mesh -> world box
    {
        CrystalScene c = crystal_scene();
        g4go_hip::FlatScene<Scene> flat(c.scene, 1e-4F);
        const auto& v = flat.volumes();
        expect(v.size() == 3, "three volumes");
        expect(v[1].half_mm[0] == 1.5F && v[1].half_mm[1] == 1.5F && v[1].half_mm[2] == 10.0F,
               "rotated crystal mesh gives half lengths 1.5 x 1.5 x 10");
        expect(v[1].center_mm[0] == 0.0F && v[1].center_mm[2] == 0.0F, "crystal centred");
        expect(std::fabs(v[2].center_mm[2] - 10.05F) < 1e-6F && v[2].half_mm[2] == 0.05F, "SiPM translated to z=10.05");
        expect(flat.view().world_volume_id == c.world && flat.view().n_bindings == 1, "world and binding");
        expect(flat.view().materials[1].rayleigh_length_mm.count == 2, "Rayleigh table passed through");
    }
    // other meshes pass as world-space meshes; the backend checks them at PrepareScene
    {
        CrystalScene c = crystal_scene();
        Geometry tetra;
        tetra.mesh.verticesMm = {ThreeVector{30, 0, 0}, ThreeVector{31, 0, 0}, ThreeVector{30, 1, 0},
                                 ThreeVector{30, 0, 1}};
        tetra.mesh.indices = {0, 1, 2, 0, 1, 3, 1, 2, 3, 0, 2, 3};
        const std::uint32_t g = c.scene.AddGeometry(tetra);
        Volume v;
        v.geometryID = g;
        v.materialID = 0;
        v.parentVolumeID = c.world;
        v.transform.translationMm = ThreeVector{0, 5, 0};
        const std::uint32_t tv = c.scene.AddVolume(v);
        {
            g4go_hip::FlatScene<Scene> f(c.scene, 1e-4F);
            expect(f.meshes().size() == 1 && f.meshes()[0].volume_id == tv && f.meshes()[0].n_triangles == 4,
                   "tetrahedron passed as a mesh");
            expect(f.meshes()[0].vertices_mm[1] == 5.0 && f.view().n_meshes == 1, "mesh translated to world space");
        }
        g4go_hip::HipTransportHost host{PhotonTransportConfig{}};
        host.PrepareScene(c.scene);
        CrystalScene d = crystal_scene();
        Volume turned;
        turned.geometryID = 2;
        turned.materialID = 0;
        turned.parentVolumeID = d.world;
        turned.transform.translationMm = ThreeVector{0, 40, 0};
        const float s = 0.70710678F;
        turned.transform.rotation = Rotation{s, -s, 0, s, s, 0, 0, 0, 1};  // 45 degrees about z
        d.scene.AddVolume(turned);
        {
            g4go_hip::FlatScene<Scene> f(d.scene, 1e-4F);
            expect(f.meshes().size() == 1 && f.meshes()[0].n_triangles == 12, "box turned 45 deg passed as a mesh");
        }
        host.PrepareScene(d.scene);
        Geometry open = tetra;
        open.mesh.indices.resize(9);  // three triangles: not closed, an open solid
        CrystalScene e = crystal_scene();
        Volume ov = v;
        ov.geometryID = e.scene.AddGeometry(open);
        e.scene.AddVolume(ov);
        host.PrepareScene(e.scene);
        Geometry degenerate = tetra;
        degenerate.mesh.indices[2] = degenerate.mesh.indices[1];  // a triangle with a repeated corner
        CrystalScene dg = crystal_scene();
        Volume dv = v;
        dv.geometryID = dg.scene.AddGeometry(degenerate);
        dg.scene.AddVolume(dv);
        expect_throws<std::invalid_argument>([&] { host.PrepareScene(dg.scene); }, "degenerate mesh rejected");
    }
    // exception types
    {
        PhotonTransportConfig bad;
        bad.boundaryToleranceMm = 0.0F;
        expect_throws<std::invalid_argument>([&] { g4go_hip::HipTransportHost h(bad); }, "zero boundary tolerance");
        g4go_hip::HipTransportHost host{PhotonTransportConfig{}};
        expect_throws<std::logic_error>([&] { host.CompleteOldestBatch(); }, "complete with an empty queue");
        CrystalScene c = crystal_scene();
        Scene broken = c.scene;
        Material m;
        m.rindex = PropertyTable{{1.0F, 2.0F}, {1.5F}};
        broken.AddMaterial(m);
        expect_throws<std::invalid_argument>([&] { host.PrepareScene(broken); }, "table size mismatch");
        Scene unknown = c.scene;
        Volume v;
        v.geometryID = 1;
        v.materialID = 77;
        v.parentVolumeID = c.world;
        unknown.AddVolume(v);
        expect_throws<std::invalid_argument>([&] { host.PrepareScene(unknown); }, "unknown material");
        host.PrepareScene(c.scene);
        std::vector<OpticalEmission> em(1);
        em[0].eventID = 5;
        em[0].volumeID = c.crystal;
        std::vector<std::uint32_t> ev = {5, 5}, idx = {0};
        expect_throws<std::invalid_argument>(
            [&] { host.EnqueueEmissions(c.scene, PhotonTransportBatch{em, ev, idx}); }, "duplicate event IDs");
        ev = {6};
        expect_throws<std::invalid_argument>(
            [&] { host.EnqueueEmissions(c.scene, PhotonTransportBatch{em, ev, idx}); }, "event index mismatch");
        std::vector<std::uint32_t> no_idx;
        expect_throws<std::invalid_argument>(
            [&] { host.EnqueueEmissions(c.scene, PhotonTransportBatch{em, ev, no_idx}); }, "index count mismatch");
        // more than 2^32 - 1 photons in a batch: overflow_error (OptiXTransportHost.cpp:621-623),
        // raised by the host-side validation before anything reaches the device
        {
            std::vector<OpticalEmission> big(2);
            for (auto& r : big) {
                r.eventID = 6;
                r.volumeID = c.crystal;
                r.materialID = 1;
                r.type = OpticalEmissionType::Direct;
                r.photonCount = 0x80000000u;
            }
            std::vector<std::uint32_t> big_ev = {6}, big_idx = {0, 0};
            expect_throws<std::overflow_error>(
                [&] { host.EnqueueEmissions(c.scene, PhotonTransportBatch{big, big_ev, big_idx}); },
                "photon count beyond 2^32 - 1");
            // an unknown material is rejected on every record type, Direct included (:658-660)
            big.resize(1);
            big[0].photonCount = 10;
            big[0].materialID = 99;
            big_idx.resize(1);
            expect_throws<std::invalid_argument>(
                [&] { host.EnqueueEmissions(c.scene, PhotonTransportBatch{big, big_ev, big_idx}); },
                "direct record with an unknown material");
            big[0].materialID = 1;
            big[0].volumeID = 999;
            expect_throws<std::invalid_argument>(
                [&] { host.EnqueueEmissions(c.scene, PhotonTransportBatch{big, big_ev, big_idx}); },
                "record with an unknown volume");
        }
        // empty batch (zero photons) completes without touching the device
        em[0].eventID = 6;
        host.EnqueueEmissions(c.scene, PhotonTransportBatch{em, ev, idx});
        expect_throws<std::logic_error>(
            [&] { host.EnqueueEmissions(c.scene, PhotonTransportBatch{em, ev, idx}); }, "second batch in flight");
        PhotonTransportOutput out = host.CompleteOldestBatch();
        expect(out.detections.empty() && out.eventStatistics.size() == 1 && out.eventStatistics[0].eventID == 6,
               "empty batch output");
        expect(out.statistics.capturedCount == 0, "nothing captured");
        // a batch whose records hold no photons: no batch fields, per event captured
        // and the terminal counts (OptiXTransportHost.cpp:587-620)
        using F = PhotonTransportStatisticField;
        const std::uint32_t terminal = StatisticFieldBit(F::Detected) | StatisticFieldBit(F::Absorbed) |
                                       StatisticFieldBit(F::Escaped) | StatisticFieldBit(F::Truncated) |
                                       StatisticFieldBit(F::MaxBounce) | StatisticFieldBit(F::InvalidState) |
                                       StatisticFieldBit(F::ZeroStep);
        expect(out.statistics.validFields == 0, "empty batch has no batch fields");
        expect(out.eventStatistics[0].statistics.validFields == (terminal | StatisticFieldBit(F::Captured)),
               "empty batch event fields");
        // no records at all: no fields anywhere
        std::vector<OpticalEmission> none;
        std::vector<std::uint32_t> none_idx;
        host.EnqueueEmissions(c.scene, PhotonTransportBatch{none, ev, none_idx});
        out = host.CompleteOldestBatch();
        expect(out.statistics.validFields == 0 && out.eventStatistics.size() == 1 &&
                   out.eventStatistics[0].statistics.validFields == 0,
               "batch without records has no fields");
    }
    std::printf("%s (%d failures)\n", failures == 0 ? "PASS" : "FAIL", failures);
    return failures == 0 ? 0 : 1;
}

template <class T>
void put(std::FILE* f, const T& x) {
    std::fwrite(&x, sizeof(T), 1, f);
}

int run_gpu(const char* path, bool tapered, bool open = false) {
    CrystalScene c = crystal_scene(tapered, open);
    std::vector<OpticalEmission> em;
    std::vector<std::uint32_t> events, index;
    make_batch(c, em, events, index);
    PhotonTransportConfig cfg;
    cfg.seed = 4242;
    g4go_hip::HipTransportHost host{cfg};
    host.PrepareScene(c.scene);
    host.EnqueueEmissions(c.scene, PhotonTransportBatch{em, events, index});
    PhotonTransportOutput out = host.CompleteOldestBatch();
    std::FILE* f = std::fopen(path, "wb");
    if (f == nullptr) {
        std::printf("cannot write %s\n", path);
        return 1;
    }
    put<std::uint64_t>(f, em.size());
    put<std::uint64_t>(f, events.size());
    put<std::uint64_t>(f, out.detections.size());
    std::fwrite(em.data(), sizeof(OpticalEmission), em.size(), f);
    std::fwrite(events.data(), sizeof(std::uint32_t), events.size(), f);
    std::fwrite(index.data(), sizeof(std::uint32_t), index.size(), f);
    std::fwrite(out.detections.data(), sizeof(PhotonDetection), out.detections.size(), f);
    auto put_stats = [&](std::uint32_t id, const PhotonTransportStatistics& s) {
        put(f, id);
        put(f, s.validFields);
        for (std::uint64_t v : {s.capturedCount, s.detectedCount, s.absorbedCount, s.escapedCount, s.truncatedCount,
                                s.maxBounceCount, s.invalidStateCount, s.zeroStepCount}) {
            put(f, v);
        }
    };
    for (const auto& e : out.eventStatistics) {
        put_stats(e.eventID, e.statistics);
    }
    put_stats(0u, out.statistics);
    put<std::uint64_t>(f, out.performance.totalBounceCount);
    if (tapered) {
        g4go_hip::FlatScene<Scene> flat(c.scene, cfg.boundaryToleranceMm);
        if (flat.meshes().size() != 1) {
            std::printf("FAIL: expected one mesh solid, got %zu\n", flat.meshes().size());
            std::fclose(f);
            return 1;
        }
        const g4go_volume_mesh& m = flat.meshes()[0];
        put<std::uint64_t>(f, m.n_vertices);
        put<std::uint64_t>(f, m.n_triangles);
        std::fwrite(m.vertices_mm, sizeof(double), 3u * m.n_vertices, f);
        std::fwrite(m.indices, sizeof(std::uint32_t), 3u * m.n_triangles, f);
    }
    std::fclose(f);
    std::printf("transported %zu records, %zu detections, kernel %.3f ms\n", em.size(), out.detections.size(),
                out.performance.optiXKernelMs);
    return 0;
}

// Electron-like scintillation records (the shape of the 1 MeV electron events of
// the single-crystal workload): per event 20 records along the crystal axis,
// ~1500 photons each, decay 40 ns.
void electron_batch(const CrystalScene& c, std::uint32_t first_event, std::uint32_t n_events,
                    std::vector<OpticalEmission>& em, std::vector<std::uint32_t>& events,
                    std::vector<std::uint32_t>& index) {
    for (std::uint32_t e = 0; e < n_events; ++e) {
        const std::uint32_t event_id = first_event + e;
        events.push_back(event_id);
        std::uint32_t next_photon = 0;
        float x = 0.0F, y = 0.0F, z = -9.999F;
        for (int k = 0; k < 20; ++k) {
            OpticalEmission r{};
            const float dx = 0.2F * (lcg() - 0.5F), dy = 0.2F * (lcg() - 0.5F);
            r.positionMm = {x, y, z};
            r.timeNs = 0.0002F * k;
            r.direction = {dx, dy, 1.0F};
            r.stepLengthMm = 0.065F;
            r.stepDeltaMm = {0.065F * dx, 0.065F * dy, 0.065F};
            r.preVelocityMmPerNs = 281.8F;
            r.deltaVelocityMmPerNs = -2.8F;
            r.eventID = event_id;
            r.firstPhotonID = next_photon;
            r.volumeID = c.crystal;
            r.materialID = 1;
            r.type = OpticalEmissionType::Scintillation;
            r.decayTimeNs = 40.0F;
            r.charge = -1.0F;
            r.photonCount = 1400 + static_cast<std::uint32_t>(200.0F * lcg());
            next_photon += r.photonCount;
            em.push_back(r);
            index.push_back(e);
            x += r.stepDeltaMm[0];
            y += r.stepDeltaMm[1];
            z += r.stepDeltaMm[2];
        }
    }
}

void accumulate(PhotonTransportPerformance& sum, const PhotonTransportPerformance& p) {
    sum.hostToDeviceMs += p.hostToDeviceMs;
    sum.deviceMemsetMs += p.deviceMemsetMs;
    sum.optiXKernelMs += p.optiXKernelMs;
    sum.deviceHitCompactionMs += p.deviceHitCompactionMs;
    sum.deviceMetadataToHostMs += p.deviceMetadataToHostMs;
    sum.deviceHitsToHostMs += p.deviceHitsToHostMs;
    sum.deviceToHostMs += p.deviceToHostMs;
    sum.hostHitCompactionMs += p.hostHitCompactionMs;
}

// Throughput of the drop-in path: batches of 33 electron events (~1e6
// photons) through the three-method interface, hits returned to the host in
// launch order and converted to PhotonDetection records, with up to `slots`
// batches in flight.  Prints one JSON line: photons/s and the mean per-batch
// performance fields.
int run_bench(int slots, int batches, int threads, const g4go_launch* launch) {
    CrystalScene c = crystal_scene();
    const int warm = 3;
    std::vector<std::vector<OpticalEmission>> em(batches + warm);
    std::vector<std::vector<std::uint32_t>> ev(batches + warm), idx(batches + warm);
    std::uint64_t photons = 0;
    for (int b = 0; b < batches + warm; ++b) {
        electron_batch(c, 1000u * static_cast<std::uint32_t>(b), 33, em[b], ev[b], idx[b]);
        if (b >= warm) {
            for (const auto& r : em[b]) {
                photons += r.photonCount;
            }
        }
    }
    PhotonTransportConfig cfg;
    cfg.seed = 99;
    g4go_hip::HipTransportHost host{cfg, launch, static_cast<std::uint32_t>(slots),
                                    static_cast<std::uint32_t>(std::max(1, threads))};
    host.PrepareScene(c.scene);
    PhotonTransportPerformance sum{};
    std::uint64_t hits = 0;
    double enqueue_ms = 0.0, complete_ms = 0.0;
    const auto ms_since = [](std::chrono::steady_clock::time_point t) {
        return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t).count();
    };
    auto complete = [&](bool record) {
        const auto t = std::chrono::steady_clock::now();
        PhotonTransportOutput out = host.CompleteOldestBatch();
        if (record) {
            complete_ms += ms_since(t);
            hits += out.detections.size();
            accumulate(sum, out.performance);
        }
    };
    auto run = [&](int b0, int b1, bool record) {
        int in_flight = 0;
        for (int b = b0; b < b1; ++b) {
            if (in_flight == slots) {
                complete(record);
                --in_flight;
            }
            const auto t = std::chrono::steady_clock::now();
            host.EnqueueEmissions(c.scene, PhotonTransportBatch{em[b], ev[b], idx[b]});
            if (record) {
                enqueue_ms += ms_since(t);
            }
            ++in_flight;
        }
        while (in_flight > 0) {
            complete(record);
            --in_flight;
        }
    };
    run(0, warm, false);
    const auto t0 = std::chrono::steady_clock::now();
    run(warm, warm + batches, true);
    const double seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    const double n = batches;
    std::printf("{\"slots\": %d, \"conversion_threads\": %d, \"batches\": %d, \"photons_per_batch\": %.1f, "
                "\"photons_per_s\": %.6g, "
                "\"ms_per_batch\": %.4f, \"hits_per_batch\": %.1f, \"enqueue_ms\": %.4f, \"complete_ms\": %.4f, "
                "\"host_to_device_ms\": %.4f, \"device_memset_ms\": %.4f, \"kernel_ms\": %.4f, "
                "\"device_hit_order_ms\": %.4f, \"metadata_to_host_ms\": %.4f, \"hits_to_host_ms\": %.4f, "
                "\"host_hit_conversion_ms\": %.4f}\n",
                slots, std::max(1, threads), batches, static_cast<double>(photons) / n, static_cast<double>(photons) / seconds,
                1e3 * seconds / n, static_cast<double>(hits) / n, enqueue_ms / n, complete_ms / n,
                sum.hostToDeviceMs / n, sum.deviceMemsetMs / n, sum.optiXKernelMs / n, sum.deviceHitCompactionMs / n,
                sum.deviceMetadataToHostMs / n, sum.deviceHitsToHostMs / n, sum.hostHitCompactionMs / n);
    return 0;
}

}  // namespace

int main(int argc, char** argv) {
    try {
        if (argc >= 2 && std::string(argv[1]) == "cpu") {
            return run_cpu();
        }
        if (argc >= 3 && std::string(argv[1]) == "gpu") {
            return run_gpu(argv[2], false);
        }
        if (argc == 3 && std::string(argv[1]) == "gpu-mesh") {
            return run_gpu(argv[2], true);
        }
        if (argc == 3 && std::string(argv[1]) == "gpu-open") {
            return run_gpu(argv[2], true, true);
        }
        if ((argc == 4 || argc == 5 || argc == 11) && std::string(argv[1]) == "bench") {
            // optional: the conversion thread count, then a persistent form
            // BLOCK CHUNK BLOCKS_PER_CU REFILL GRAPH PREGENERATE (else the handle's default)
            g4go_launch form{1, 512, 256, 4, 2, 0, 0, 0};
            const int threads = argc >= 5 ? std::atoi(argv[4]) : 4;
            if (argc == 11) {
                form.block_size = static_cast<std::uint32_t>(std::atoi(argv[5]));
                form.chunk = static_cast<std::uint32_t>(std::atoi(argv[6]));
                form.blocks_per_cu = static_cast<std::uint32_t>(std::atoi(argv[7]));
                form.refill_threshold = static_cast<std::uint32_t>(std::atoi(argv[8]));
                form.graph = static_cast<std::uint32_t>(std::atoi(argv[9]));
                form.pregenerate = static_cast<std::uint32_t>(std::atoi(argv[10]));
            }
            return run_bench(std::atoi(argv[2]), std::atoi(argv[3]), threads, argc == 11 ? &form : nullptr);
        }
    } catch (const std::exception& e) {
        std::printf("FAIL: uncaught %s: %s\n", typeid(e).name(), e.what());
        return 1;
    }
    std::printf("usage: shim_test cpu | shim_test gpu OUT.bin | shim_test gpu-mesh OUT.bin | shim_test gpu-open OUT.bin | shim_test bench SLOTS BATCHES [THREADS [BLOCK CHUNK BLOCKS_PER_CU REFILL GRAPH PREGENERATE]]\n");
    return 2;
}
