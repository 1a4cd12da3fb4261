// Type check of the C++ shim instantiated with the reference's own data types
// (include/g4go/optical/PhotonTransport.hpp, PhotonTransportConfig.hpp and
// Scene.hpp of the reference tree), exactly as the type bundle a maintainer
// adds to the reference declares it.  Compiled with -fsyntax-only against the
// reference's headers where they lie; the explicit instantiation makes every
// member function of BasicHipTransportHost<HipTypes> compile.
#include "g4go/optical/PhotonTransport.hpp"
#include "g4go/optical/PhotonTransportConfig.hpp"
#include "g4go_hip_host.hpp"

namespace G4GO::Optical {
struct HipTypes {
    using Scene = G4GO::Optical::Scene;
    using PhotonTransportConfig = G4GO::Optical::PhotonTransportConfig;
    using PhotonTransportBatch = G4GO::Optical::PhotonTransportBatch;
    using PhotonTransportOutput = G4GO::Optical::PhotonTransportOutput;
    using PhotonDetection = G4GO::Optical::PhotonDetection;
    using OpticalEmission = G4GO::Optical::OpticalEmission;
    using PhotonTransportEventStatistics = G4GO::Optical::PhotonTransportEventStatistics;
    using PhotonTransportStatisticField = G4GO::Optical::PhotonTransportStatisticField;
};
}  // namespace G4GO::Optical

template class g4go_hip::BasicHipTransportHost<G4GO::Optical::HipTypes>;
template class g4go_hip::FlatScene<G4GO::Optical::Scene>;

// The scheduler's calls (OpticalBatchScheduler.cpp:552, 567-568, 453).
[[maybe_unused]] static G4GO::Optical::PhotonTransportOutput scheduler_calls(
    const G4GO::Optical::PhotonTransportConfig& config, const G4GO::Optical::Scene& scene,
    const G4GO::Optical::PhotonTransportBatch& batch) {
    g4go_hip::BasicHipTransportHost<G4GO::Optical::HipTypes> host{config};
    host.PrepareScene(scene);
    host.EnqueueEmissions(scene, batch);
    return host.CompleteOldestBatch();
}
