// C entry points around the reference's own boundary-physics header,
// src/optical/optix/OptiXBoundaryPhysics.cuh of G4GO, compiled from where it
// lies in the reference tree (see oracle/build.py).  Test infrastructure only:
// it lets the oracle's restatement be compared with the reference code itself.
//
// The header needs only OptiXDeviceData.cuh and the C++ standard library; it
// is compiled as host code (G4GO_OPTICAL_HD expands to nothing without nvcc,
// OptiXBoundaryPhysics.cuh:7-11).
#include "OptiXBoundaryPhysics.cuh"

namespace P = G4GO::Optical::BoundaryPhysics;
using G4GO::Optical::DeviceVector3;

namespace {
DeviceVector3 v3(const float* a) { return {a[0], a[1], a[2]}; }
void put(float* o, const DeviceVector3& v) {
    o[0] = v.at(0);
    o[1] = v.at(1);
    o[2] = v.at(2);
}
}  // namespace

extern "C" {

// 11 inputs (dir[3], pol[3], normal[3], n1, n2) -> 14 outputs
// (transmittance, tir, reflected dir, transmitted dir, reflected pol, transmitted pol).
void ref_compute_fresnel(const float* in, float* out) {
    const P::FresnelOutput f = P::ComputeFresnel(v3(in), v3(in + 3), v3(in + 6), in[9], in[10]);
    out[0] = f.transmittance;
    out[1] = f.totalInternalReflection ? 1.0f : 0.0f;
    put(out + 2, f.reflectedDirection);
    put(out + 5, f.transmittedDirection);
    put(out + 8, f.reflectedPolarization);
    put(out + 11, f.transmittedPolarization);
}

void ref_project_polarization(const float* pol, const float* dir, float* out) {
    put(out, P::ProjectPolarization(v3(pol), v3(dir)));
}

void ref_reflect_direction(const float* dir, const float* normal, float* out) {
    put(out, P::ReflectDirection(v3(dir), v3(normal)));
}

void ref_reflect_polarization(const float* pol, const float* normal, const float* dir, float* out) {
    put(out, P::ReflectPolarization(v3(pol), v3(normal), v3(dir)));
}

int ref_classify_surface(float r, float t, int has_r, int has_t, float sample) {
    return static_cast<int>(P::ClassifySurface(r, t, has_r != 0, has_t != 0, sample));
}

int ref_evaluate_absorption(float efficiency, int sensor, float sample) {
    return static_cast<int>(P::EvaluateAbsorption(efficiency, sensor != 0, sample));
}

void ref_sample_lambertian(const float* normal, float radial, float azimuth, float* out) {
    put(out, P::SampleLambertian(v3(normal), radial, azimuth));
}

}  // extern "C"
