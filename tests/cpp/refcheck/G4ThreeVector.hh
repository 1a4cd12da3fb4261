// Declaration-only stand-in for CLHEP's G4ThreeVector (absent from this
// image), holding just what the reference's Scene.hpp uses: construction from
// three doubles and x(), y(), z().  It lets tests/cpp/reference_types_check.cpp
// type-check the C++ shim against the reference's own data-type headers with
// -fsyntax-only; nothing is built or run from it.
#pragma once

class G4ThreeVector {
public:
    G4ThreeVector() = default;
    G4ThreeVector(double x, double y, double z) : dx(x), dy(y), dz(z) {}
    double x() const { return dx; }
    double y() const { return dy; }
    double z() const { return dz; }

private:
    double dx{}, dy{}, dz{};
};
