// Optical-photon physics shared by every transport kernel of the gfx950 backend.
//
// All arithmetic here is written so that a CPU restatement compiled with
// IEEE single precision, no FP contraction and the same operation order
// reproduces it bit for bit:
//   * add, sub, mul, div and sqrt are correctly rounded on both sides
//     (hipcc: -ffp-contract=off -fhip-fp32-correctly-rounded-divide-sqrt);
//   * fused multiply-adds are written explicitly (fmaf), never contracted;
//   * log, exp, sin and cos are project-written from Cephes (S. L. Moshier,
//     "Methods and Programs for Mathematical Functions", 1989) minimax
//     polynomials instead of the device library, whose last-ulp behaviour
//     differs from any host libm.
//
// Reference algorithm: src/optical/optix/OptiXBoundaryPhysics.cuh and
// src/optical/optix/OptiXTransportDevice.cu of G4GO.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define G4GO_HD __host__ __device__ __forceinline__

namespace g4go {

constexpr uint32_t kInvalid = 0xffffffffu;
constexpr float kInf = __builtin_huge_valf();
constexpr float kSpeedOfLightMmPerNs = 299.792458f;
constexpr float kHcEvNm = 1239.84198f;  // h*c in eV*nm

G4GO_HD uint32_t f2u(float x) { return __builtin_bit_cast(uint32_t, x); }
G4GO_HD float u2f(uint32_t x) { return __builtin_bit_cast(float, x); }
G4GO_HD bool finitef(float x) { return __builtin_isfinite(x); }

// ---------------------------------------------------------------------------
// Philox4x32-10 (Salmon et al., SC'11).  Counter layout follows the reference
// Uniform() (OptiXTransportDevice.cu:254-286):
//   counter = {lo32(photonKey), hi32(photonKey), step, (process << 16) ^ block}
//   key     = {lo32(seed), hi32(seed)}
// with photonKey = (eventID << 32) | photonID.  One block yields four words;
// word k of block b is draw 4b+k of that (step, process).  Draw 4b is the
// reference's Uniform(seed, key, step, process, 4b) exactly.
// ---------------------------------------------------------------------------
struct U4 {
    uint32_t x, y, z, w;
};

// a ^ b ^ c.  On gfx950 one v_bitop3_b32 (truth table 0x96) instead of two
// XORs; the host overload is the plain expression.
__device__ __forceinline__ uint32_t xor3(uint32_t a, uint32_t b, uint32_t c) {
    return __builtin_amdgcn_bitop3_b32(a, b, c, 0x96);
}
__host__ inline uint32_t xor3(uint32_t a, uint32_t b, uint32_t c) { return a ^ b ^ c; }

G4GO_HD U4 philox(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t k0, uint32_t k1) {
#pragma unroll
    for (int round = 0; round < 10; ++round) {
        const uint64_t p0 = static_cast<uint64_t>(0xD2511F53u) * c0;
        const uint64_t p1 = static_cast<uint64_t>(0xCD9E8D57u) * c2;
        const uint32_t n0 = xor3(static_cast<uint32_t>(p1 >> 32), c1, k0);
        const uint32_t n2 = xor3(static_cast<uint32_t>(p0 >> 32), c3, k1);
        c1 = static_cast<uint32_t>(p1);
        c3 = static_cast<uint32_t>(p0);
        c0 = n0;
        c2 = n2;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    return {c0, c1, c2, c3};
}

// (float(word) + 0.5) / 2^32, bit-identical to the reference mapping
// (OptiXTransportDevice.cu:285); the result lies in (0, 1].
G4GO_HD float u01(uint32_t word) { return fmaf(static_cast<float>(word), 0x1p-32f, 0x1p-33f); }

// Process identifiers of this backend's draw map.
enum Process : uint32_t {
    kProcStep = 1,       // per step: w0 distance, w1 bulk process, w2 surface class, w3 Fresnel/detect
    kProcRayleigh = 2,   // Rayleigh angular sampling (one block: w0-w2 median and sign, w3 azimuth)
    kProcFacet = 3,      // ground-surface facet normal, block = attempt
    kProcLobe = 4,       // unified-model lobe choice and Lambertian draws
    kProcRough = 5,      // roughness test and Lambertian draws
    kProcSynthetic = 6,  // synthetic source generation (step 0)
    kProcEmission = 7,   // emission-record expansion (step 0)
};

struct Rng {
    uint32_t seed_lo, seed_hi;
    G4GO_HD U4 block(uint32_t key_lo, uint32_t key_hi, uint32_t step, uint32_t process, uint32_t blk) const {
        return philox(key_lo, key_hi, step, (process << 16) ^ blk, seed_lo, seed_hi);
    }
};

// ---------------------------------------------------------------------------
// Elementary functions (Cephes single precision, evaluated with explicit fma).
// ---------------------------------------------------------------------------

// Natural logarithm; x > 0.  Cephes logf: x = m * 2^e, m in [sqrt(1/2), sqrt(2)),
// log(1+f) = f - f^2/2 + f^3 P(f), Cody-Waite split of ln 2.
G4GO_HD float g4_logf(float x) {
    if (!(x > 0.0f)) {
        return x == 0.0f ? -kInf : u2f(0x7fc00000u);
    }
    if (!finitef(x)) {
        return x;
    }
    uint32_t b = f2u(x);
    int e = 0;
    if (b < 0x00800000u) {  // subnormal input
        x = x * 33554432.0f;
        b = f2u(x);
        e = -25;
    }
    e += static_cast<int>(b >> 23) - 126;
    float m = u2f((b & 0x007fffffu) | 0x3f000000u);  // [0.5, 1)
    // branch-free: the lanes of a wave fall on both sides
    const bool low = m < 0.70710678118654752f;
    e -= low ? 1 : 0;
    m = low ? (m + m) - 1.0f : m - 1.0f;
    const float z = m * m;
    float y = 7.0376836292e-2f;
    y = fmaf(y, m, -1.1514610310e-1f);
    y = fmaf(y, m, 1.1676998740e-1f);
    y = fmaf(y, m, -1.2420140846e-1f);
    y = fmaf(y, m, 1.4249322787e-1f);
    y = fmaf(y, m, -1.6668057665e-1f);
    y = fmaf(y, m, 2.0000714765e-1f);
    y = fmaf(y, m, -2.4999993993e-1f);
    y = fmaf(y, m, 3.3333331174e-1f);
    y = (y * m) * z;
    const float fe = static_cast<float>(e);
    y = fmaf(fe, -2.12194440e-4f, y);
    y = fmaf(-0.5f, z, y);
    float r = m + y;
    r = fmaf(fe, 0.693359375f, r);
    return r;
}

// Exponential.  Cephes expf: n = round(x / ln 2), Cody-Waite reduction,
// degree-6 polynomial, exact scaling by 2^n.  Results below the normal
// range are flushed to zero (the backend only uses exp(-a) probabilities).
G4GO_HD float g4_expf(float x) {
    if (x != x) {
        return x;
    }
    if (x > 88.7228391f) {
        return kInf;
    }
    if (x < -87.3365447f) {
        return 0.0f;
    }
    const float n = floorf(fmaf(x, 1.44269504088896341f, 0.5f));
    float r = fmaf(n, -0.693359375f, x);
    r = fmaf(n, 2.12194440e-4f, r);
    const float z = r * r;
    float y = 1.9875691500e-4f;
    y = fmaf(y, r, 1.3981999507e-3f);
    y = fmaf(y, r, 8.3334519073e-3f);
    y = fmaf(y, r, 4.1665795894e-2f);
    y = fmaf(y, r, 1.6666665459e-1f);
    y = fmaf(y, r, 5.0000001201e-1f);
    y = fmaf(y, z, r);
    y = y + 1.0f;
    int ni = static_cast<int>(n);
    if (ni > 127) {
        y = y * 2.0f;
        ni -= 1;
    }
    return y * u2f(static_cast<uint32_t>(ni + 127) << 23);
}

// sin and cos of 2*pi*u.  The reduction is done in turns (exact), then the
// Cephes sinf/cosf minimax polynomials on [-pi/4, pi/4].
G4GO_HD void g4_sincos_turn(float u, float& s, float& c) {
    const float t = u * 4.0f;
    const float q = floorf(t + 0.5f);
    const float x = (t - q) * 1.57079632679489662f;
    const float z = x * x;
    const float sp = fmaf(fmaf(-1.9515295891e-4f, z, 8.3321608736e-3f), z, -1.6666654611e-1f);
    const float sn = fmaf(sp * z, x, x);
    const float cp = fmaf(fmaf(2.443315711809948e-5f, z, -1.388731625493765e-3f), z, 4.166664568298827e-2f);
    const float cs = fmaf(cp * z, z, fmaf(-0.5f, z, 1.0f));
    // quadrant by selects, not branches (the lanes of a wave fall in every
    // quadrant): (s, c) = (sn, cs), (cs, -sn), (-sn, -cs), (-cs, sn)
    const int quadrant = static_cast<int>(q) & 3;
    const bool swap = (quadrant & 1) != 0;
    const float a = swap ? cs : sn;
    const float b = swap ? sn : cs;
    s = (quadrant & 2) != 0 ? -a : a;
    c = ((quadrant + 1) & 2) != 0 ? -b : b;
}

// ---------------------------------------------------------------------------
// Vectors (same operation order as OptiXBoundaryPhysics.cuh:16-45).
// ---------------------------------------------------------------------------
struct V3 {
    float x, y, z;
};

G4GO_HD V3 add(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
G4GO_HD V3 sub(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
G4GO_HD V3 scale(V3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
G4GO_HD float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
G4GO_HD V3 cross(V3 a, V3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
G4GO_HD float axis_of(V3 a, int k) { return k == 0 ? a.x : (k == 1 ? a.y : a.z); }

// OptiXBoundaryPhysics.cuh:47-54
G4GO_HD V3 normalize(V3 a) {
    const float l2 = dot(a, a);
    if (!(l2 > 0.0f)) {
        return {0.0f, 0.0f, 0.0f};
    }
    return scale(a, 1.0f / sqrtf(l2));
}

G4GO_HD float clamp01(float v) { return fminf(fmaxf(v, 0.0f), 1.0f); }

// Specular reflection at a face normal to `axis` (0, 1, 2): d - 2 (d.n) n
// negates the direction's axis component, and -p + 2 (p.n) n negates the
// polarisation's other two components.  Both are exact sign flips, so norms
// and p.d are kept and no renormalisation is needed.
G4GO_HD V3 negate_axis(V3 a, int axis) {
    return {axis == 0 ? -a.x : a.x, axis == 1 ? -a.y : a.y, axis == 2 ? -a.z : a.z};
}
G4GO_HD V3 negate_off_axis(V3 a, int axis) {
    return {axis == 0 ? a.x : -a.x, axis == 1 ? a.y : -a.y, axis == 2 ? a.z : -a.z};
}

// OptiXBoundaryPhysics.cuh:57-67
G4GO_HD V3 project_polarization(V3 pol, V3 dir) {
    V3 out = sub(pol, scale(dir, dot(pol, dir)));
    if (dot(out, out) < 1.0e-12f) {
        const V3 helper = fabsf(dir.z) < 0.9f ? V3{0.0f, 0.0f, 1.0f} : V3{1.0f, 0.0f, 0.0f};
        out = cross(dir, helper);
    }
    return normalize(out);
}

// OptiXBoundaryPhysics.cuh:69-71
G4GO_HD V3 reflect_direction(V3 dir, V3 n) { return normalize(sub(dir, scale(n, 2.0f * dot(dir, n)))); }

// OptiXBoundaryPhysics.cuh:73-77
G4GO_HD V3 reflect_polarization(V3 pol, V3 n, V3 dir) {
    const V3 r = add(scale(pol, -1.0f), scale(n, 2.0f * dot(pol, n)));
    return project_polarization(r, dir);
}

enum SurfaceOutcome : int { kAbsorb = 0, kDetect = 1, kDirectTransmit = 2, kInteraction = 3 };

// OptiXBoundaryPhysics.cuh:79-94
G4GO_HD int classify_surface(float reflectivity, float transmittance, bool has_r, bool has_t, float sample) {
    const float u = clamp01(sample);
    const float r = has_r ? clamp01(reflectivity) : 0.0f;
    const float t = has_t ? clamp01(transmittance) : 0.0f;
    if (!has_r && !has_t) {
        return kInteraction;
    }
    if (u > r + t) {
        return kAbsorb;
    }
    if (u > r) {
        return kDirectTransmit;
    }
    return kInteraction;
}

// OptiXBoundaryPhysics.cuh:96-101
G4GO_HD int evaluate_absorption(float efficiency, bool sensor, float sample) {
    return (sensor && sample < clamp01(efficiency)) ? kDetect : kAbsorb;
}

// OptiXBoundaryPhysics.cuh:103-120 (cos/sin of the azimuth via g4_sincos_turn)
G4GO_HD V3 sample_lambertian(V3 n, float radial_sample, float azimuth_sample) {
    n = normalize(n);
    radial_sample = clamp01(radial_sample);
    azimuth_sample = clamp01(azimuth_sample);
    const float radial = sqrtf(radial_sample);
    float s, c;
    g4_sincos_turn(azimuth_sample, s, c);
    const float normal_component = sqrtf(fmaxf(0.0f, 1.0f - radial_sample));
    const V3 helper = fabsf(n.z) < 0.9f ? V3{0.0f, 0.0f, 1.0f} : V3{1.0f, 0.0f, 0.0f};
    const V3 tangent = normalize(cross(helper, n));
    const V3 bitangent = cross(n, tangent);
    return normalize(add(add(scale(tangent, radial * c), scale(bitangent, radial * s)), scale(n, normal_component)));
}

struct Fresnel {
    V3 reflected_dir, transmitted_dir, reflected_pol, transmitted_pol;
    double transmittance;
    bool tir;
};

struct D3 {
    double x, y, z;
};
G4GO_HD D3 dvec(V3 a) { return {static_cast<double>(a.x), static_cast<double>(a.y), static_cast<double>(a.z)}; }
G4GO_HD double ddot(D3 a, D3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
G4GO_HD D3 dcross(D3 a, D3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }

// Fresnel equations with polarisation (OptiXBoundaryPhysics.cuh:122-178),
// in two phases: fresnel_state() gives the reflect/transmit probability, then
// fresnel_reflect() or fresnel_transmit() builds only the outgoing direction
// and polarisation that the draw selects.  Directions and polarisations are
// FP32 with the reference's operation order.  The probability and the
// total-internal-reflection test are evaluated in FP64 from the same FP32
// state (as Geant4's G4OpBoundaryProcess does), which keeps them accurate
// near the critical angle where 1 - sin^2(theta_t) cancels.  The s/p split
// of the polarisation uses |p.s|^2 = (p.(d x n))^2 / |d x n|^2 and
// |p.s|^2 + |p.p|^2 = |p|^2 - (p.d)^2 / |d|^2 (the squared projections on the
// unit s and p directions, both perpendicular to d).
struct FresnelState {
    V3 dir, n, pol;  // normalised direction and normal, polarisation projected perpendicular to dir
    float cos_i, eta, cos_t;
    double transmittance;
    bool tir;
    bool normal_incidence;  // |dir x n|^2 < 1e-12: no s/p split
};

G4GO_HD FresnelState fresnel_state(V3 dir, V3 pol, V3 n, float n1, float n2) {
    FresnelState f;
    f.dir = normalize(dir);
    f.n = normalize(n);
    f.pol = project_polarization(pol, f.dir);
    const D3 dd = dvec(f.dir), dn = dvec(f.n);
    const double n1d = n1, n2d = n2;
    const double ci_d = fmin(fmax(-ddot(dd, dn), 0.0), 1.0);
    const double eta_d = n1d / n2d;
    const double st2_d = eta_d * eta_d * (1.0 - ci_d * ci_d);
    f.cos_i = clamp01(-dot(f.dir, f.n));
    f.eta = n1 / n2;
    const float sin_t2 = f.eta * f.eta * (1.0f - f.cos_i * f.cos_i);
    f.tir = st2_d >= 1.0;
    f.transmittance = 0.0;
    f.cos_t = sqrtf(fmaxf(0.0f, 1.0f - sin_t2));
    const V3 transverse = cross(f.dir, f.n);
    f.normal_incidence = dot(transverse, transverse) < 1.0e-12f;
    if (f.tir) {
        return f;
    }
    if (f.normal_incidence) {
        const double ra = (n1d - n2d) / (n1d + n2d);
        f.transmittance = fmin(fmax(1.0 - ra * ra, 0.0), 1.0);
        return f;
    }
    const double ct_d = sqrt(fmax(0.0, 1.0 - st2_d));
    const D3 dp = dvec(f.pol);
    const D3 c = dcross(dd, dn);
    const double pc = ddot(dp, c);
    const double es2 = pc * pc / ddot(c, c);
    const double pdd = ddot(dp, dd);
    const double perp2 = ddot(dp, dp) - pdd * pdd / ddot(dd, dd);
    const double rs = (n1d * ci_d - n2d * ct_d) / (n1d * ci_d + n2d * ct_d);
    const double rp = (n2d * ci_d - n1d * ct_d) / (n2d * ci_d + n1d * ct_d);
    f.transmittance = fmin(fmax(1.0 - (es2 * (rs * rs) + (perp2 - es2) * (rp * rp)), 0.0), 1.0);
    return f;
}

G4GO_HD void fresnel_reflect(const FresnelState& f, float n1, float n2, V3& dir, V3& pol) {
    dir = reflect_direction(f.dir, f.n);
    if (f.tir) {
        pol = reflect_polarization(f.pol, f.n, dir);
        return;
    }
    if (f.normal_incidence) {
        pol = n2 > n1 ? scale(f.pol, -1.0f) : f.pol;
        return;
    }
    const V3 s_dir = normalize(cross(f.dir, f.n));
    const V3 p_dir = normalize(cross(f.dir, s_dir));
    const float in_s = dot(f.pol, s_dir);
    const float in_p = dot(f.pol, p_dir);
    const float rs = (n1 * f.cos_i - n2 * f.cos_t) / (n1 * f.cos_i + n2 * f.cos_t);
    const float rp = (n2 * f.cos_i - n1 * f.cos_t) / (n2 * f.cos_i + n1 * f.cos_t);
    const V3 rp_dir = normalize(cross(dir, s_dir));
    pol = normalize(add(scale(s_dir, in_s * rs), scale(rp_dir, in_p * rp)));
}

// Not defined under total internal reflection.
G4GO_HD void fresnel_transmit(const FresnelState& f, float n1, float n2, V3& dir, V3& pol) {
    dir = normalize(add(scale(f.dir, f.eta), scale(f.n, f.eta * f.cos_i - f.cos_t)));
    if (f.normal_incidence) {
        pol = f.pol;
        return;
    }
    const V3 s_dir = normalize(cross(f.dir, f.n));
    const V3 p_dir = normalize(cross(f.dir, s_dir));
    const float in_s = dot(f.pol, s_dir);
    const float in_p = dot(f.pol, p_dir);
    const float ts = 2.0f * n1 * f.cos_i / (n1 * f.cos_i + n2 * f.cos_t);
    const float tp = 2.0f * n1 * f.cos_i / (n2 * f.cos_i + n1 * f.cos_t);
    const V3 tp_dir = normalize(cross(dir, s_dir));
    pol = normalize(add(scale(s_dir, in_s * ts), scale(tp_dir, in_p * tp)));
}

// Both outcomes at once (for the evaluation entry point).
G4GO_HD Fresnel compute_fresnel(V3 dir, V3 pol, V3 n, float n1, float n2) {
    const FresnelState f = fresnel_state(dir, pol, n, n1, n2);
    Fresnel out{};
    out.tir = f.tir;
    out.transmittance = f.transmittance;
    fresnel_reflect(f, n1, n2, out.reflected_dir, out.reflected_pol);
    if (!f.tir) {
        fresnel_transmit(f, n1, n2, out.transmitted_dir, out.transmitted_pol);
    }
    return out;
}

// Rotates v from the frame whose z axis is u into the global frame
// (CLHEP Hep3Vector::rotateUz, used by Geant4 G4OpRayleigh).
G4GO_HD V3 rotate_uz(V3 v, V3 u) {
    const float up2 = u.x * u.x + u.y * u.y;
    if (up2 > 0.0f) {
        const float up = sqrtf(up2);
        return {(u.x * u.z * v.x - u.y * v.y) / up + u.x * v.z, (u.y * u.z * v.x + u.x * v.y) / up + u.y * v.z,
                -up * v.x + u.z * v.z};
    }
    if (u.z < 0.0f) {
        return {-v.x, v.y, -v.z};
    }
    return v;
}

}  // namespace g4go
