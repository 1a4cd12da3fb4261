// Per-photon transport: property lookup, box geometry, photon generation and
// one transport step.  Used by the plain, persistent and adjoint kernels.
//
// Reference: src/optical/optix/OptiXTransportDevice.cu:198-595 (helpers) and
// :660-941 (the bounce loop of __raygen__rg).
#pragma once

#include "physics.h"
#include "region_prof.h"
#include "scene.h"

namespace g4go {

// OpticalEmission, 128 bytes (include/g4go/optical/PhotonTransport.hpp:40-69).
struct alignas(16) Emission {
    float pos[3];
    float time;
    float dir[3];
    float step_length;
    float step_delta[3];
    float pre_velocity;
    float pol[3];
    float delta_velocity;
    float energy;
    float pre_mean;
    float post_mean;
    float decay;
    float rise;
    float charge;
    uint32_t event_id;
    uint32_t first_photon_id;
    uint32_t photon_count;
    uint32_t volume;  // device volume index (remapped on the host)
    uint32_t material;
    uint32_t spectrum;
    uint8_t type;  // 0 direct, 1 Cerenkov, 2 scintillation
    uint8_t pad[15];
};
static_assert(sizeof(Emission) == 128, "OpticalEmission layout");

// Detected-photon record written by the kernels (48 bytes, FP64 time).
struct Hit {
    float pos[3];
    float dir[3];
    float energy;
    uint32_t sensor;
    uint32_t event_id;
    uint32_t photon_id;
    double time;
};
static_assert(sizeof(Hit) == 48, "hit layout");

// Synthetic source: photons uniform in a box, isotropic, random linear
// polarisation, Gaussian wavelength and exponential emission time.
struct SyntheticSource {
    float lo[3], hi[3];
    float wavelength_mean_nm, wavelength_sigma_nm;
    float decay_ns;
    uint32_t volume;
};

struct Photon {
    V3 pos, dir, pol;
    double time;
    float energy;
    uint32_t volume;
    uint32_t step;
    uint32_t photon_id, event_id;
    uint32_t material;  // material whose values are cached below
    float rindex, inv_velocity, interaction_length, absorb_fraction;
    // node of the photon's energy on the scene's hot grid: grid_i = kInvalid
    // until located; grid_f < 0 marks a clamped value (values[grid_i] itself)
    uint32_t grid_i;
    float grid_f;
};

enum StepStatus : int {
    kAlive = 0,
    kDetected = 1,
    kAbsorbedBulk = 2,
    kAbsorbedSurface = 3,
    kEscaped = 4,
    kInvalidState = 5,
};

// What one step did, for the path-replay adjoint (likelihood-ratio scores).
struct StepTrace {
    float segment;      // path length travelled in the bulk of `material`
    uint32_t material;  // material of that segment
    uint32_t surface;   // surface met at the boundary (kInvalid = none or a bulk event)
    int event;          // kTraceNone, kTraceRayleigh, kTraceInteraction, kTraceTransmit, kTraceDetect
};
enum TraceEvent : int {
    kTraceNone = 0,
    kTraceRayleigh = 1,     // bulk Rayleigh scatter (probability mu_s / mu)
    kTraceInteraction = 2,  // surface classified as interaction (probability R)
    kTraceTransmit = 3,     // surface classified as direct transmission (probability T)
    kTraceDetect = 4,       // surface absorption (1 - R - T) followed by detection (efficiency)
};

// ---------------------------------------------------------------------------
// Property tables (OptiXTransportDevice.cu:198-216): clamped piecewise-linear
// interpolation over ascending energies.  The bracketing node is found by a
// binary search, which selects the same node as the reference's linear scan.
// ---------------------------------------------------------------------------
// One 16-byte read of a property descriptor (instead of one load per field
// on each branch that needs it).
G4GO_HD Prop load_prop(const Prop& p) {
    Prop q;
    __builtin_memcpy(&q, &p, sizeof(Prop));
    return q;
}

G4GO_HD float prop_eval(const float* pool, const Prop& pref, float e, float fallback) {
    const Prop p = load_prop(pref);
    if (p.count == 0) {
        return fallback;
    }
    if ((p.flags & kPropConstant) != 0) {
        return p.value0;
    }
    const float* en = pool + p.offset;
    const float* va = en + p.count;
    if (e <= en[0]) {
        return va[0];
    }
    const uint32_t last = p.count - 1;
    if (e >= en[last]) {
        return va[last];
    }
    uint32_t lo = 1, hi = last;  // first j in [1, last] with e <= en[j]
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (e > en[mid]) {
            lo = mid + 1;
        } else {
            hi = mid;
        }
    }
    const uint32_t i = lo - 1;
    const float f = (e - en[i]) / (en[i + 1] - en[i]);
    return va[i] + f * (va[i + 1] - va[i]);
}

// Node of energy e on a grid of `count` ascending energies, as prop_eval
// selects it: f < 0 marks a clamped value (e at or below the first node, or at
// or above the last), otherwise the value is va[i] + f (va[i+1] - va[i]) with
// the f prop_eval computes.
G4GO_HD void grid_locate(const float* en, uint32_t count, float e, uint32_t& i, float& f) {
    const uint32_t last = count - 1;
    if (count <= 1 || e <= en[0]) {
        i = 0;
        f = -1.0f;
        return;
    }
    if (e >= en[last]) {
        i = last;
        f = -1.0f;
        return;
    }
    uint32_t lo = 1, hi = last;
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (e > en[mid]) {
            lo = mid + 1;
        } else {
            hi = mid;
        }
    }
    i = lo - 1;
    f = (e - en[i]) / (en[i + 1] - en[i]);
}

// prop_eval at the photon's energy, through its located hot-grid node when the
// table lies on the hot grid (bit-identical to prop_eval).
G4GO_HD float prop_eval_ph(const float* pool, const Prop& pref, const Photon& ph, float fallback) {
    const Prop p = load_prop(pref);
    if (p.count == 0) {
        return fallback;
    }
    if ((p.flags & kPropConstant) != 0) {
        return p.value0;
    }
    if ((p.flags & kPropHotGrid) != 0 && ph.grid_i != kInvalid) {
        const float* va = pool + p.offset + p.count;
        const uint32_t i = ph.grid_i;
        if (ph.grid_f < 0.0f) {
            return va[i];
        }
        return va[i] + ph.grid_f * (va[i + 1] - va[i]);
    }
    return prop_eval(pool, p, ph.energy, fallback);
}

// Interpolation node and weight of prop_eval (for parameter gradients):
// value = (1 - w) * va[i] + w * va[i + 1] in exact arithmetic.  The constant
// shortcut is ignored: a parameterised table may hold equal values, and its
// value is the same either way.
G4GO_HD void prop_node(const float* pool, const Prop& p, float e, uint32_t& i, float& w) {
    const float* en = pool + p.offset;
    const uint32_t last = p.count - 1;
    if (p.count <= 1 || e <= en[0]) {
        i = 0;
        w = 0.0f;
        return;
    }
    if (e >= en[last]) {
        i = last;
        w = 0.0f;
        return;
    }
    uint32_t lo = 1, hi = last;
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (e > en[mid]) {
            lo = mid + 1;
        } else {
            hi = mid;
        }
    }
    i = lo - 1;
    w = (e - en[i]) / (en[i + 1] - en[i]);
}

// Inverse-CDF sampling of a normalised spectrum (OptiXTransportDevice.cu:322-344).
G4GO_HD float sample_spectrum(const float* pool, const Prop& p, float random) {
    if (p.count == 0) {
        return u2f(0x7fc00000u);
    }
    const float* en = pool + p.offset;
    const float* cdf = en + p.count;
    const float u = clamp01(random);
    if (p.count == 1 || u <= cdf[0]) {
        return en[0];
    }
    const uint32_t last = p.count - 1;
    if (u >= cdf[last]) {
        return en[last];
    }
    uint32_t lo = 1, hi = last;  // first j with u <= cdf[j]
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (u > cdf[mid]) {
            lo = mid + 1;
        } else {
            hi = mid;
        }
    }
    const uint32_t i = lo - 1;
    const float den = cdf[i + 1] - cdf[i];
    if (!(den > 0.0f)) {
        return en[i];
    }
    const float f = (u - cdf[i]) / den;
    return en[i] + f * (en[i + 1] - en[i]);
}

// Caches the bulk properties of `material` at the photon energy.
// Returns false when the material cannot transport (reference :720-727).
G4GO_HD bool cache_material(Photon& ph, const SceneView& sc, uint32_t material) {
    if (material >= sc.n_materials) {
        return false;
    }
    const Material& m = sc.materials[material];
    const float n = prop_eval_ph(sc.pool, m.rindex, ph, u2f(0x7fc00000u));
    const bool n_ok = finitef(n) && n > 0.0f;
    const float v = prop_eval_ph(sc.pool, m.group_velocity, ph,
                              n_ok ? kSpeedOfLightMmPerNs / n : u2f(0x7fc00000u));
    if (!(v > 0.0f) || !finitef(v)) {
        return false;
    }
    const float la = prop_eval_ph(sc.pool, m.abs_length, ph, kInf);
    const float lr = prop_eval_ph(sc.pool, m.rayleigh_length, ph, kInf);
    const bool a_ok = finitef(la) && la > 0.0f;
    const bool r_ok = finitef(lr) && lr > 0.0f;
    ph.material = material;
    ph.rindex = n;
    ph.inv_velocity = 1.0f / v;
    if (a_ok && r_ok) {
        ph.interaction_length = (la * lr) / (la + lr);
        ph.absorb_fraction = lr / (la + lr);
    } else if (a_ok) {
        ph.interaction_length = la;
        ph.absorb_fraction = 1.0f;
    } else if (r_ok) {
        ph.interaction_length = lr;
        ph.absorb_fraction = 0.0f;
    } else {
        ph.interaction_length = kInf;
        ph.absorb_fraction = 1.0f;
    }
    return true;
}

// ---------------------------------------------------------------------------
// Geometry: nearest boundary inside an axis-aligned box volume.
// ---------------------------------------------------------------------------

// Exit distance from the box along dir.  The face is chosen by comparing
// (distance to plane) * |d_other| cross products, so only one division is
// needed.  Axes with d == 0 never win.
G4GO_HD void box_exit(const Volume& v, V3 p, V3 d, float& dist, int& axis) {
    const float ax = fabsf(d.x), ay = fabsf(d.y), az = fabsf(d.z);
    const float nx = d.x > 0.0f ? fmaxf(v.hi[0] - p.x, 0.0f) : (d.x < 0.0f ? fmaxf(p.x - v.lo[0], 0.0f) : 1.0f);
    const float ny = d.y > 0.0f ? fmaxf(v.hi[1] - p.y, 0.0f) : (d.y < 0.0f ? fmaxf(p.y - v.lo[1], 0.0f) : 1.0f);
    const float nz = d.z > 0.0f ? fmaxf(v.hi[2] - p.z, 0.0f) : (d.z < 0.0f ? fmaxf(p.z - v.lo[2], 0.0f) : 1.0f);
    int best = 0;
    float bn = nx, ba = ax;
    if (ny * ba < bn * ay) {
        best = 1;
        bn = ny;
        ba = ay;
    }
    if (nz * ba < bn * az) {
        best = 2;
        bn = nz;
        ba = az;
    }
    axis = best;
    dist = bn / ba;
}

// Slab entry distance into box b from outside; returns kInf on a miss or when
// the entry is closer than tmin.  inv = 1/d per axis (inf for d == 0).
G4GO_HD float box_entry(const float* lo, const float* hi, V3 p, V3 inv, float tmin, int& axis) {
    const float pc[3] = {p.x, p.y, p.z};
    const float ic[3] = {inv.x, inv.y, inv.z};
    float t_near = -kInf, t_far = kInf;
    int a = 0;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        if (finitef(ic[k])) {
            const float t1 = (lo[k] - pc[k]) * ic[k];
            const float t2 = (hi[k] - pc[k]) * ic[k];
            const float tn = fminf(t1, t2);
            const float tf = fmaxf(t1, t2);
            if (tn > t_near) {
                t_near = tn;
                a = k;
            }
            t_far = fminf(t_far, tf);
        } else if (pc[k] < lo[k] || pc[k] > hi[k]) {
            return kInf;
        }
    }
    if (t_near >= tmin && t_near <= t_far) {
        axis = a;
        return t_near;
    }
    return kInf;
}

// BVH node test: entry distance (may be negative when p is inside), kInf on a
// miss or when the node lies entirely before tmin.
G4GO_HD float node_entry(const float* lo, const float* hi, V3 p, V3 inv, float tmin) {
    const float pc[3] = {p.x, p.y, p.z};
    const float ic[3] = {inv.x, inv.y, inv.z};
    float t_near = -kInf, t_far = kInf;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        if (finitef(ic[k])) {
            const float t1 = (lo[k] - pc[k]) * ic[k];
            const float t2 = (hi[k] - pc[k]) * ic[k];
            t_near = fmaxf(t_near, fminf(t1, t2));
            t_far = fminf(t_far, fmaxf(t1, t2));
        } else if (pc[k] < lo[k] || pc[k] > hi[k]) {
            return kInf;
        }
    }
    return (t_near <= t_far && t_far >= tmin) ? t_near : kInf;
}

G4GO_HD bool box_contains(const Volume& v, V3 q) {
    return q.x >= v.lo[0] && q.x <= v.hi[0] && q.y >= v.lo[1] && q.y <= v.hi[1] && q.z >= v.lo[2] && q.z <= v.hi[2];
}

// ---------------------------------------------------------------------------
// Geometry: closed convex mesh solids, held as the planes of their faces.  A
// ray from inside leaves through the plane with the smallest exit distance
// among the planes it moves towards (n.d > 0); a ray from outside enters at
// the largest entry distance among the planes it moves against (n.d < 0) when
// that stays below the smallest exit distance (Cyrus-Beck clipping).  For a
// closed convex mesh both are the closest triangle hit of the reference's
// trace (OptiXTransportDevice.cu:105-196, 728-741) up to rounding at edges.
// Distances are compared as cross products, with one division per call.
// ---------------------------------------------------------------------------
G4GO_HD float4 load_plane(const Plane& pl) {
    float4 q;
    __builtin_memcpy(&q, &pl, sizeof(q));
    return q;
}

// Exit plane and distance from p inside the solid; plane = kInvalid and
// dist = kInf when no plane faces away from d.
G4GO_HD void mesh_exit(const Plane* planes, uint32_t begin, uint32_t count, V3 p, V3 d, float& dist,
                       uint32_t& plane) {
    float bn = 1.0f, ba = 0.0f;  // best distance bn / ba; ba = 0 until a plane qualifies
    uint32_t best = kInvalid;
    for (uint32_t k = begin; k < begin + count; ++k) {
        const float4 q = load_plane(planes[k]);
        const V3 n = {q.x, q.y, q.z};
        const float nd = dot(n, d);
        if (nd > 0.0f) {
            const float num = fmaxf(q.w - dot(n, p), 0.0f);
            if (num * ba < bn * nd) {
                best = k;
                bn = num;
                ba = nd;
            }
        }
    }
    plane = best;
    dist = best == kInvalid ? kInf : bn / ba;
}

// Entry distance into the solid from p outside it, kInf on a miss or when the
// entry is closer than tmin.
template <bool kBatch>
G4GO_HD float mesh_entry(const Plane* planes, uint32_t begin, uint32_t count, V3 p, V3 d, float tmin,
                         uint32_t& plane) {
    float a = -1.0f, b = 0.0f;  // entry a / b over n.d < 0 (b = -n.d > 0)
    float c = 1.0f, e = 0.0f;   // exit c / e over n.d > 0
    uint32_t best = kInvalid;
    bool outside = false;  // parallel to a face, outside its plane: a miss
    const auto test = [&](const float4 q, uint32_t k) {
        const V3 n = {q.x, q.y, q.z};
        const float nd = dot(n, d);
        const float num = q.w - dot(n, p);
        if (nd < 0.0f) {
            if (-num * b > a * -nd) {
                best = k;
                a = -num;
                b = -nd;
            }
        } else if (nd > 0.0f) {
            if (num * e < c * nd) {
                c = num;
                e = nd;
            }
        } else {
            outside = outside || num < 0.0f;
        }
    };
    // The planes in order.  kBatch (a search spread over a wave, where the
    // latency of each load is the wave's): three loads issued together, and
    // no early exit.  Else one by one, leaving at the first outside parallel
    // face (three at a time made the single-lane steps of a mesh crystal 20%
    // slower).
    const uint32_t end = begin + count;
    uint32_t k = begin;
    if constexpr (kBatch) {
        for (; k + 2u < end; k += 3u) {
            const float4 q0 = load_plane(planes[k]), q1 = load_plane(planes[k + 1u]), q2 = load_plane(planes[k + 2u]);
            test(q0, k);
            test(q1, k + 1u);
            test(q2, k + 2u);
        }
        for (; k < end; ++k) {
            test(load_plane(planes[k]), k);
        }
    } else {
        for (; k < end; ++k) {
            test(load_plane(planes[k]), k);
            if (outside) {
                return kInf;
            }
        }
    }
    if (best == kInvalid || outside) {
        return kInf;
    }
    const float t_near = a / b;
    if (!(t_near >= tmin) || (e != 0.0f && t_near > c / e)) {
        return kInf;
    }
    plane = best;
    return t_near;
}

G4GO_HD bool mesh_contains(const Plane* planes, uint32_t begin, uint32_t count, V3 q) {
    for (uint32_t k = begin; k < begin + count; ++k) {
        const float4 pl = load_plane(planes[k]);
        if (dot(V3{pl.x, pl.y, pl.z}, q) > pl.w) {
            return false;
        }
    }
    return true;
}

// ---------------------------------------------------------------------------
// Geometry: closed non-convex mesh solids, held as their triangles (wound so
// that (v1 - v0) x (v2 - v0) points outwards) under a BVH per solid.  Hits use
// the watertight ray/triangle test of Woop, Benthin and Wald (Journal of
// Computer Graphics Techniques 2(1), 2013): the ray is sheared onto its
// dominant axis, the triangle's three edge functions are evaluated in 2D (in
// FP64 when one is exactly zero), so a ray through a shared edge or vertex
// never passes between the triangles there.  With outward winding the sign of
// the determinant separates leaving hits (det < 0) from entering ones
// (det > 0).  A point lies inside a solid when the first triangle along
// kContainDir is a leaving one.
// ---------------------------------------------------------------------------
struct WtRay {
    V3 o;
    int kx, ky, kz;
    float sx, sy, sz;
};

G4GO_HD WtRay wt_ray(V3 o, V3 d) {
    const float ax = fabsf(d.x), ay = fabsf(d.y), az = fabsf(d.z);
    WtRay r;
    r.o = o;
    r.kz = (ax >= ay && ax >= az) ? 0 : (ay >= az ? 1 : 2);
    r.kx = r.kz == 2 ? 0 : r.kz + 1;
    r.ky = r.kx == 2 ? 0 : r.kx + 1;
    const float dz = axis_of(d, r.kz);
    if (dz < 0.0f) {  // keep the winding
        const int k = r.kx;
        r.kx = r.ky;
        r.ky = k;
    }
    r.sx = axis_of(d, r.kx) / dz;
    r.sy = axis_of(d, r.ky) / dz;
    r.sz = 1.0f / dz;
    return r;
}

// Hit distance of the ray on triangle tr and whether it leaves the solid;
// false on a miss.
G4GO_HD bool wt_hit(const WtRay& r, const Tri& tr, float& t, bool& leaving) {
    const V3 a = sub(V3{tr.v0[0], tr.v0[1], tr.v0[2]}, r.o);
    const V3 b = sub(V3{tr.v1[0], tr.v1[1], tr.v1[2]}, r.o);
    const V3 c = sub(V3{tr.v2[0], tr.v2[1], tr.v2[2]}, r.o);
    const float az = axis_of(a, r.kz), bz = axis_of(b, r.kz), cz = axis_of(c, r.kz);
    const float ax = axis_of(a, r.kx) - r.sx * az, ay = axis_of(a, r.ky) - r.sy * az;
    const float bx = axis_of(b, r.kx) - r.sx * bz, by = axis_of(b, r.ky) - r.sy * bz;
    const float cx = axis_of(c, r.kx) - r.sx * cz, cy = axis_of(c, r.ky) - r.sy * cz;
    float u = cx * by - cy * bx;
    float v = ax * cy - ay * cx;
    float w = bx * ay - by * ax;
    if (u == 0.0f || v == 0.0f || w == 0.0f) {
        u = static_cast<float>(static_cast<double>(cx) * by - static_cast<double>(cy) * bx);
        v = static_cast<float>(static_cast<double>(ax) * cy - static_cast<double>(ay) * cx);
        w = static_cast<float>(static_cast<double>(bx) * ay - static_cast<double>(by) * ax);
    }
    if ((u < 0.0f || v < 0.0f || w < 0.0f) && (u > 0.0f || v > 0.0f || w > 0.0f)) {
        return false;
    }
    const float det = u + v + w;
    if (det == 0.0f) {
        return false;
    }
    t = (u * (r.sz * az) + v * (r.sz * bz) + w * (r.sz * cz)) / det;
    leaving = det < 0.0f;
    return true;
}

enum TriMode : int { kTriLeaving = 0, kTriEntering = 1, kTriAny = 2 };

// Closest hit so far of a triangle search.
struct TriBest {
    float t;
    uint32_t index, key;
    bool leaving;
};

G4GO_HD void tri_test(const Tri* tris, const WtRay& ray, uint32_t i, float tmin, int mode, TriBest& b) {
    float t;
    bool leave;
    const Tri& tr = tris[i];
    if (wt_hit(ray, tr, t, leave) && t >= tmin && (mode == kTriAny || leave == (mode == kTriLeaving)) &&
        (t < b.t || (t == b.t && tr.key < b.key))) {
        b.t = t;
        b.index = i;
        b.key = tr.key;
        b.leaving = leave;
    }
}

// Near-first walk of a triangle solid's BVH, as for the daughter BVH: nodes
// beyond the best hit so far are skipped.  Out of line: inlined at the call
// sites its traversal stack costs the triangle-level kernels more in spills
// than the calls do.
static __host__ __device__ __attribute__((noinline)) void tri_walk(const SceneView& sc, uint32_t root, V3 p, V3 d,
                                                                  const WtRay& ray, float tmin, int mode, TriBest& b) {
    const Tri* tris = scene_tris(sc);
    const V3 inv = {1.0f / d.x, 1.0f / d.y, 1.0f / d.z};
    uint32_t stack[kBvhMaxDepth + 1];
    float stack_t[kBvhMaxDepth + 1];
    int top = 0;
    {
        const BvhNode& n0 = sc.bvh[root];
        const float t = node_entry(n0.lo, n0.hi, p, inv, tmin);
        if (t < kInf) {
            stack[0] = root;
            stack_t[0] = t;
            top = 1;
        }
    }
    while (top > 0) {
        --top;
        if (!(stack_t[top] <= b.t)) {
            continue;
        }
        const BvhNode& node = sc.bvh[stack[top]];
        if ((node.right_or_count & kBvhLeaf) != 0) {
            const uint32_t n = node.right_or_count & ~kBvhLeaf;
            for (uint32_t k = 0; k < n; ++k) {
                tri_test(tris, ray, node.left + k, tmin, mode, b);
            }
        } else {
            const uint32_t l = node.left, rr = node.right_or_count;
            const BvhNode& nl = sc.bvh[l];
            const BvhNode& nr = sc.bvh[rr];
            const float tl = node_entry(nl.lo, nl.hi, p, inv, tmin);
            const float tr = node_entry(nr.lo, nr.hi, p, inv, tmin);
            const bool left_near = tl <= tr;
            const float t_far = left_near ? tr : tl;
            const float t_near = left_near ? tl : tr;
            if (t_far < kInf && t_far <= b.t && top <= kBvhMaxDepth) {
                stack[top] = left_near ? rr : l;
                stack_t[top] = t_far;
                ++top;
            }
            if (t_near < kInf && t_near <= b.t && top <= kBvhMaxDepth) {
                stack[top] = left_near ? l : rr;
                stack_t[top] = t_near;
                ++top;
            }
        }
    }
}

// Closest hit at t >= tmin on the triangles of triangle solid `vol` with the
// orientation `mode` asks for: distance (kInf on none), face (kFaceTri | the
// triangle's index, kInvalid on none) and whether it leaves the solid.  Ties
// go to the lower key.  A solid without a BVH (at most 64 triangles) is
// tested in a loop, inline.
G4GO_HD float tri_closest(const SceneView& sc, const Volume& vol, V3 p, V3 d, float tmin, int mode, uint32_t& face,
                          bool& leaving) {
    const Tri* tris = scene_tris(sc);
    const WtRay ray = wt_ray(p, d);
    TriBest b{kInf, kInvalid, kInvalid, false};
    const uint32_t root = vol.faces[0].neighbor;
    if (root == kInvalid) {
        const uint32_t begin = vol.plane_begin, count = vol.plane_count & kSolidCountMask;
        for (uint32_t i = begin; i < begin + count; ++i) {
            tri_test(tris, ray, i, tmin, mode, b);
        }
    } else {
        tri_walk(sc, root, p, d, ray, tmin, mode, b);
    }
    face = b.index == kInvalid ? kInvalid : (b.index | kFaceTri);
    leaving = b.leaving;
    return b.t;
}

// Closest hit at t >= tmin on the triangles of an open triangle solid, from
// either side.  Out of line, so that the closed solids' searches (inlined,
// with their orientation fixed) keep their code and registers.
static __host__ __device__ __attribute__((noinline)) float tri_closest_open(const SceneView& sc, const Volume& vol,
                                                                         V3 p, V3 d, float tmin, uint32_t& face) {
    bool leaving;
    return tri_closest(sc, vol, p, d, tmin, kTriAny, face, leaving);
}

// Fallback exit from a triangle solid when no triangle ahead is a leaving
// hit: among the triangles the ray moves away from (n.d > 0), the one whose
// plane lies nearest the point (ties: lower key); the distance to that plane
// (0 when the point is beyond it).  kInf when none.
static __host__ __device__ __attribute__((noinline)) float tri_nearest_leaving(const SceneView& sc, const Volume& vol,
                                                                            V3 p, V3 d, uint32_t& face) {
    const Tri* tris = scene_tris(sc);
    const uint32_t begin = vol.plane_begin, count = vol.plane_count & kSolidCountMask;
    float best_gap = kInf, best_num = 0.0f, best_nd = 1.0f;
    uint32_t best_i = kInvalid, best_key = kInvalid;
    for (uint32_t i = begin; i < begin + count; ++i) {
        const Plane& f = tris[i].face;
        const V3 n = {f.n[0], f.n[1], f.n[2]};
        const float nd = dot(n, d);
        if (nd > 0.0f) {
            const float num = f.h - dot(n, p);
            const float gap = fabsf(num);
            if (gap < best_gap || (gap == best_gap && tris[i].key < best_key)) {
                best_gap = gap;
                best_num = num;
                best_nd = nd;
                best_i = i;
                best_key = tris[i].key;
            }
        }
    }
    face = best_i == kInvalid ? kInvalid : (best_i | kFaceTri);
    return best_i == kInvalid ? kInf : fmaxf(best_num, 0.0f) / best_nd;
}

// A direction that meets no axis-aligned edge or face of a mesh edge-on:
// (1, 2, 3) / sqrt(14).
constexpr float kContainDir[3] = {0.26726124f, 0.53452248f, 0.80178373f};

G4GO_HD bool tri_contains(const SceneView& sc, const Volume& vol, V3 q) {
    uint32_t face;
    bool leave;
    const float t = tri_closest(sc, vol, q, V3{kContainDir[0], kContainDir[1], kContainDir[2]}, 0.0f, kTriAny, face,
                                leave);
    return t < kInf && leave;
}

// Geometry levels of the step code: kGeomBoxes (box solids only),
// kGeomConvex (also convex mesh solids), kGeomTriangles (also closed
// triangle solids), kGeomOpen (also open triangle solids).  A scene runs the
// lowest level that holds it, so the closed-solid kernels carry no code of
// open solids.
enum GeomLevel : int { kGeomBoxes = 0, kGeomConvex = 1, kGeomTriangles = 2, kGeomOpen = 3 };

// Point inside a mesh solid (convex, or a triangle solid at kGeomTriangles
// and kGeomOpen; an open solid holds no point).
template <int kGeom>
G4GO_HD bool solid_contains(const SceneView& sc, const Volume& v, V3 q) {
    if (kGeom >= 2 && (v.plane_count & kSolidTriangles) != 0u) {
        return !(kGeom == kGeomOpen && (v.plane_count & kSolidOpen) != 0u) && tri_contains(sc, v, q);
    }
    return mesh_contains(sc.planes, v.plane_begin, v.plane_count, q);
}

// The face record of a mesh face reference (a plane or a triangle).
template <int kGeom>
G4GO_HD const Plane& face_record(const SceneView& sc, uint32_t face) {
    return kGeom >= 2 && (face & kFaceTri) != 0u ? scene_tris(sc)[face & ~kFaceTri].face : sc.planes[face];
}

// Border binding lookup with skin precedence (OptiXTransportDevice.cu:218-252).
G4GO_HD uint32_t find_surface(const SceneView& sc, uint32_t from, uint32_t to) {
    uint32_t first = 0, last = sc.n_bindings;
    while (first < last) {
        const uint32_t mid = first + (last - first) / 2;
        const Binding& b = sc.bindings[mid];
        if (b.from < from || (b.from == from && b.to < to)) {
            first = mid + 1;
        } else {
            last = mid;
        }
    }
    if (first < sc.n_bindings && sc.bindings[first].from == from && sc.bindings[first].to == to) {
        return sc.bindings[first].surface;
    }
    if (from >= sc.n_volumes || to >= sc.n_volumes) {
        return kInvalid;
    }
    const Volume& f = sc.volumes[from];
    const Volume& t = sc.volumes[to];
    if (t.parent == from) {
        return t.skin != kInvalid ? t.skin : f.skin;
    }
    return f.skin != kInvalid ? f.skin : t.skin;
}


// A daughter box read with two 16-byte loads (its fields one by one would be
// eight loads, each waited for where a test reads it).
G4GO_HD DaughterBox load_dbox(const DaughterBox& src) {
    DaughterBox b;
    __builtin_memcpy(&b, &src, sizeof(b));
    return b;
}

// Entry distance into one daughter: its box, or (kGeom >= kGeomConvex) its
// mesh when it has one (plane = the entered face, else kInvalid).
template <int kGeom, bool kBatch = false>
G4GO_HD float daughter_entry(const SceneView& sc, const DaughterBox& b, V3 p, V3 d, V3 inv, float tmin, int& axis,
                             uint32_t& plane) {
    if (kGeom != kGeomBoxes && b.mesh != 0u) {
        if (b.mesh != kMeshInVolume) {
            return mesh_entry<kBatch>(sc.planes, b.mesh >> kMeshBeginShift, b.mesh & kMeshInVolume, p, d, tmin, plane);
        }
        const Volume& v = sc.volumes[b.volume];
        if (kGeom >= kGeomTriangles && (v.plane_count & kSolidTriangles) != 0u) {
            if (kGeom == kGeomOpen && (v.plane_count & kSolidOpen) != 0u) {
                return tri_closest_open(sc, v, p, d, tmin, plane);
            }
            bool leave;
            return tri_closest(sc, v, p, d, tmin, kTriEntering, plane, leave);
        }
        return mesh_entry<kBatch>(sc.planes, v.plane_begin, v.plane_count, p, d, tmin, plane);
    }
    plane = kInvalid;
    return box_entry(b.lo, b.hi, p, inv, tmin, axis);
}

// In scenes with convex mesh solids (kGeomConvex) a volume with at most
// kScanDaughters daughters is searched by testing every daughter, BVH or not:
// the smallest entry distance, ties to the lower volume index.  The
// transport kernels split that test over the lanes of a wave for the few
// lanes that need it (wave_daughter_search); every other caller runs it per
// lane.  Both give the same answer bit for bit, whatever order the daughters
// are tested in.  Larger volumes, and volumes of the other geometry levels
// with more than eight daughters, take the BVH walk everywhere.
constexpr uint32_t kScanDaughters = 128;

G4GO_HD bool scans_daughters(int geom, const Volume& vol) {
    return vol.bvh_root == kInvalid || (geom == kGeomConvex && vol.daughter_count <= kScanDaughters);
}

// A daughter search's answer: entry distance (kInf: none), the daughter
// entered (kInvalid: none) and the box axis or mesh face it enters through.
// which == kDaughterUnsearched: not searched yet.
constexpr uint32_t kDaughterUnsearched = 0xfffffffdu;
struct DaughterHit {
    float t;
    uint32_t which;
    int axis;
    uint32_t plane;
};

// Nearest daughter entry of volume `vol` (BVH walk when the volume has one and
// scans_daughters is false).  `plane` (mesh levels only) is the entered mesh
// face, kInvalid for a box face.
template <int kGeom>
G4GO_HD float nearest_daughter(const SceneView& sc, const Volume& vol, V3 p, V3 d, float tmin, uint32_t& which,
                               int& axis, uint32_t& plane) {
    const V3 inv = {1.0f / d.x, 1.0f / d.y, 1.0f / d.z};
    float best = kInf;
    which = kInvalid;
    if (scans_daughters(kGeom, vol)) {
        for (uint32_t k = 0; k < vol.daughter_count; ++k) {
            const DaughterBox b = load_dbox(sc.dbox[vol.daughter_begin + k]);
            const uint32_t j = b.volume;
            int a = 0;
            uint32_t f = kInvalid;
            const float t = daughter_entry<kGeom>(sc, b, p, d, inv, tmin, a, f);
            if (t < best || (t == best && t < kInf && j < which)) {  // ties: lower volume index
                best = t;
                which = j;
                axis = a;
                plane = f;
            }
        }
        return best;
    }
    // Near-first traversal: the nearer child is visited first and every node
    // whose entry distance exceeds the best hit so far is skipped, so the walk
    // usually ends after the leaves around the answer.  The stack holds at most
    // one entry per level (the farther child) plus one; the host limits the
    // tree depth to kBvhMaxDepth.
    uint32_t stack[kBvhMaxDepth + 1];
    float stack_t[kBvhMaxDepth + 1];
    int top = 0;
    {
        const BvhNode& root = sc.bvh[vol.bvh_root];
        const float t = node_entry(root.lo, root.hi, p, inv, tmin);
        if (t < kInf) {
            stack[0] = vol.bvh_root;
            stack_t[0] = t;
            top = 1;
        }
    }
    while (top > 0) {
        --top;
        if (!(stack_t[top] <= best)) {
            continue;
        }
        const BvhNode& node = sc.bvh[stack[top]];
        if ((node.right_or_count & kBvhLeaf) != 0) {
            const uint32_t count = node.right_or_count & ~kBvhLeaf;
            for (uint32_t k = 0; k < count; ++k) {
                const DaughterBox b = load_dbox(sc.dbox[node.left + k]);
                const uint32_t j = b.volume;
                int a = 0;
                uint32_t f = kInvalid;
                const float t = daughter_entry<kGeom>(sc, b, p, d, inv, tmin, a, f);
                if (t < best || (t == best && t < kInf && j < which)) {  // ties: lower volume index
                    best = t;
                    which = j;
                    axis = a;
                    plane = f;
                }
            }
        } else {
            const uint32_t l = node.left, r = node.right_or_count;
            const BvhNode& nl = sc.bvh[l];
            const BvhNode& nr = sc.bvh[r];
            const float tl = node_entry(nl.lo, nl.hi, p, inv, tmin);
            const float tr = node_entry(nr.lo, nr.hi, p, inv, tmin);
            const bool left_near = tl <= tr;
            const float t_far = left_near ? tr : tl;
            const float t_near = left_near ? tl : tr;
            if (t_far < kInf && t_far <= best && top <= kBvhMaxDepth) {
                stack[top] = left_near ? r : l;
                stack_t[top] = t_far;
                ++top;
            }
            if (t_near < kInf && t_near <= best && top <= kBvhMaxDepth) {
                stack[top] = left_near ? l : r;
                stack_t[top] = t_near;
                ++top;
            }
        }
    }
    return best;
}

// Neighbour across a face found at run time: the first sibling solid that
// holds the point just outside the face, else the parent volume.
template <int kGeom>
G4GO_HD uint32_t point_neighbor(const SceneView& sc, uint32_t current, V3 q) {
    const Volume& cur = sc.volumes[current];
    if (cur.parent == kInvalid) {
        return kInvalid;
    }
    const Volume& par = sc.volumes[cur.parent];
    for (uint32_t k = 0; k < par.daughter_count; ++k) {
        const DaughterBox b = load_dbox(sc.dbox[par.daughter_begin + k]);
        if (b.volume != current && q.x >= b.lo[0] && q.x <= b.hi[0] && q.y >= b.lo[1] && q.y <= b.hi[1] &&
            q.z >= b.lo[2] && q.z <= b.hi[2]) {
            if (kGeom != kGeomBoxes && b.mesh != 0u && !solid_contains<kGeom>(sc, sc.volumes[b.volume], q)) {
                continue;
            }
            return b.volume;
        }
    }
    return cur.parent;
}

// point_neighbor over a face's listed siblings (daughter-list slots in
// ascending order, the current volume not among them): every sibling left out
// has a box that cannot hold q, so the first listed one that holds it is the
// first of all.
template <int kGeom>
G4GO_HD uint32_t listed_neighbor(const SceneView& sc, const uint32_t* slots, uint32_t count, uint32_t parent, V3 q) {
    for (uint32_t k = 0; k < count; ++k) {
        const DaughterBox b = load_dbox(sc.dbox[slots[k]]);
        if (q.x >= b.lo[0] && q.x <= b.hi[0] && q.y >= b.lo[1] && q.y <= b.hi[1] && q.z >= b.lo[2] &&
            q.z <= b.hi[2]) {
            if (kGeom != kGeomBoxes && b.mesh != 0u && !solid_contains<kGeom>(sc, sc.volumes[b.volume], q)) {
                continue;
            }
            return b.volume;
        }
    }
    return parent;
}

// ---------------------------------------------------------------------------
// Ground-surface sampling (OptiXTransportDevice.cu:492-595).  Each rejection
// attempt consumes one Philox block of its own, so a rejected glisur smear is
// redrawn (the reference re-evaluates the same draws inside its inner loop).
// ---------------------------------------------------------------------------
G4GO_HD V3 sample_facet_normal(const Surface& s, V3 n, V3 incoming, const Rng& rng, const Photon& ph) {
    n = normalize(n);
    if (s.model == kGlisur) {
        const float polish = clamp01(s.model_value);
        if (polish >= 1.0f) {
            return n;
        }
        for (uint32_t attempt = 0; attempt < 128u; ++attempt) {
            const U4 w = rng.block(ph.photon_id, ph.event_id, ph.step, kProcFacet, attempt);
            const V3 smear = {2.0f * u01(w.x) - 1.0f, 2.0f * u01(w.y) - 1.0f, 2.0f * u01(w.z) - 1.0f};
            if (dot(smear, smear) > 1.0f) {
                continue;
            }
            const V3 facet = normalize(add(n, scale(smear, 1.0f - polish)));
            if (dot(incoming, facet) < 0.0f) {
                return facet;
            }
        }
        return n;
    }
    if (s.model != kUnified) {
        return n;
    }
    const float sigma = fmaxf(s.model_value, 0.0f);
    if (!(sigma > 0.0f)) {
        return n;
    }
    const V3 helper = fabsf(n.z) < 0.9f ? V3{0.0f, 0.0f, 1.0f} : V3{1.0f, 0.0f, 0.0f};
    const V3 tangent = normalize(cross(n, helper));
    const V3 bitangent = normalize(cross(n, tangent));
    const float sigma2 = sigma * sigma;
    for (uint32_t attempt = 0; attempt < 256u; ++attempt) {
        const U4 w = rng.block(ph.photon_id, ph.event_id, ph.step, kProcFacet, attempt);
        const float quarter = 0.25f * u01(w.x);  // alpha = (pi/2) u in turns
        const float alpha = 1.57079632679489662f * u01(w.x);
        float sa, ca;
        g4_sincos_turn(quarter, sa, ca);
        const float gaussian = g4_expf(-0.5f * alpha * alpha / sigma2);
        if (u01(w.y) > gaussian * sa) {
            continue;
        }
        float sp, cp;
        g4_sincos_turn(u01(w.z), sp, cp);
        const V3 facet = normalize(add(scale(n, ca), add(scale(tangent, sa * cp), scale(bitangent, sa * sp))));
        if (dot(incoming, facet) < 0.0f) {
            return facet;
        }
    }
    return n;
}

G4GO_HD V3 sample_unified_reflection(const SceneView& sc, const Surface& s, V3 dir, V3 n, const Rng& rng,
                                     const Photon& ph) {
    const bool ground = s.finish == kGround;
    const V3 facet = sample_facet_normal(s, n, dir, rng, ph);
    const float spike = clamp01(prop_eval_ph(sc.pool, s.spike, ph, ground ? 0.0f : 1.0f));
    const float lobe = clamp01(prop_eval_ph(sc.pool, s.lobe, ph, 0.0f));
    const float back = clamp01(prop_eval_ph(sc.pool, s.backscatter, ph, 0.0f));
    const U4 w = rng.block(ph.photon_id, ph.event_id, ph.step, kProcLobe, 0);
    const float u = u01(w.x);
    if (u < spike) {
        return reflect_direction(dir, n);
    }
    if (u < spike + lobe) {
        return reflect_direction(dir, facet);
    }
    if (u < spike + lobe + back) {
        return scale(dir, -1.0f);
    }
    return sample_lambertian(n, u01(w.y), u01(w.z));
}

G4GO_HD V3 sample_surface_reflection(const SceneView& sc, const Surface& s, V3 dir, V3 n, const Rng& rng,
                                     const Photon& ph) {
    if (s.finish == kPolished) {
        return reflect_direction(dir, n);
    }
    if (s.model == kGlisur) {
        return reflect_direction(dir, sample_facet_normal(s, n, dir, rng, ph));
    }
    return sample_unified_reflection(sc, s, dir, n, rng, ph);
}

// Rayleigh scattering with the dipole distribution relative to the
// polarisation (Geant4 11 G4OpRayleigh::PostStepDoIt): the new direction k'
// has density proportional to 1 - (p.k')^2 about the old polarisation p, and
// the new polarisation is the unit projection of p perpendicular to k', with
// a random sign.  Geant4 reaches this by rejection (uniform k', accepted with
// probability (p'.p)^2 = 1 - (p.k')^2); here it is sampled directly from one
// Philox block, so a wave never loops on its slowest lane:
//   * c = p.k' is the median of three uniforms on (-1, 1], whose density is
//     3/4 (1 - c^2).  u01 is monotone, so that median is u01 of the median of
//     the words w0, w1, w2;
//   * the azimuth of k' about p is uniform (w3);
//   * the sign of the new polarisation is the parity of the order of w0, w1,
//     w2: a uniformly random permutation, independent of its median.
G4GO_HD void rayleigh_scatter(Photon& ph, const Rng& rng) {
    const V3 old_pol = ph.pol;
    const U4 w = rng.block(ph.photon_id, ph.event_id, ph.step, kProcRayleigh, 0);
    const uint32_t lo = min(w.x, w.y), hi = max(w.x, w.y);
    const uint32_t med = max(lo, min(hi, w.z));
    const bool flip = ((w.x < w.y) != (w.y < w.z)) != (w.x < w.z);
    const float c = 2.0f * u01(med) - 1.0f;
    const float s = sqrtf(fmaxf(0.0f, 1.0f - c * c));
    float sp, cp;
    g4_sincos_turn(u01(w.w), sp, cp);
    const V3 new_dir = rotate_uz(V3{s * cp, s * sp, c}, old_pol);
    V3 new_pol = sub(old_pol, scale(new_dir, dot(new_dir, old_pol)));
    const float l2 = dot(new_pol, new_pol);
    if (!(l2 > 0.0f)) {
        float sq, cq;
        g4_sincos_turn(static_cast<float>(w.w & 0x00ffffffu) * 0x1p-24f, sq, cq);
        new_pol = rotate_uz(V3{cq, sq, 0.0f}, new_dir);
    } else {
        new_pol = scale(new_pol, 1.0f / sqrtf(l2));
        if (flip) {
            new_pol = scale(new_pol, -1.0f);
        }
    }
    ph.dir = new_dir;
    ph.pol = new_pol;
}

// ---------------------------------------------------------------------------
// Photon generation.
// ---------------------------------------------------------------------------

// Random linear polarisation perpendicular to (sin t cos p, sin t sin p, cos t):
// cos(psi) e_theta + sin(psi) e_phi.
G4GO_HD void isotropic(float cos_t, float sp, float cp, float spsi, float cpsi, V3& dir, V3& pol) {
    const float sin_t = sqrtf(fmaxf(0.0f, 1.0f - cos_t * cos_t));
    dir = V3{sin_t * cp, sin_t * sp, cos_t};
    const V3 e_theta = {cos_t * cp, cos_t * sp, -sin_t};
    const V3 e_phi = {-sp, cp, 0.0f};
    pol = add(scale(e_theta, cpsi), scale(e_phi, spsi));
}

G4GO_HD void generate_synthetic(Photon& ph, const SyntheticSource& src, const Rng& rng, uint32_t event_id,
                                uint32_t photon_id) {
    ph.photon_id = photon_id;
    ph.event_id = event_id;
    ph.step = 0;
    ph.volume = src.volume;
    ph.material = kInvalid;
    ph.grid_i = kInvalid;
    const U4 a = rng.block(photon_id, event_id, 0, kProcSynthetic, 0);
    const U4 b = rng.block(photon_id, event_id, 0, kProcSynthetic, 1);
    const float cos_t = 1.0f - 2.0f * u01(a.x);
    float sp, cp, spsi, cpsi, sg, cg;
    g4_sincos_turn(u01(a.y), sp, cp);
    g4_sincos_turn(u01(a.z), spsi, cpsi);
    isotropic(cos_t, sp, cp, spsi, cpsi, ph.dir, ph.pol);
    // Box-Muller wavelength
    const float radius = sqrtf(-2.0f * g4_logf(u01(a.w)));
    g4_sincos_turn(u01(b.x), sg, cg);
    const float wavelength = src.wavelength_mean_nm + src.wavelength_sigma_nm * (radius * cg);
    ph.energy = kHcEvNm / wavelength;
    ph.pos = V3{src.lo[0] + (src.hi[0] - src.lo[0]) * u01(b.y), src.lo[1] + (src.hi[1] - src.lo[1]) * u01(b.z),
                src.lo[2] + (src.hi[2] - src.lo[2]) * u01(b.w)};
    const U4 c = rng.block(photon_id, event_id, 0, kProcSynthetic, 2);
    ph.time = static_cast<double>(-src.decay_ns * g4_logf(u01(c.x)));
}

// Expansion of one emission record into photon `local` of that record
// (OptiXTransportDevice.cu:346-481).  Draws: block 0 {spectrum or
// Cerenkov energy, cos theta, phi, polarisation angle}, block 1
// {position fraction, decay}, further blocks for rejection loops.
G4GO_HD void generate_emission(Photon& ph, const Emission& em, const SceneView& sc, const Rng& rng, uint32_t local) {
    ph.photon_id = em.first_photon_id + local;
    ph.event_id = em.event_id;
    ph.step = 0;
    ph.volume = em.volume;
    ph.material = kInvalid;
    ph.grid_i = kInvalid;
    ph.pos = V3{em.pos[0], em.pos[1], em.pos[2]};
    ph.time = static_cast<double>(em.time);
    ph.dir = normalize(V3{em.dir[0], em.dir[1], em.dir[2]});
    ph.pol = V3{em.pol[0], em.pol[1], em.pol[2]};
    ph.energy = em.energy;
    if (em.type == 0) {
        return;
    }
    if (em.material >= sc.n_materials) {
        ph.energy = u2f(0x7fc00000u);
        return;
    }
    const Material& mat = sc.materials[em.material];
    const uint32_t key_lo = ph.photon_id, key_hi = ph.event_id;
    float frac = 1.0f;
    uint32_t blk = 2;
    if (em.type == 1) {
        // Cerenkov: energy by rejection on sin^2(theta), then the position along
        // the step by rejection on the linearly varying mean photon count.
        const float beta = (em.pre_velocity + 0.5f * em.delta_velocity) / kSpeedOfLightMmPerNs;
        if (!(beta > 0.0f) || !(beta < 1.0f) || mat.rindex.count == 0) {
            ph.energy = u2f(0x7fc00000u);
            return;
        }
        const float* en = sc.pool + mat.rindex.offset;
        const float e_min = en[0];
        const float e_max = en[mat.rindex.count - 1];
        const float n_max = mat.rindex_max;
        const float beta_inv = 1.0f / beta;
        const float max_cos = beta_inv / n_max;
        const float max_sin2 = 1.0f - max_cos * max_cos;
        const float de = e_max - e_min;
        if (!(n_max > 0.0f) || !finitef(n_max) || !(max_cos > 0.0f) || !(max_cos < 1.0f) || !(de >= 0.0f) ||
            !finitef(de) || !(max_sin2 > 0.0f)) {
            ph.energy = u2f(0x7fc00000u);
            return;
        }
        float energy = e_min, cos_t = 0.0f, sin2 = 0.0f;
        bool accepted = false;
        for (uint32_t attempt = 0; attempt < 1024u; ++attempt) {
            const U4 w = rng.block(key_lo, key_hi, 0, kProcEmission, 16u + attempt);
            energy = e_min + u01(w.x) * de;
            const float n = prop_eval(sc.pool, mat.rindex, energy, u2f(0x7fc00000u));
            cos_t = beta_inv / n;
            sin2 = 1.0f - cos_t * cos_t;
            if (finitef(n) && cos_t > 0.0f && cos_t < 1.0f && sin2 > 0.0f && u01(w.y) * max_sin2 <= sin2) {
                accepted = true;
                break;
            }
        }
        if (!accepted) {
            ph.energy = u2f(0x7fc00000u);
            return;
        }
        const U4 w0 = rng.block(key_lo, key_hi, 0, kProcEmission, 0);
        float sp, cp;
        g4_sincos_turn(u01(w0.z), sp, cp);
        const float sin_t = sqrtf(fmaxf(sin2, 0.0f));
        const V3 axis = V3{em.dir[0], em.dir[1], em.dir[2]};
        ph.dir = rotate_uz(V3{sin_t * cp, sin_t * sp, cos_t}, normalize(axis));
        ph.pol = rotate_uz(V3{cos_t * cp, cos_t * sp, -sin_t}, normalize(axis));
        ph.dir = normalize(ph.dir);
        ph.pol = normalize(ph.pol);
        const float max_mean = fmaxf(em.pre_mean, em.post_mean);
        const float d_mean = em.pre_mean - em.post_mean;
        for (uint32_t attempt = 0; attempt < 1024u; ++attempt) {
            const U4 w = rng.block(key_lo, key_hi, 0, kProcEmission, 2048u + attempt);
            frac = u01(w.x);
            if (!(u01(w.y) * max_mean > em.pre_mean - frac * d_mean)) {
                break;
            }
        }
        ph.energy = energy;
    } else if (em.type == 2) {
        if (em.spectrum >= 3u) {
            ph.energy = u2f(0x7fc00000u);
            return;
        }
        const U4 w0 = rng.block(key_lo, key_hi, 0, kProcEmission, 0);
        const U4 w1 = rng.block(key_lo, key_hi, 0, kProcEmission, 1);
        ph.energy = sample_spectrum(sc.pool, mat.spectrum[em.spectrum], u01(w0.x));
        const float cos_t = 1.0f - 2.0f * u01(w0.y);
        float sp, cp, spsi, cpsi;
        g4_sincos_turn(u01(w0.z), sp, cp);
        g4_sincos_turn(u01(w0.w), spsi, cpsi);
        isotropic(cos_t, sp, cp, spsi, cpsi, ph.dir, ph.pol);
        frac = em.charge == 0.0f ? 1.0f : u01(w1.x);
        const float decay = em.decay;
        if (decay > 0.0f && finitef(decay)) {
            if (!(em.rise > 0.0f) || !finitef(em.rise)) {
                ph.time += static_cast<double>(-decay * g4_logf(fmaxf(u01(w1.y), 1.0e-7f)));
            } else {
                float sample = 0.0f;
                for (uint32_t attempt = 0; attempt < 1024u; ++attempt) {
                    const U4 w = rng.block(key_lo, key_hi, 0, kProcEmission, blk + attempt);
                    sample = -decay * g4_logf(fmaxf(1.0e-7f, 1.0f - u01(w.x)));
                    if (!(u01(w.y) > 1.0f - g4_expf(-sample / em.rise))) {
                        break;
                    }
                }
                ph.time += static_cast<double>(sample);
            }
        }
    } else {
        ph.energy = u2f(0x7fc00000u);
        return;
    }
    ph.pos = add(V3{em.pos[0], em.pos[1], em.pos[2]}, scale(V3{em.step_delta[0], em.step_delta[1], em.step_delta[2]}, frac));
    const float velocity = em.pre_velocity + 0.5f * frac * em.delta_velocity;
    if (velocity > 0.0f && finitef(velocity)) {
        ph.time += static_cast<double>(frac * em.step_length / velocity);
    }
}

// Validity checks of a generated photon (OptiXTransportDevice.cu:693-702).
G4GO_HD bool start_photon(Photon& ph) {
    const float l2 = dot(ph.dir, ph.dir);
    if (!(ph.energy > 0.0f) || !finitef(ph.energy) || !(l2 > 0.0f) || !finitef(l2)) {
        return false;
    }
    ph.dir = normalize(ph.dir);
    ph.pol = project_polarization(ph.pol, ph.dir);
    return true;
}

// The first step's property setup, done when the photon is generated: its
// node on the hot grid and the bulk properties of its start volume's material
// (transport_step computes the same values when it finds them missing; a
// photon whose volume or material cannot transport is left to fail there).
G4GO_HD void prepare_start(Photon& ph, const SceneView& sc) {
    if (ph.volume >= sc.n_volumes) {
        return;
    }
    if (sc.hot_count != 0) {
        grid_locate(sc.pool + sc.hot_offset, sc.hot_count, ph.energy, ph.grid_i, ph.grid_f);
    }
    cache_material(ph, sc, sc.volumes[ph.volume].material);
}

// ---------------------------------------------------------------------------
// One transport step (one pass of the bounce loop, OptiXTransportDevice.cu:708-931).
// ---------------------------------------------------------------------------
struct StepConfig {
    float eps;  // boundary tolerance (mm)
};

// The lean box step: a photon in a box volume with its material cached and
// its node on the hot grid located, leaving through a face that
// Volume::box_step_faces marks (no daughters, a static neighbour, a polished
// dielectric_metal surface).  It is transport_step's arithmetic for that case
// in the same order, so it gives the same bits, without the tests and
// branches of every other case, which lengthen the dependent chain of a step
// and the register pressure of the loop.  Returns kStepFallback, having
// changed nothing, when the step is not such a step.
constexpr int kStepFallback = -2;

G4GO_HD int box_step(Photon& ph, const SceneView& sc, const Rng& rng, const StepConfig& cfg, bool& zero_step,
                     uint32_t& sensor_out) {
    if (ph.volume >= sc.n_volumes || (ph.grid_i == kInvalid && sc.hot_count != 0)) {
        return kStepFallback;
    }
    const Volume& vol = sc.volumes[ph.volume];
    float dist;
    int axis = 0;
    box_exit(vol, ph.pos, ph.dir, dist, axis);
    const bool up = axis_of(ph.dir, axis) > 0.0f;
    const int face = 2 * axis + (up ? 1 : 0);
    if (vol.material != ph.material || ((vol.box_step_faces >> face) & 1u) == 0u || !finitef(dist)) {
        return kStepFallback;
    }
    zero_step = dist <= cfg.eps;

    const U4 w = rng.block(ph.photon_id, ph.event_id, ph.step, kProcStep, 0);
    // bulk interaction (an infinite interaction length gives an infinite or
    // NaN distance, which never wins)
    const float d_int = -g4_logf(u01(w.x)) * ph.interaction_length;
    if (d_int < dist) {
        ph.time += static_cast<double>(d_int * ph.inv_velocity);
        if (u01(w.y) <= ph.absorb_fraction) {
            return kAbsorbedBulk;
        }
        ph.pos = add(ph.pos, scale(ph.dir, d_int));
        rayleigh_scatter(ph, rng);
        return kAlive;
    }
    ph.time += static_cast<double>(dist * ph.inv_velocity);
    // onto the face plane
    const V3 bpos = add(ph.pos, scale(ph.dir, dist));
    const float fplane = up ? (axis == 0 ? vol.hi[0] : (axis == 1 ? vol.hi[1] : vol.hi[2]))
                            : (axis == 0 ? vol.lo[0] : (axis == 1 ? vol.lo[1] : vol.lo[2]));
    ph.pos = V3{axis == 0 ? fplane : bpos.x, axis == 1 ? fplane : bpos.y, axis == 2 ? fplane : bpos.z};

    Face fr;
    __builtin_memcpy(&fr, &vol.faces[face], sizeof(Face));
    const Volume& nvol = sc.volumes[fr.neighbor];
    const Surface& srf = sc.surfaces[fr.surface];
    const float refl = prop_eval_ph(sc.pool, srf.reflectivity, ph, 0.0f);
    const float trans = prop_eval_ph(sc.pool, srf.transmittance, ph, 0.0f);
    const int outcome =
        classify_surface(refl, trans, srf.reflectivity.count != 0, srf.transmittance.count != 0, u01(w.z));
    if (outcome == kAbsorb) {
        const float eff = clamp01(prop_eval_ph(sc.pool, srf.efficiency, ph, 0.0f));
        if (evaluate_absorption(eff, nvol.sensor != kInvalid || vol.sensor != kInvalid, u01(w.w)) == kDetect) {
            sensor_out = nvol.sensor != kInvalid ? nvol.sensor : vol.sensor;
            return kDetected;
        }
        return kAbsorbedSurface;
    }
    if (outcome == kDirectTransmit) {
        ph.volume = fr.neighbor;
        return kAlive;
    }
    // specular reflection at the axis-aligned face
    ph.dir = negate_axis(ph.dir, axis);
    ph.pol = negate_off_axis(ph.pol, axis);
    return kAlive;
}

// kGeom: the geometry level (kGeomConvex: convex mesh solids, whose faces
// are found by clipping; kGeomTriangles: also triangle solids, whose faces
// are found by the watertight triangle test).  On mesh faces the photon is not snapped onto a mesh face,
// and a daughter entered up to eps beyond the current volume's exit wins (a
// daughter face coincident with the wall, which the reference settles by its
// surface rank re-trace, OptiXTransportDevice.cu:105-127, 742-777).
//
// `pre` (kGeomConvex) is the daughter search of this step when the caller has
// made it already (wave_daughter_search), else which == kDaughterUnsearched.
template <bool kTrace, int kGeom = kGeomBoxes>
G4GO_HD int transport_step(Photon& ph, const SceneView& sc, const Rng& rng, const StepConfig& cfg, bool& zero_step,
                           uint32_t& sensor_out, StepTrace* trace,
                           DaughterHit pre = DaughterHit{kInf, kDaughterUnsearched, 0, kInvalid}) {
    if constexpr (!kTrace && kGeom == kGeomBoxes) {
        const int lean = box_step(ph, sc, rng, cfg, zero_step, sensor_out);
        if (lean != kStepFallback) {
            return lean;
        }
    }
    zero_step = false;
    if (ph.volume >= sc.n_volumes) {
        return kInvalidState;
    }
    G4GO_REGION_BEGIN(kRegionGeometry);
    if (ph.grid_i == kInvalid && sc.hot_count != 0) {
        grid_locate(sc.pool + sc.hot_offset, sc.hot_count, ph.energy, ph.grid_i, ph.grid_f);
    }
    const Volume& vol = sc.volumes[ph.volume];
    if (vol.material != ph.material && !cache_material(ph, sc, vol.material)) {
        return kInvalidState;
    }
    if (kTrace) {
        trace->segment = 0.0f;
        trace->material = ph.material;
        trace->surface = kInvalid;
        trace->event = 0;
    }

    // nearest boundary: leave the solid, or enter a daughter; `plane` is the
    // mesh face crossed (mesh levels only), kInvalid for a box face
    float dist;
    int axis = 0;
    uint32_t plane = kInvalid;
    constexpr bool kMesh = kGeom != kGeomBoxes;
    bool open_solid = false;  // in an open triangle solid, with no triangle of it ahead
    if (kGeom >= kGeomTriangles && (vol.plane_count & kSolidTriangles) != 0u) {
        // A closed solid: its closest leaving hit.  An open one: any of its
        // triangles beyond the boundary tolerance, from either side (the
        // reference's trace from the photon, tmin = eps).
        const bool open = kGeom == kGeomOpen && (vol.plane_count & kSolidOpen) != 0u;
        if (open) {
            dist = tri_closest_open(sc, vol, ph.pos, ph.dir, cfg.eps, plane);
            open_solid = !finitef(dist);
        } else {
            bool leave;
            dist = tri_closest(sc, vol, ph.pos, ph.dir, 0.0f, kTriLeaving, plane, leave);
        }
        if (!open && !finitef(dist)) {
            // No leaving triangle ahead: rounding left the photon a hair
            // outside its solid at an edge (after a reflection there), moving
            // out past the edge.  It leaves through the nearest face it moves
            // away from, as the clipping of a convex solid does.
            dist = tri_nearest_leaving(sc, vol, ph.pos, ph.dir, plane);
            if (!finitef(dist)) {
                return kInvalidState;
            }
        }
    } else if (kMesh && vol.plane_count != 0u) {
        mesh_exit(sc.planes, vol.plane_begin, vol.plane_count, ph.pos, ph.dir, dist, plane);
    } else {
        box_exit(vol, ph.pos, ph.dir, dist, axis);
    }
    uint32_t next = kInvalid;
    uint32_t surface = kInvalid;
    bool entering = false;
    if (vol.daughter_count != 0) {
        uint32_t which;
        int a2 = 0;
        uint32_t p2 = kInvalid;
        float t;
        if (kGeom == kGeomConvex && pre.which != kDaughterUnsearched) {
            t = pre.t;
            which = pre.which;
            a2 = pre.axis;
            p2 = pre.plane;
        } else {
            t = nearest_daughter<kGeom>(sc, vol, ph.pos, ph.dir, cfg.eps, which, a2, p2);
        }
        if (which != kInvalid && (kMesh ? t <= dist + cfg.eps : t <= dist)) {
            dist = t;
            axis = a2;
            plane = p2;
            next = which;
            surface = sc.volumes[which].entry_surface;
            entering = true;
        }
    }
    if (!finitef(dist)) {
        if (kGeom == kGeomOpen && open_solid && vol.parent != kInvalid) {
            // out of an open solid where none of its triangles lies ahead:
            // into the parent where the photon is, without a boundary
            ph.volume = vol.parent;
            zero_step = true;
            return kAlive;
        }
        return kEscaped;
    }
    if (dist <= cfg.eps) {
        zero_step = true;
    }
    G4GO_REGION_END(kRegionGeometry);

    const U4 w = rng.block(ph.photon_id, ph.event_id, ph.step, kProcStep, 0);

    // bulk interaction: absorption or Rayleigh (competing exponentials)
    if (finitef(ph.interaction_length)) {
        const float d_int = -g4_logf(u01(w.x)) * ph.interaction_length;
        if (d_int < dist) {
            ph.time += static_cast<double>(d_int * ph.inv_velocity);
            if (kTrace) {
                trace->segment = d_int;
            }
            if (u01(w.y) <= ph.absorb_fraction) {
                return kAbsorbedBulk;
            }
            ph.pos = add(ph.pos, scale(ph.dir, d_int));
            G4GO_REGION_BEGIN(kRegionRayleigh);
            rayleigh_scatter(ph, rng);
            G4GO_REGION_END(kRegionRayleigh);
            if (kTrace) {
                trace->event = kTraceRayleigh;
            }
            return kAlive;
        }
    }
    G4GO_REGION_BEGIN(kRegionSurface);
    ph.time += static_cast<double>(dist * ph.inv_velocity);
    if (kTrace) {
        trace->segment = dist;
    }
    V3 bpos = add(ph.pos, scale(ph.dir, dist));
    V3 normal;
    const bool mesh_face = kMesh && plane != kInvalid;
    const float dsign = axis_of(ph.dir, axis) > 0.0f ? 1.0f : -1.0f;
    if (mesh_face) {
        // the face's boundary normal, turned to face the incoming photon (reference :789-791)
        const Plane& pl = face_record<kGeom>(sc, plane);
        const V3 sn = {pl.shade[0], pl.shade[1], pl.shade[2]};
        normal = dot(ph.dir, sn) > 0.0f ? scale(sn, -1.0f) : sn;
    } else {
        // snap onto the face plane
        const Volume& face_box = entering ? sc.volumes[next] : vol;
        const float fplane = entering ? (dsign > 0.0f ? face_box.lo[axis] : face_box.hi[axis])
                                      : (dsign > 0.0f ? face_box.hi[axis] : face_box.lo[axis]);
        if (axis == 0) {
            bpos.x = fplane;
        } else if (axis == 1) {
            bpos.y = fplane;
        } else {
            bpos.z = fplane;
        }
        // normal facing the incoming photon (reference :789-791)
        normal = axis == 0 ? V3{-dsign, 0.0f, 0.0f} : (axis == 1 ? V3{0.0f, -dsign, 0.0f} : V3{0.0f, 0.0f, -dsign});
    }
    ph.pos = bpos;

    if (!entering) {
        if (mesh_face) {
            // the point just outside the face, along its outward plane normal
            // (an open solid's triangle: its normal's side the photon moves to)
            const Plane& pl = face_record<kGeom>(sc, plane);
            const V3 pn = {pl.n[0], pl.n[1], pl.n[2]};
            const bool back = kGeom == kGeomOpen && (vol.plane_count & kSolidOpen) != 0u && dot(pn, ph.dir) < 0.0f;
            const V3 q = add(bpos, scale(pn, back ? -cfg.eps : cfg.eps));
            const uint32_t across = pl.across;
            if (across == kFaceDynamic) {
                next = pl.sib_count != 0u
                           ? listed_neighbor<kGeom>(sc, scene_face_sibs(sc) + pl.sib_first, pl.sib_count, vol.parent, q)
                           : point_neighbor<kGeom>(sc, ph.volume, q);
                surface = next == kInvalid ? kInvalid : find_surface(sc, ph.volume, next);
            } else {
                next = vol.parent;
                surface = pl.surface_parent;
                if (across != kInvalid) {
                    const DaughterBox b = load_dbox(sc.dbox[across]);
                    if (q.x >= b.lo[0] && q.x <= b.hi[0] && q.y >= b.lo[1] && q.y <= b.hi[1] && q.z >= b.lo[2] &&
                        q.z <= b.hi[2]) {
                        if (b.mesh == 0u || solid_contains<kGeom>(sc, sc.volumes[b.volume], q)) {
                            next = b.volume;
                            surface = pl.surface_across;
                        }
                    }
                }
            }
        } else {
            const int face = 2 * axis + (dsign > 0.0f ? 1 : 0);
            Face fr;
            __builtin_memcpy(&fr, &vol.faces[face], sizeof(Face));
            next = fr.neighbor;
            if (next == kFaceDynamic) {
                const V3 q = add(bpos, scale(normal, -cfg.eps));
                next = point_neighbor<kGeom>(sc, ph.volume, q);
                surface = next == kInvalid ? kInvalid : find_surface(sc, ph.volume, next);
            } else {
                surface = fr.surface;
            }
        }
        if (next == kInvalid) {
            return ph.volume == sc.world ? kEscaped : kInvalidState;
        }
    }
    const Volume& nvol = sc.volumes[next];
    if (nvol.material >= sc.n_materials) {
        return kInvalidState;
    }
    const Material& nmat = sc.materials[nvol.material];
    const float n1 = ph.rindex;
    const float e = ph.energy;

    bool reflected = false, detected = false, absorbed = false, transmitted = false;
    V3 dir = ph.dir, pol = ph.pol;
    if (surface < sc.n_surfaces) {
        const Surface& s = sc.surfaces[surface];
        if (kTrace) {
            trace->surface = surface;
        }
        // classify_surface clamps both to [0, 1]
        const float refl = prop_eval_ph(sc.pool, s.reflectivity, ph, 0.0f);
        const float trans = prop_eval_ph(sc.pool, s.transmittance, ph, 0.0f);
        const int outcome = classify_surface(refl, trans, s.reflectivity.count != 0, s.transmittance.count != 0, u01(w.z));
        if (kTrace) {
            trace->event = outcome == kInteraction ? kTraceInteraction
                                                   : (outcome == kDirectTransmit ? kTraceTransmit : kTraceNone);
        }
        if (outcome == kAbsorb) {
            G4GO_REGION_BEGIN(kRegionAbsorbDetect);
            absorbed = true;
            const float eff = clamp01(prop_eval_ph(sc.pool, s.efficiency, ph, 0.0f));
            detected = evaluate_absorption(eff, nvol.sensor != kInvalid || vol.sensor != kInvalid, u01(w.w)) == kDetect;
            G4GO_REGION_END(kRegionAbsorbDetect);
        } else if (outcome == kDirectTransmit) {
            transmitted = true;
        } else if (s.type == kDielectricDielectric) {
            const float n2 = prop_eval_ph(sc.pool, nmat.rindex, ph, u2f(0x7fc00000u));
            if (!(n1 > 0.0f) || !finitef(n1) || !(n2 > 0.0f) || !finitef(n2)) {
                absorbed = true;
            } else {
                const V3 inormal = s.finish == kPolished ? normalize(normal) : sample_facet_normal(s, normal, dir, rng, ph);
                const float rough = prop_eval_ph(sc.pool, s.roughness, ph, 0.0f);
                bool rough_pass = true;
                U4 wr = {0u, 0u, 0u, 0u};
                if (rough > 0.0f && n1 > n2 && finitef(rough) && finitef(e) && e > 0.0f) {
                    const float cost1 = fmaxf(0.0f, -dot(dir, inormal));
                    const float wavelength_mm = 1.239841984e-3f / e;
                    const float expo = 4.0f * 3.14159265358979323846f * rough * n1 * cost1 / wavelength_mm;
                    wr = rng.block(ph.photon_id, ph.event_id, ph.step, kProcRough, 0);
                    rough_pass = u01(wr.x) < g4_expf(-expo * expo);
                }
                const FresnelState f = fresnel_state(dir, pol, inormal, n1, n2);
                reflected = f.tir || static_cast<double>(u01(w.w)) > f.transmittance;
                if (reflected) {
                    if (rough_pass) {
                        if (s.model == kUnified && s.finish != kPolished) {
                            dir = sample_unified_reflection(sc, s, dir, normal, rng, ph);
                            pol = reflect_polarization(pol, normal, dir);
                        } else {
                            fresnel_reflect(f, n1, n2, dir, pol);
                        }
                    } else {
                        dir = sample_lambertian(normal, u01(wr.y), u01(wr.z));
                        pol = project_polarization(pol, dir);
                    }
                } else {
                    fresnel_transmit(f, n1, n2, dir, pol);
                    transmitted = true;
                }
            }
        } else {
            reflected = true;
            if (s.finish == kPolished && mesh_face) {
                dir = reflect_direction(dir, normal);
                pol = reflect_polarization(pol, normal, dir);
            } else if (s.finish == kPolished) {
                // specular reflection at the axis-aligned face
                dir = negate_axis(dir, axis);
                pol = negate_off_axis(pol, axis);
            } else {
                dir = sample_surface_reflection(sc, s, dir, normal, rng, ph);
                pol = reflect_polarization(pol, normal, dir);
            }
        }
    } else {
        const float n2 = prop_eval_ph(sc.pool, nmat.rindex, ph, u2f(0x7fc00000u));
        if (!(n1 > 0.0f) || !finitef(n1) || !(n2 > 0.0f) || !finitef(n2)) {
            absorbed = true;
        } else {
            const FresnelState f = fresnel_state(dir, pol, normal, n1, n2);
            reflected = static_cast<double>(u01(w.w)) > f.transmittance;
            if (reflected) {
                fresnel_reflect(f, n1, n2, dir, pol);
            } else {
                fresnel_transmit(f, n1, n2, dir, pol);
                transmitted = true;
            }
        }
    }

    G4GO_REGION_END(kRegionSurface);
    if (detected) {
        sensor_out = nvol.sensor != kInvalid ? nvol.sensor : vol.sensor;
        if (sensor_out == kInvalid) {
            return kInvalidState;
        }
        if (kTrace) {
            trace->event = kTraceDetect;
        }
        return kDetected;
    }
    if (absorbed) {
        return kAbsorbedSurface;
    }
    ph.dir = dir;
    ph.pol = pol;
    if (transmitted) {
        ph.volume = next;
    }
    return kAlive;
}

}  // namespace g4go
