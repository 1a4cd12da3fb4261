/*
 * oracle.c -- CPU restatement of G4GO's optical-photon transport, used ONLY
 * as the reference the HIP backend is tested against (tests, smoke check, CPU
 * baseline leg of bench.py).  It is never linked into or called by the product.
 *
 * The same source builds twice:
 *   liboracle32.so  FP32 state, FP64 time: bit-exact replay of the HIP
 *                   kernels on the same Philox streams (IEEE float, no FP
 *                   contraction, explicit fmaf, Cephes elementary functions);
 *   liboracle64.so  FP64 everywhere with libm elementary functions: the
 *                   reference semantics without FP32 rounding, used for the
 *                   statistical gates and the FP64 replay gradient.
 *
 * Reference algorithm, function by function (paths relative to the G4GO tree):
 *   or_philox / or_uniform       src/optical/optix/OptiXTransportDevice.cu:254-286
 *   or_prop                      OptiXTransportDevice.cu:198-216 (linear scan)
 *   or_spectrum                  OptiXTransportDevice.cu:322-344,
 *                                CDF: src/optical/optix/OptiXTransportHost.cpp:371-410
 *   or_normalize .. or_fresnel   src/optical/optix/OptiXBoundaryPhysics.cuh:47-178
 *   or_classify / or_absorption  OptiXBoundaryPhysics.cuh:79-101
 *   or_lambertian                OptiXBoundaryPhysics.cuh:103-120
 *   or_find_surface              OptiXTransportDevice.cu:218-252, src/optical/Scene.cpp:120-150
 *   or_facet / or_unified ...    OptiXTransportDevice.cu:492-595
 *   or_generate_emission         OptiXTransportDevice.cu:346-481
 *   mesh_solid / mesh_exit ..    closed convex mesh solids as face planes, other meshes
 *   wt_hit / tri_closest         as triangles (watertight test, Woop et al. 2013), in
 *                                place of the closest triangle hit,
 *                                OptiXTransportDevice.cu:105-196, triangles from
 *                                OptiXTransportHost.cpp:1084-1200; open meshes hit
 *                                from either side beyond eps as the reference's
 *                                trace (OptiXTransportDevice.cu:742-746)
 *   or_step                      OptiXTransportDevice.cu:708-931 (bounce loop body)
 *   or_run                       OptiXTransportDevice.cu:660-941 + the batch
 *                                statistics of OptiXTransportHost.cpp:809-844
 *   Rayleigh (absent from the reference GPU path): Geant4 11 G4OpRayleigh,
 *   rotateUz: CLHEP Hep3Vector::rotateUz.
 *
 * Deliberate differences from the reference, shared with the HIP backend:
 * the draw map (four draws per Philox block, see physics.h), box solids with
 * analytic faces and convex mesh solids as face planes instead of OptiX
 * triangles, no epsilon push-off (the photon is snapped onto a box face
 * plane), the competing-exponential bulk step
 * (absorption + Rayleigh), FP64 time, sin/cos of 2*pi*u by turns.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "../include/g4go_hip.h"

#ifdef ORACLE_DOUBLE
typedef double real;
#define RF(x) x
#define R_SQRT sqrt
#define R_FABS fabs
#define R_FMIN fmin
#define R_FMAX fmax
#define R_FLOOR floor
#define R_FMA fma
#define R_INF ((double)INFINITY)
#define ORACLE_NAME(n) oracle64_##n
#else
typedef float real;
#define RF(x) x##f
#define R_SQRT sqrtf
#define R_FABS fabsf
#define R_FMIN fminf
#define R_FMAX fmaxf
#define R_FLOOR floorf
#define R_FMA fmaf
#define R_INF INFINITY
#define ORACLE_NAME(n) oracle32_##n
#endif

#define INVALID 0xffffffffu

/* ------------------------------------------------------------------ */
/* Philox4x32-10                                                       */
/* ------------------------------------------------------------------ */
void ORACLE_NAME(philox)(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
    uint32_t x0 = ctr[0], x1 = ctr[1], x2 = ctr[2], x3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; ++r) {
        const uint64_t a = (uint64_t)0xD2511F53u * x0;
        const uint64_t b = (uint64_t)0xCD9E8D57u * x2;
        const uint32_t y0 = (uint32_t)(b >> 32) ^ x1 ^ k0;
        const uint32_t y1 = (uint32_t)b;
        const uint32_t y2 = (uint32_t)(a >> 32) ^ x3 ^ k1;
        const uint32_t y3 = (uint32_t)a;
        x0 = y0;
        x1 = y1;
        x2 = y2;
        x3 = y3;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    out[0] = x0;
    out[1] = x1;
    out[2] = x2;
    out[3] = x3;
}

/* The reference Uniform(): one block per draw, word draw % 4. */
float ORACLE_NAME(uniform)(uint64_t seed, uint64_t photon_key, uint32_t bounce, uint32_t process, uint32_t draw) {
    const uint32_t ctr[4] = {(uint32_t)photon_key, (uint32_t)(photon_key >> 32), bounce, (process << 16) ^ draw};
    const uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    uint32_t w[4];
    ORACLE_NAME(philox)(ctr, key, w);
    return ((float)w[draw % 4u] + 0.5f) / 4294967296.0f;
}

typedef struct {
    uint32_t k0, k1;
} rng_t;

static void block4(const rng_t* r, uint32_t photon, uint32_t event, uint32_t step, uint32_t process, uint32_t blk,
                   uint32_t w[4]) {
    const uint32_t ctr[4] = {photon, event, step, (process << 16) ^ blk};
    const uint32_t key[2] = {r->k0, r->k1};
    ORACLE_NAME(philox)(ctr, key, w);
}

enum { P_STEP = 1, P_RAYLEIGH = 2, P_FACET = 3, P_LOBE = 4, P_ROUGH = 5, P_SYNTH = 6, P_EMISSION = 7 };

#ifdef ORACLE_DOUBLE
static real u01(uint32_t w) { return ((double)w + 0.5) / 4294967296.0; }
static real r_log(real x) { return log(x); }
static real r_exp(real x) { return exp(x); }
static void r_sincos_turn(real u, real* s, real* c) {
    const double a = 6.283185307179586476925 * u;
    *s = sin(a);
    *c = cos(a);
}
#else
/* (float(w) + 0.5) / 2^32 written as one fma: identical rounding. */
static real u01(uint32_t w) { return fmaf((float)w, 0x1p-32f, 0x1p-33f); }

static uint32_t fbits(float x) {
    uint32_t b;
    memcpy(&b, &x, 4);
    return b;
}
static float bitsf(uint32_t b) {
    float x;
    memcpy(&x, &b, 4);
    return x;
}

/* Cephes logf (Moshier 1989). */
static float r_log(float x) {
    if (!(x > 0.0f)) {
        return x == 0.0f ? -INFINITY : NAN;
    }
    if (isinf(x)) {
        return x;
    }
    uint32_t b = fbits(x);
    int e = 0;
    if (b < 0x00800000u) {
        x = x * 33554432.0f;
        b = fbits(x);
        e = -25;
    }
    e += (int)(b >> 23) - 126;
    float m = bitsf((b & 0x007fffffu) | 0x3f000000u);
    if (m < 0.70710678118654752f) {
        e -= 1;
        m = (m + m) - 1.0f;
    } else {
        m = m - 1.0f;
    }
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
    const float fe = (float)e;
    y = fmaf(fe, -2.12194440e-4f, y);
    y = fmaf(-0.5f, z, y);
    float r = m + y;
    r = fmaf(fe, 0.693359375f, r);
    return r;
}

/* Cephes expf with exact 2^n scaling; flush below the normal range. */
static float r_exp(float x) {
    if (x != x) {
        return x;
    }
    if (x > 88.7228391f) {
        return INFINITY;
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
    int ni = (int)n;
    if (ni > 127) {
        y = y * 2.0f;
        ni -= 1;
    }
    return y * bitsf((uint32_t)(ni + 127) << 23);
}

/* sin, cos of 2*pi*u: reduction in turns, Cephes sinf/cosf polynomials. */
static void r_sincos_turn(float u, float* s, float* c) {
    const float t = u * 4.0f;
    const float q = floorf(t + 0.5f);
    const float x = (t - q) * 1.57079632679489662f;
    const float z = x * x;
    const float sp = fmaf(fmaf(-1.9515295891e-4f, z, 8.3321608736e-3f), z, -1.6666654611e-1f);
    const float sn = fmaf(sp * z, x, x);
    const float cp = fmaf(fmaf(2.443315711809948e-5f, z, -1.388731625493765e-3f), z, 4.166664568298827e-2f);
    const float cs = fmaf(cp * z, z, fmaf(-0.5f, z, 1.0f));
    const int quadrant = ((int)q) & 3;
    if (quadrant == 0) {
        *s = sn;
        *c = cs;
    } else if (quadrant == 1) {
        *s = cs;
        *c = -sn;
    } else if (quadrant == 2) {
        *s = -sn;
        *c = -cs;
    } else {
        *s = -cs;
        *c = sn;
    }
}
#endif

real ORACLE_NAME(log)(real x) { return r_log(x); }
real ORACLE_NAME(exp)(real x) { return r_exp(x); }
void ORACLE_NAME(sincos_turn)(real u, real* s, real* c) { r_sincos_turn(u, s, c); }

/* ------------------------------------------------------------------ */
/* Vectors and boundary physics (OptiXBoundaryPhysics.cuh)             */
/* ------------------------------------------------------------------ */
typedef struct {
    real x, y, z;
} v3;

static v3 mk(real x, real y, real z) {
    v3 v = {x, y, z};
    return v;
}
static v3 vadd(v3 a, v3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
static v3 vsub(v3 a, v3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
static v3 vscale(v3 a, real s) { return mk(a.x * s, a.y * s, a.z * s); }
static real vdot(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static v3 vcross(v3 a, v3 b) { return mk(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x); }
static real vaxis(v3 a, int k) { return k == 0 ? a.x : (k == 1 ? a.y : a.z); }
static int finite_r(real x) { return isfinite(x); }

static v3 or_normalize(v3 a) {
    const real l2 = vdot(a, a);
    if (!(l2 > RF(0.0))) {
        return mk(RF(0.0), RF(0.0), RF(0.0));
    }
    return vscale(a, RF(1.0) / R_SQRT(l2));
}

static real clamp01(real v) { return R_FMIN(R_FMAX(v, RF(0.0)), RF(1.0)); }

static v3 or_project(v3 pol, v3 dir) {
    v3 out = vsub(pol, vscale(dir, vdot(pol, dir)));
    if (vdot(out, out) < RF(1.0e-12)) {
        const v3 helper = R_FABS(dir.z) < RF(0.9) ? mk(RF(0.0), RF(0.0), RF(1.0)) : mk(RF(1.0), RF(0.0), RF(0.0));
        out = vcross(dir, helper);
    }
    return or_normalize(out);
}

static v3 or_reflect_dir(v3 d, v3 n) { return or_normalize(vsub(d, vscale(n, RF(2.0) * vdot(d, n)))); }

static v3 or_reflect_pol(v3 p, v3 n, v3 d) {
    const v3 r = vadd(vscale(p, RF(-1.0)), vscale(n, RF(2.0) * vdot(p, n)));
    return or_project(r, d);
}

enum { O_ABSORB = 0, O_DETECT = 1, O_TRANSMIT = 2, O_INTERACT = 3 };

int ORACLE_NAME(classify)(real r, real t, int has_r, int has_t, real sample) {
    const real u = clamp01(sample);
    const real rr = has_r ? clamp01(r) : RF(0.0);
    const real tt = has_t ? clamp01(t) : RF(0.0);
    if (!has_r && !has_t) {
        return O_INTERACT;
    }
    if (u > rr + tt) {
        return O_ABSORB;
    }
    if (u > rr) {
        return O_TRANSMIT;
    }
    return O_INTERACT;
}

int ORACLE_NAME(absorption)(real eff, int sensor, real sample) {
    return (sensor && sample < clamp01(eff)) ? O_DETECT : O_ABSORB;
}

static v3 or_lambertian(v3 n, real radial_s, real azim_s) {
    n = or_normalize(n);
    radial_s = clamp01(radial_s);
    azim_s = clamp01(azim_s);
    const real radial = R_SQRT(radial_s);
    real s, c;
    r_sincos_turn(azim_s, &s, &c);
    const real nc = R_SQRT(R_FMAX(RF(0.0), RF(1.0) - radial_s));
    const v3 helper = R_FABS(n.z) < RF(0.9) ? mk(RF(0.0), RF(0.0), RF(1.0)) : mk(RF(1.0), RF(0.0), RF(0.0));
    const v3 t = or_normalize(vcross(helper, n));
    const v3 b = vcross(n, t);
    return or_normalize(vadd(vadd(vscale(t, radial * c), vscale(b, radial * s)), vscale(n, nc)));
}

void ORACLE_NAME(lambertian)(const real n[3], real radial, real azimuth, real out[3]) {
    const v3 d = or_lambertian(mk(n[0], n[1], n[2]), radial, azimuth);
    out[0] = d.x;
    out[1] = d.y;
    out[2] = d.z;
}

typedef struct {
    v3 rdir, tdir, rpol, tpol;
    double t; /* transmission probability, evaluated in FP64 as Geant4 does */
    int tir;
} fresnel_t;

typedef struct {
    double x, y, z;
} d3;
static d3 dv(v3 a) {
    d3 r = {(double)a.x, (double)a.y, (double)a.z};
    return r;
}
static double ddot(d3 a, d3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static d3 dcross(d3 a, d3 b) {
    d3 r = {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
    return r;
}

/* ComputeFresnel (OptiXBoundaryPhysics.cuh:122-178): vectors in the state
 * precision with the reference's operation order; the reflect/transmit
 * probability and the TIR test in FP64 from that state. */
static fresnel_t or_fresnel(v3 d, v3 p, v3 n, real n1, real n2) {
    fresnel_t o;
    memset(&o, 0, sizeof(o));
    d = or_normalize(d);
    n = or_normalize(n);
    p = or_project(p, d);
    const d3 dd = dv(d), dn = dv(n), dp = dv(p);
    const double n1d = n1, n2d = n2;
    const double ci_d = fmin(fmax(-ddot(dd, dn), 0.0), 1.0);
    const double eta_d = n1d / n2d;
    const double st2_d = eta_d * eta_d * (1.0 - ci_d * ci_d);
    const real ci = clamp01(-vdot(d, n));
    const real eta = n1 / n2;
    const real st2 = eta * eta * (RF(1.0) - ci * ci);
    o.rdir = or_reflect_dir(d, n);
    o.tir = st2_d >= 1.0;
    o.t = 0.0;
    if (o.tir) {
        o.rpol = or_reflect_pol(p, n, o.rdir);
        return o;
    }
    const double ct_d = sqrt(fmax(0.0, 1.0 - st2_d));
    const real ct = R_SQRT(R_FMAX(RF(0.0), RF(1.0) - st2));
    o.tdir = or_normalize(vadd(vscale(d, eta), vscale(n, eta * ci - ct)));
    const v3 tr = vcross(d, n);
    if (vdot(tr, tr) < RF(1.0e-12)) {
        const double ra = (n1d - n2d) / (n1d + n2d);
        o.t = fmin(fmax(1.0 - ra * ra, 0.0), 1.0);
        o.rpol = n2 > n1 ? vscale(p, RF(-1.0)) : p;
        o.tpol = p;
        return o;
    }
    {
        /* squared projections of p on the unit s and p directions (both
         * perpendicular to d): |p.s|^2 = (p.(d x n))^2 / |d x n|^2 and
         * |p.s|^2 + |p.p|^2 = |p|^2 - (p.d)^2 / |d|^2 */
        const d3 c = dcross(dd, dn);
        const double pc = ddot(dp, c);
        const double es2 = pc * pc / ddot(c, c);
        const double pdd = ddot(dp, dd);
        const double perp2 = ddot(dp, dp) - pdd * pdd / ddot(dd, dd);
        const double rs = (n1d * ci_d - n2d * ct_d) / (n1d * ci_d + n2d * ct_d);
        const double rp = (n2d * ci_d - n1d * ct_d) / (n2d * ci_d + n1d * ct_d);
        o.t = fmin(fmax(1.0 - (es2 * (rs * rs) + (perp2 - es2) * (rp * rp)), 0.0), 1.0);
    }
    const v3 sd = or_normalize(tr);
    const v3 pd = or_normalize(vcross(d, sd));
    const real is = vdot(p, sd);
    const real ip = vdot(p, pd);
    const real sden = n1 * ci + n2 * ct;
    const real pden = n2 * ci + n1 * ct;
    const real rs = (n1 * ci - n2 * ct) / sden;
    const real rp = (n2 * ci - n1 * ct) / pden;
    const real ts = RF(2.0) * n1 * ci / sden;
    const real tp = RF(2.0) * n1 * ci / pden;
    const v3 rpd = or_normalize(vcross(o.rdir, sd));
    const v3 tpd = or_normalize(vcross(o.tdir, sd));
    o.rpol = or_normalize(vadd(vscale(sd, is * rs), vscale(rpd, ip * rp)));
    o.tpol = or_normalize(vadd(vscale(sd, is * ts), vscale(tpd, ip * tp)));
    return o;
}

/* 11 inputs -> 14 outputs, same packing as g4go_eval_fresnel. */
void ORACLE_NAME(fresnel)(const real* in, real* out) {
    const fresnel_t f = or_fresnel(mk(in[0], in[1], in[2]), mk(in[3], in[4], in[5]), mk(in[6], in[7], in[8]), in[9], in[10]);
    out[0] = (real)f.t;
    out[1] = f.tir ? RF(1.0) : RF(0.0);
    const v3 vs[4] = {f.rdir, f.tdir, f.rpol, f.tpol};
    for (int k = 0; k < 4; ++k) {
        out[2 + 3 * k] = vs[k].x;
        out[3 + 3 * k] = vs[k].y;
        out[4 + 3 * k] = vs[k].z;
    }
}

/* CLHEP Hep3Vector::rotateUz. */
static v3 or_rotate_uz(v3 v, v3 u) {
    const real up2 = u.x * u.x + u.y * u.y;
    if (up2 > RF(0.0)) {
        const real up = R_SQRT(up2);
        return mk((u.x * u.z * v.x - u.y * v.y) / up + u.x * v.z, (u.y * u.z * v.x + u.x * v.y) / up + u.y * v.z,
                  -up * v.x + u.z * v.z);
    }
    if (u.z < RF(0.0)) {
        return mk(-v.x, v.y, -v.z);
    }
    return v;
}

/* ------------------------------------------------------------------ */
/* Scene (built from the public g4go_scene description)               */
/* ------------------------------------------------------------------ */
typedef struct {
    uint32_t n;
    real* e;
    real* v;
    int constant;
} prop_t;

typedef struct {
    prop_t rindex, groupvel, abslen, rayleigh, spectrum[3];
    real rindex_max;
} material_t;

typedef struct {
    uint32_t type, model, finish;
    real model_value;
    prop_t refl, eff, trans, lobe, spike, back, rough;
} surface_t;

typedef struct {
    float lo[3], hi[3]; /* box bounds (a mesh's bounding box) are FP32 in both builds (as uploaded) */
    uint32_t id, material, parent, skin, sensor;
    uint32_t plane_begin, plane_count; /* mesh solid: planes (or triangles when tri); plane_count 0: the box */
    int tri;
    int open; /* triangle solid without an inside (open mesh): hit from either side, holds no point */
} volume_t;

/* Face plane of a convex mesh solid: inside where n.x <= h; s = boundary normal. */
typedef struct {
    real n[3], h, s[3];
} plane_t;

/* Triangle of a non-convex mesh solid, wound outwards (its plane in face). */
typedef struct {
    plane_t face;
    real v[3][3];
} otri_t;

typedef struct {
    uint32_t n_mat, n_surf, n_vol, n_bind, world;
    material_t* mat;
    surface_t* surf;
    volume_t* vol;
    uint32_t* bind; /* triples (from, to, surface) in input order */
    plane_t* planes;
    uint32_t n_planes;
    otri_t* tris;
    uint32_t n_tris;
} oscene_t;

static prop_t make_prop(const g4go_property* p) {
    prop_t o;
    memset(&o, 0, sizeof(o));
    if (p->count == 0) {
        return o;
    }
    o.n = p->count;
    o.e = (real*)malloc(sizeof(real) * p->count);
    o.v = (real*)malloc(sizeof(real) * p->count);
    o.constant = 1;
    for (uint32_t i = 0; i < p->count; ++i) {
        o.e[i] = p->energy_ev[i];
        o.v[i] = p->values[i];
        if (p->values[i] != p->values[0]) {
            o.constant = 0;
        }
    }
    return o;
}

/* Trapezoid CDF in FP32 exactly as the host code computes it. */
static prop_t make_spectrum(const g4go_property* p) {
    prop_t o;
    memset(&o, 0, sizeof(o));
    if (p->count == 0) {
        return o;
    }
    float* cdf = (float*)calloc(p->count, sizeof(float));
    if (p->count == 1) {
        cdf[0] = 1.0f;
    } else {
        for (uint32_t i = 1; i < p->count; ++i) {
            const float de = p->energy_ev[i] - p->energy_ev[i - 1];
            const float v0 = p->values[i - 1] > 0.0f ? p->values[i - 1] : 0.0f;
            const float v1 = p->values[i] > 0.0f ? p->values[i] : 0.0f;
            if (!(de >= 0.0f) || !isfinite(de) || !isfinite(v0) || !isfinite(v1)) {
                free(cdf);
                return o;
            }
            cdf[i] = cdf[i - 1] + 0.5f * de * (v0 + v1);
        }
        const float integral = cdf[p->count - 1];
        if (!(integral > 0.0f) || !isfinite(integral)) {
            free(cdf);
            return o;
        }
        for (uint32_t i = 0; i < p->count; ++i) {
            cdf[i] /= integral;
        }
        cdf[p->count - 1] = 1.0f;
    }
    o.n = p->count;
    o.e = (real*)malloc(sizeof(real) * p->count);
    o.v = (real*)malloc(sizeof(real) * p->count);
    for (uint32_t i = 0; i < p->count; ++i) {
        o.e[i] = p->energy_ev[i];
        o.v[i] = cdf[i];
    }
    free(cdf);
    return o;
}

static void free_prop(prop_t* p) {
    free(p->e);
    free(p->v);
}

static uint32_t vol_index(const oscene_t* s, uint32_t id) {
    for (uint32_t i = 0; i < s->n_vol; ++i) {
        if (s->vol[i].id == id) {
            return i;
        }
    }
    return INVALID;
}

/* Solid of a closed mesh (OptiXTransportHost.cpp:1084-1200 builds the
 * reference's triangles).  In FP64: per triangle the unit normal of
 * (v1 - v0) x (v2 - v0) and offset u.v0, turned so that the vertex centroid
 * lies inside; triangles whose normals, offsets (relative to the largest
 * centroid-vertex coordinate distance) and boundary normals agree within 1e-9
 * share the first one's plane (at most 1024 planes).  The mesh is convex when
 * the centroid lies more than 1e-6 of that size inside every plane and no
 * vertex more than that outside one: it becomes planes.  Otherwise it becomes
 * its triangles in mesh order (key = index).  A closed one is wound
 * consistently (mesh_topology) and all reversed when the signed volume about
 * the centroid is negative; an open, non-orientable or flat one is an open
 * solid, its triangles as given.  Sets v's range and bounding box (rounded
 * outwards to FP32). */
#define MAX_PLANES 1024

/* Winding of a closed mesh made consistent (OptiX ignores the winding,
 * OptiXTransportHost.cpp:1118-1200; the library's closed triangle solids
 * need it): vertices are welded by equal coordinates; two triangles sharing a
 * welded edge must traverse it in opposite directions, so rev[u] = rev[t] ^
 * dir_t ^ dir_u ^ 1 along every shared edge, with rev[0] = 0.  On a connected
 * orientable surface that fixes every triangle whatever the visiting order;
 * here a depth-first walk (the library walks breadth-first). */
typedef struct {
    uint64_t key; /* welded (min, max) vertex pair */
    uint32_t tri, dir;
} oedge_t;
static int oedge_cmp(const void* a, const void* b) {
    const oedge_t* x = (const oedge_t*)a;
    const oedge_t* y = (const oedge_t*)b;
    if (x->key != y->key) return x->key < y->key ? -1 : 1;
    return x->tri < y->tri ? -1 : (x->tri > y->tri);
}
static uint32_t weld_of(const double* x, uint32_t nv, uint32_t i) {
    for (uint32_t j = 0; j < nv; ++j) { /* the first vertex with the same coordinates */
        if (x[3 * j] == x[3 * i] && x[3 * j + 1] == x[3 * i + 1] && x[3 * j + 2] == x[3 * i + 2]) return j;
    }
    return i;
}
/* Topology of a mesh: returns 1 when it is closed (every welded edge in
 * exactly two triangles), else 0.  For a closed mesh, rev[t] = 1 when
 * triangle t is reversed to wind the mesh consistently, and *orientable = 0
 * when no winding is (a triangle forced both ways, or a part the walk from
 * triangle 0 does not reach while the given winding is inconsistent). */
static int mesh_topology(const g4go_volume_mesh* m, uint8_t* rev, int* orientable) {
    const uint32_t nt = m->n_triangles, nv = m->n_vertices;
    uint32_t* w = (uint32_t*)malloc(nv * sizeof(uint32_t));
    for (uint32_t i = 0; i < nv; ++i) w[i] = weld_of(m->vertices_mm, nv, i);
    oedge_t* e = (oedge_t*)malloc(3u * nt * sizeof(oedge_t));
    for (uint32_t t = 0; t < nt; ++t) {
        for (int k = 0; k < 3; ++k) {
            const uint32_t a = w[m->indices[3 * t + k]], b = w[m->indices[3 * t + (k + 1) % 3]];
            e[3 * t + k].key = ((uint64_t)(a < b ? a : b) << 32) | (a < b ? b : a);
            e[3 * t + k].tri = t;
            e[3 * t + k].dir = a > b;
        }
    }
    qsort(e, 3u * nt, sizeof(oedge_t), oedge_cmp);
    int closed = 1, consistent = 1;
    for (uint32_t i = 0; i < 3u * nt;) {
        uint32_t j = i;
        while (j < 3u * nt && e[j].key == e[i].key) ++j;
        if (j - i != 2) {
            closed = 0;
        } else if (e[i].dir == e[i + 1].dir) {
            consistent = 0;
        }
        i = j;
    }
    memset(rev, 0, nt);
    *orientable = 1;
    if (closed && !consistent) {
        /* per triangle, the positions of its three edges in e (an edge's two
         * entries are e[2i], e[2i + 1]) */
        uint32_t* at = (uint32_t*)malloc(3u * nt * sizeof(uint32_t));
        uint32_t* cnt = (uint32_t*)calloc(nt, sizeof(uint32_t));
        for (uint32_t i = 0; i < 3u * nt; ++i) at[3 * e[i].tri + cnt[e[i].tri]++] = i;
        uint8_t* seen = (uint8_t*)calloc(nt, 1);
        uint32_t* stack = (uint32_t*)malloc(nt * sizeof(uint32_t));
        uint32_t top = 0, reached = 1;
        stack[top++] = 0;
        seen[0] = 1;
        while (top > 0) {
            const uint32_t t = stack[--top];
            for (int k = 0; k < 3; ++k) {
                const uint32_t i = at[3 * t + k];
                const oedge_t* o = &e[i ^ 1u];
                const uint8_t want = (uint8_t)(rev[t] ^ e[i].dir ^ o->dir ^ 1u);
                if (!seen[o->tri]) {
                    seen[o->tri] = 1;
                    rev[o->tri] = want;
                    stack[top++] = o->tri;
                    ++reached;
                } else if (rev[o->tri] != want) {
                    *orientable = 0;
                }
            }
        }
        if (reached != nt) *orientable = 0;
        if (!*orientable) memset(rev, 0, nt);
        free(stack);
        free(seen);
        free(cnt);
        free(at);
    }
    free(e);
    free(w);
    return closed;
}

static void mesh_solid(oscene_t* s, const g4go_volume_mesh* m, volume_t* v) {
    const uint32_t nv = m->n_vertices, nt = m->n_triangles;
    const double* x = m->vertices_mm;
    double c[3] = {0.0, 0.0, 0.0}, lo[3], hi[3], size = 0.0;
    for (uint32_t i = 0; i < nv; ++i) {
        for (int k = 0; k < 3; ++k) {
            c[k] += x[3 * i + k];
        }
    }
    for (int k = 0; k < 3; ++k) {
        c[k] /= (double)nv;
        lo[k] = hi[k] = x[k];
    }
    for (uint32_t i = 0; i < nv; ++i) {
        for (int k = 0; k < 3; ++k) {
            const double d = fabs(x[3 * i + k] - c[k]);
            size = d > size ? d : size;
            lo[k] = x[3 * i + k] < lo[k] ? x[3 * i + k] : lo[k];
            hi[k] = x[3 * i + k] > hi[k] ? x[3 * i + k] : hi[k];
        }
    }
    const double tol = 1e-6 * size;
    double(*f64)[7] = (double(*)[7])calloc(nt, sizeof(*f64)); /* u[3], h, s[3] */
    uint32_t nf = 0;
    int convex = 1;
    for (uint32_t t = 0; t < nt; ++t) {
        const double* p0 = x + 3 * m->indices[3 * t];
        const double* p1 = x + 3 * m->indices[3 * t + 1];
        const double* p2 = x + 3 * m->indices[3 * t + 2];
        const double a0 = p1[0] - p0[0], a1 = p1[1] - p0[1], a2 = p1[2] - p0[2];
        const double b0 = p2[0] - p0[0], b1 = p2[1] - p0[1], b2 = p2[2] - p0[2];
        const double g0 = a1 * b2 - a2 * b1, g1 = a2 * b0 - a0 * b2, g2 = a0 * b1 - a1 * b0;
        const double len = sqrt(g0 * g0 + g1 * g1 + g2 * g2);
        double f[7] = {g0 / len, g1 / len, g2 / len, 0.0, 0.0, 0.0, 0.0};
        f[3] = f[0] * p0[0] + f[1] * p0[1] + f[2] * p0[2];
        if (f[0] * c[0] + f[1] * c[1] + f[2] * c[2] > f[3]) {
            f[0] = -f[0];
            f[1] = -f[1];
            f[2] = -f[2];
            f[3] = -f[3];
        }
        if (m->triangle_normals) {
            const double* n = m->triangle_normals + 3 * t;
            const double nl = sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
            f[4] = n[0] / nl;
            f[5] = n[1] / nl;
            f[6] = n[2] / nl;
        } else {
            f[4] = f[0];
            f[5] = f[1];
            f[6] = f[2];
        }
        int dup = 0;
        for (uint32_t j = 0; j < nf && !dup; ++j) {
            int same = fabs(f64[j][3] - f[3]) <= 1e-9 * size;
            for (int k = 0; k < 3; ++k) {
                same = same && fabs(f64[j][k] - f[k]) <= 1e-9 && fabs(f64[j][4 + k] - f[4 + k]) <= 1e-9;
            }
            dup = same;
        }
        if (!dup) {
            if (nf == MAX_PLANES) {
                convex = 0;
                break;
            }
            memcpy(f64[nf++], f, sizeof(f));
        }
    }
    uint8_t* rev = (uint8_t*)malloc(nt);
    int orientable;
    const int closed = mesh_topology(m, rev, &orientable);
    convex = convex && closed;
    for (uint32_t j = 0; convex && j < nf; ++j) {
        convex = f64[j][3] - (f64[j][0] * c[0] + f64[j][1] * c[1] + f64[j][2] * c[2]) > tol;
        for (uint32_t i = 0; convex && i < nv; ++i) {
            convex = !(f64[j][0] * x[3 * i] + f64[j][1] * x[3 * i + 1] + f64[j][2] * x[3 * i + 2] - f64[j][3] > tol);
        }
    }
    if (convex) {
        s->planes = (plane_t*)realloc(s->planes, (s->n_planes + nf) * sizeof(plane_t));
        v->plane_begin = s->n_planes;
        v->plane_count = nf;
        for (uint32_t j = 0; j < nf; ++j) {
            plane_t* pl = &s->planes[s->n_planes + j];
            for (int k = 0; k < 3; ++k) {
                pl->n[k] = (real)f64[j][k];
                pl->s[k] = (real)f64[j][4 + k];
            }
            pl->h = (real)f64[j][3];
        }
        s->n_planes += nf;
    } else {
        /* an open mesh, one without a consistent winding, or a flat one is an
         * open solid: its triangles as given */
        int open = !closed || !orientable;
        /* corner j of triangle t in its consistent winding */
#define CORNER(t, j) m->indices[3 * (t) + (rev[t] && (j) ? 3 - (j) : (j))]
        double vol6 = 0.0;
        for (uint32_t t = 0; t < nt; ++t) {
            double q[3][3];
            for (int j = 0; j < 3; ++j) {
                for (int k = 0; k < 3; ++k) {
                    q[j][k] = x[3 * CORNER(t, j) + k] - c[k];
                }
            }
            vol6 += q[0][0] * (q[1][1] * q[2][2] - q[1][2] * q[2][1]) + q[0][1] * (q[1][2] * q[2][0] - q[1][0] * q[2][2]) +
                    q[0][2] * (q[1][0] * q[2][1] - q[1][1] * q[2][0]);
        }
        open = open || !(fabs(vol6) > tol * size * size);
        const int flip = !open && vol6 < 0.0;
        s->tris = (otri_t*)realloc(s->tris, (s->n_tris + nt) * sizeof(otri_t));
        v->plane_begin = s->n_tris;
        v->plane_count = nt;
        v->tri = 1;
        v->open = open;
        for (uint32_t t = 0; t < nt; ++t) {
            const double* p0 = x + 3 * CORNER(t, 0);
            const double* p1 = x + 3 * CORNER(t, flip ? 2 : 1);
            const double* p2 = x + 3 * CORNER(t, flip ? 1 : 2);
            const double a0 = p1[0] - p0[0], a1 = p1[1] - p0[1], a2 = p1[2] - p0[2];
            const double b0 = p2[0] - p0[0], b1 = p2[1] - p0[1], b2 = p2[2] - p0[2];
            const double g0 = a1 * b2 - a2 * b1, g1 = a2 * b0 - a0 * b2, g2 = a0 * b1 - a1 * b0;
            const double len = sqrt(g0 * g0 + g1 * g1 + g2 * g2);
            const double u[3] = {g0 / len, g1 / len, g2 / len};
            otri_t* tr = &s->tris[s->n_tris + t];
            for (int k = 0; k < 3; ++k) {
                tr->face.n[k] = (real)u[k];
                if (m->triangle_normals) {
                    const double* n = m->triangle_normals + 3 * t;
                    const double nl = sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
                    tr->face.s[k] = (real)(n[k] / nl);
                } else {
                    tr->face.s[k] = (real)u[k];
                }
                tr->v[0][k] = (real)p0[k];
                tr->v[1][k] = (real)p1[k];
                tr->v[2][k] = (real)p2[k];
            }
            tr->face.h = (real)(u[0] * p0[0] + u[1] * p0[1] + u[2] * p0[2]);
        }
#undef CORNER
        s->n_tris += nt;
    }
    free(rev);
    free(f64);
    for (int k = 0; k < 3; ++k) {
        v->lo[k] = (float)lo[k];
        if ((double)v->lo[k] > lo[k]) {
            v->lo[k] = nextafterf(v->lo[k], -INFINITY);
        }
        v->hi[k] = (float)hi[k];
        if ((double)v->hi[k] < hi[k]) {
            v->hi[k] = nextafterf(v->hi[k], INFINITY);
        }
    }
}

static oscene_t* build(const g4go_scene* in) {
    oscene_t* s = (oscene_t*)calloc(1, sizeof(oscene_t));
    s->n_mat = in->n_materials;
    s->n_surf = in->n_surfaces;
    s->n_vol = in->n_volumes;
    s->n_bind = in->n_bindings;
    s->mat = (material_t*)calloc(s->n_mat + 1, sizeof(material_t));
    s->surf = (surface_t*)calloc(s->n_surf + 1, sizeof(surface_t));
    s->vol = (volume_t*)calloc(s->n_vol + 1, sizeof(volume_t));
    s->bind = (uint32_t*)calloc(3 * s->n_bind + 3, sizeof(uint32_t));
    for (uint32_t i = 0; i < s->n_mat; ++i) {
        const g4go_material* m = &in->materials[i];
        material_t* d = &s->mat[i];
        d->rindex = make_prop(&m->rindex);
        d->groupvel = make_prop(&m->group_velocity_mm_per_ns);
        d->abslen = make_prop(&m->abs_length_mm);
        d->rayleigh = make_prop(&m->rayleigh_length_mm);
        for (int k = 0; k < 3; ++k) {
            d->spectrum[k] = make_spectrum(&m->scintillation_spectrum[k]);
        }
        d->rindex_max = RF(1.0);
        if (m->rindex.count) {
            float mx = m->rindex.values[0];
            for (uint32_t j = 1; j < m->rindex.count; ++j) {
                mx = m->rindex.values[j] > mx ? m->rindex.values[j] : mx;
            }
            d->rindex_max = mx;
        }
    }
    for (uint32_t i = 0; i < s->n_surf; ++i) {
        const g4go_surface* f = &in->surfaces[i];
        surface_t* d = &s->surf[i];
        d->type = f->type;
        d->model = f->model;
        d->finish = f->finish;
        d->model_value = f->model_value;
        d->refl = make_prop(&f->reflectivity);
        d->eff = make_prop(&f->efficiency);
        d->trans = make_prop(&f->transmittance);
        d->lobe = make_prop(&f->specular_lobe);
        d->spike = make_prop(&f->specular_spike);
        d->back = make_prop(&f->backscatter);
        d->rough = make_prop(&f->surface_roughness);
    }
    for (uint32_t i = 0; i < s->n_vol; ++i) {
        const g4go_volume* v = &in->volumes[i];
        volume_t* d = &s->vol[i];
        for (int k = 0; k < 3; ++k) {
            d->lo[k] = (float)((double)v->center_mm[k] - v->half_mm[k]);
            d->hi[k] = (float)((double)v->center_mm[k] + v->half_mm[k]);
        }
        d->id = v->volume_id;
        d->material = v->material_id;
        d->skin = v->skin_surface_id;
        d->sensor = v->sensor_id;
        for (uint32_t j = 0; j < in->n_meshes; ++j) {
            if (in->meshes[j].volume_id == v->volume_id) {
                mesh_solid(s, &in->meshes[j], d);
            }
        }
    }
    for (uint32_t i = 0; i < s->n_vol; ++i) {
        s->vol[i].parent = in->volumes[i].parent_volume_id == INVALID ? INVALID : vol_index(s, in->volumes[i].parent_volume_id);
    }
    s->world = vol_index(s, in->world_volume_id);
    for (uint32_t i = 0; i < s->n_bind; ++i) {
        s->bind[3 * i] = vol_index(s, in->bindings[i].from_volume_id);
        s->bind[3 * i + 1] = vol_index(s, in->bindings[i].to_volume_id);
        s->bind[3 * i + 2] = in->bindings[i].surface_id;
    }
    return s;
}

static void destroy(oscene_t* s) {
    for (uint32_t i = 0; i < s->n_mat; ++i) {
        free_prop(&s->mat[i].rindex);
        free_prop(&s->mat[i].groupvel);
        free_prop(&s->mat[i].abslen);
        free_prop(&s->mat[i].rayleigh);
        for (int k = 0; k < 3; ++k) {
            free_prop(&s->mat[i].spectrum[k]);
        }
    }
    for (uint32_t i = 0; i < s->n_surf; ++i) {
        surface_t* d = &s->surf[i];
        free_prop(&d->refl);
        free_prop(&d->eff);
        free_prop(&d->trans);
        free_prop(&d->lobe);
        free_prop(&d->spike);
        free_prop(&d->back);
        free_prop(&d->rough);
    }
    free(s->mat);
    free(s->surf);
    free(s->vol);
    free(s->bind);
    free(s->planes);
    free(s->tris);
    free(s);
}

/* SampleProperty: clamped linear interpolation, node found by linear scan. */
static real or_prop(const prop_t* p, real e, real fallback) {
    if (p->n == 0) {
        return fallback;
    }
    if (p->constant || e <= p->e[0]) {
        return p->v[0];
    }
    const uint32_t last = p->n - 1;
    if (e >= p->e[last]) {
        return p->v[last];
    }
    uint32_t i = 0;
    while (i + 1 < p->n && e > p->e[i + 1]) {
        ++i;
    }
    const real f = (e - p->e[i]) / (p->e[i + 1] - p->e[i]);
    return p->v[i] + f * (p->v[i + 1] - p->v[i]);
}

/* SampleSpectrum: inverse CDF by linear scan. */
static real or_spectrum(const prop_t* p, real random) {
    if (p->n == 0) {
        return NAN;
    }
    const real u = clamp01(random);
    if (p->n == 1 || u <= p->v[0]) {
        return p->e[0];
    }
    const uint32_t last = p->n - 1;
    if (u >= p->v[last]) {
        return p->e[last];
    }
    uint32_t i = 0;
    while (i + 1 < p->n && u > p->v[i + 1]) {
        ++i;
    }
    const real den = p->v[i + 1] - p->v[i];
    if (!(den > RF(0.0))) {
        return p->e[i];
    }
    const real f = (u - p->v[i]) / den;
    return p->e[i] + f * (p->e[i + 1] - p->e[i]);
}

/* FindSurfaceID / Scene::FindBoundarySurface. */
static uint32_t or_find_surface(const oscene_t* s, uint32_t from, uint32_t to) {
    for (uint32_t i = 0; i < s->n_bind; ++i) {
        if (s->bind[3 * i] == from && s->bind[3 * i + 1] == to) {
            return s->bind[3 * i + 2];
        }
    }
    if (from >= s->n_vol || to >= s->n_vol) {
        return INVALID;
    }
    const volume_t* f = &s->vol[from];
    const volume_t* t = &s->vol[to];
    if (t->parent == from) {
        return t->skin != INVALID ? t->skin : f->skin;
    }
    return f->skin != INVALID ? f->skin : t->skin;
}

/* Public wrapper for the surface-precedence fixtures (index space = input order). */
uint32_t ORACLE_NAME(find_surface)(const g4go_scene* scene, uint32_t from_id, uint32_t to_id) {
    oscene_t* s = build(scene);
    const uint32_t r = or_find_surface(s, vol_index(s, from_id), vol_index(s, to_id));
    destroy(s);
    return r;
}

/* ------------------------------------------------------------------ */
/* Photon state and geometry                                           */
/* ------------------------------------------------------------------ */
typedef struct {
    v3 pos, dir, pol;
    double time;
    real energy;
    uint32_t volume, step, photon, event, material;
    real rindex, inv_v, ilen, afrac;
} photon_t;

static int cache_material(photon_t* ph, const oscene_t* s, uint32_t m) {
    if (m >= s->n_mat) {
        return 0;
    }
    const material_t* mt = &s->mat[m];
    const real n = or_prop(&mt->rindex, ph->energy, NAN);
    const int n_ok = finite_r(n) && n > RF(0.0);
    const real v = or_prop(&mt->groupvel, ph->energy, n_ok ? RF(299.792458) / n : NAN);
    if (!(v > RF(0.0)) || !finite_r(v)) {
        return 0;
    }
    const real la = or_prop(&mt->abslen, ph->energy, R_INF);
    const real lr = or_prop(&mt->rayleigh, ph->energy, R_INF);
    const int a_ok = finite_r(la) && la > RF(0.0);
    const int r_ok = finite_r(lr) && lr > RF(0.0);
    ph->material = m;
    ph->rindex = n;
    ph->inv_v = RF(1.0) / v;
    if (a_ok && r_ok) {
        ph->ilen = (la * lr) / (la + lr);
        ph->afrac = lr / (la + lr);
    } else if (a_ok) {
        ph->ilen = la;
        ph->afrac = RF(1.0);
    } else if (r_ok) {
        ph->ilen = lr;
        ph->afrac = RF(0.0);
    } else {
        ph->ilen = R_INF;
        ph->afrac = RF(1.0);
    }
    return 1;
}

/* Exit face of a box: the smallest (distance to plane)/|d| chosen with cross
 * products, x then y then z, strict comparisons. */
static void box_exit(const volume_t* v, v3 p, v3 d, real* dist, int* axis) {
    const real ax = R_FABS(d.x), ay = R_FABS(d.y), az = R_FABS(d.z);
    const real nx = d.x > RF(0.0) ? R_FMAX((real)v->hi[0] - p.x, RF(0.0))
                                  : (d.x < RF(0.0) ? R_FMAX(p.x - (real)v->lo[0], RF(0.0)) : RF(1.0));
    const real ny = d.y > RF(0.0) ? R_FMAX((real)v->hi[1] - p.y, RF(0.0))
                                  : (d.y < RF(0.0) ? R_FMAX(p.y - (real)v->lo[1], RF(0.0)) : RF(1.0));
    const real nz = d.z > RF(0.0) ? R_FMAX((real)v->hi[2] - p.z, RF(0.0))
                                  : (d.z < RF(0.0) ? R_FMAX(p.z - (real)v->lo[2], RF(0.0)) : RF(1.0));
    int best = 0;
    real bn = nx, ba = ax;
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
    *axis = best;
    *dist = bn / ba;
}

static real box_entry(const volume_t* b, v3 p, v3 inv, real tmin, int* axis) {
    const real pc[3] = {p.x, p.y, p.z};
    const real ic[3] = {inv.x, inv.y, inv.z};
    real tn_max = -R_INF, tf_min = R_INF;
    int a = 0;
    for (int k = 0; k < 3; ++k) {
        if (finite_r(ic[k])) {
            const real t1 = ((real)b->lo[k] - pc[k]) * ic[k];
            const real t2 = ((real)b->hi[k] - pc[k]) * ic[k];
            const real tn = R_FMIN(t1, t2);
            const real tf = R_FMAX(t1, t2);
            if (tn > tn_max) {
                tn_max = tn;
                a = k;
            }
            tf_min = R_FMIN(tf_min, tf);
        } else if (pc[k] < (real)b->lo[k] || pc[k] > (real)b->hi[k]) {
            return R_INF;
        }
    }
    if (tn_max >= tmin && tn_max <= tf_min) {
        *axis = a;
        return tn_max;
    }
    return R_INF;
}

static int box_contains(const volume_t* v, v3 q) {
    return q.x >= (real)v->lo[0] && q.x <= (real)v->hi[0] && q.y >= (real)v->lo[1] && q.y <= (real)v->hi[1] &&
           q.z >= (real)v->lo[2] && q.z <= (real)v->hi[2];
}

/* Convex mesh solids by clipping against their face planes: the exit is the
 * smallest (h - n.p)/(n.d) over the planes with n.d > 0, the entry the largest
 * over n.d < 0 if it does not pass the smallest exit; candidates compared by
 * cross products (first plane kept on ties), one division at the end.  For a
 * closed convex mesh this is the closest triangle hit of the reference's
 * trace (OptiXTransportDevice.cu:105-196). */
static void mesh_exit(const oscene_t* s, const volume_t* v, v3 p, v3 d, real* dist, uint32_t* plane) {
    real bn = RF(1.0), ba = RF(0.0);
    uint32_t best = INVALID;
    for (uint32_t k = v->plane_begin; k < v->plane_begin + v->plane_count; ++k) {
        const plane_t* pl = &s->planes[k];
        const v3 n = mk(pl->n[0], pl->n[1], pl->n[2]);
        const real nd = vdot(n, d);
        if (nd > RF(0.0)) {
            const real num = R_FMAX(pl->h - vdot(n, p), RF(0.0));
            if (num * ba < bn * nd) {
                best = k;
                bn = num;
                ba = nd;
            }
        }
    }
    *plane = best;
    *dist = best == INVALID ? R_INF : bn / ba;
}

static real mesh_entry(const oscene_t* s, const volume_t* v, v3 p, v3 d, real tmin, uint32_t* plane) {
    real a = RF(-1.0), b = RF(0.0), c = RF(1.0), e = RF(0.0);
    uint32_t best = INVALID;
    for (uint32_t k = v->plane_begin; k < v->plane_begin + v->plane_count; ++k) {
        const plane_t* pl = &s->planes[k];
        const v3 n = mk(pl->n[0], pl->n[1], pl->n[2]);
        const real nd = vdot(n, d);
        const real num = pl->h - vdot(n, p);
        if (nd < RF(0.0)) {
            if (-num * b > a * -nd) {
                best = k;
                a = -num;
                b = -nd;
            }
        } else if (nd > RF(0.0)) {
            if (num * e < c * nd) {
                c = num;
                e = nd;
            }
        } else if (num < RF(0.0)) {
            return R_INF;
        }
    }
    if (best == INVALID) {
        return R_INF;
    }
    const real t = a / b;
    if (!(t >= tmin) || (e != RF(0.0) && t > c / e)) {
        return R_INF;
    }
    *plane = best;
    return t;
}

static int mesh_contains(const oscene_t* s, const volume_t* v, v3 q) {
    for (uint32_t k = v->plane_begin; k < v->plane_begin + v->plane_count; ++k) {
        const plane_t* pl = &s->planes[k];
        if (vdot(mk(pl->n[0], pl->n[1], pl->n[2]), q) > pl->h) {
            return 0;
        }
    }
    return 1;
}

/* Non-convex mesh solids as their triangles: the watertight ray/triangle test
 * of Woop, Benthin and Wald (JCGT 2(1), 2013): the ray sheared onto its
 * dominant axis, the edge functions in 2D (in FP64 when one is exactly
 * zero); det < 0 marks a hit leaving the solid (outward winding).  Closest
 * hits by brute force in mesh order (ties: the lower index).  A point is
 * inside when the first triangle along (1, 2, 3)/sqrt(14) is a leaving one.
 * The reference traces the triangles themselves (OptiXTransportDevice.cu:
 * 105-196). */
#define FACE_TRI 0x80000000u
static real ax3(v3 a, int k) { return k == 0 ? a.x : (k == 1 ? a.y : a.z); }

static int wt_hit(v3 o, v3 d, const otri_t* tr, real* t, int* leaving) {
    const real adx = R_FABS(d.x), ady = R_FABS(d.y), adz = R_FABS(d.z);
    const int kz = (adx >= ady && adx >= adz) ? 0 : (ady >= adz ? 1 : 2);
    int kx = kz == 2 ? 0 : kz + 1;
    int ky = kx == 2 ? 0 : kx + 1;
    const real dz = ax3(d, kz);
    if (dz < RF(0.0)) {
        const int k = kx;
        kx = ky;
        ky = k;
    }
    const real sx = ax3(d, kx) / dz, sy = ax3(d, ky) / dz, sz = RF(1.0) / dz;
    const v3 a = vsub(mk(tr->v[0][0], tr->v[0][1], tr->v[0][2]), o);
    const v3 b = vsub(mk(tr->v[1][0], tr->v[1][1], tr->v[1][2]), o);
    const v3 c = vsub(mk(tr->v[2][0], tr->v[2][1], tr->v[2][2]), o);
    const real az = ax3(a, kz), bz = ax3(b, kz), cz = ax3(c, kz);
    const real ax_ = ax3(a, kx) - sx * az, ay_ = ax3(a, ky) - sy * az;
    const real bx = ax3(b, kx) - sx * bz, by = ax3(b, ky) - sy * bz;
    const real cx = ax3(c, kx) - sx * cz, cy = ax3(c, ky) - sy * cz;
    real u = cx * by - cy * bx;
    real vv = ax_ * cy - ay_ * cx;
    real w = bx * ay_ - by * ax_;
    if (u == RF(0.0) || vv == RF(0.0) || w == RF(0.0)) {
        u = (real)((double)cx * (double)by - (double)cy * (double)bx);
        vv = (real)((double)ax_ * (double)cy - (double)ay_ * (double)cx);
        w = (real)((double)bx * (double)ay_ - (double)by * (double)ax_);
    }
    if ((u < RF(0.0) || vv < RF(0.0) || w < RF(0.0)) && (u > RF(0.0) || vv > RF(0.0) || w > RF(0.0))) {
        return 0;
    }
    const real det = u + vv + w;
    if (det == RF(0.0)) {
        return 0;
    }
    *t = (u * (sz * az) + vv * (sz * bz) + w * (sz * cz)) / det;
    *leaving = det < RF(0.0);
    return 1;
}

/* mode: 0 leaving, 1 entering, 2 either; returns R_INF and *face INVALID on none */
static real tri_closest(const oscene_t* s, const volume_t* v, v3 p, v3 d, real tmin, int mode, uint32_t* face,
                        int* leaving) {
    real best = R_INF;
    uint32_t bi = INVALID;
    int bl = 0;
    for (uint32_t k = v->plane_begin; k < v->plane_begin + v->plane_count; ++k) {
        real t;
        int l;
        if (wt_hit(p, d, &s->tris[k], &t, &l) && t >= tmin && (mode == 2 || l == (mode == 0)) && t < best) {
            best = t;
            bi = k;
            bl = l;
        }
    }
    *face = bi == INVALID ? INVALID : (bi | FACE_TRI);
    *leaving = bl;
    return best;
}

/* fallback exit when no triangle ahead is a leaving hit: among the triangles
 * with n.d > 0 the one whose plane is nearest the point (ties: lower index),
 * at max(h - n.p, 0) / n.d */
static real tri_nearest_leaving(const oscene_t* s, const volume_t* v, v3 p, v3 d, uint32_t* face) {
    real best_gap = R_INF, best_num = RF(0.0), best_nd = RF(1.0);
    uint32_t bi = INVALID;
    for (uint32_t k = v->plane_begin; k < v->plane_begin + v->plane_count; ++k) {
        const plane_t* f = &s->tris[k].face;
        const v3 n = mk(f->n[0], f->n[1], f->n[2]);
        const real nd = vdot(n, d);
        if (nd > RF(0.0)) {
            const real num = f->h - vdot(n, p);
            const real gap = R_FABS(num);
            if (gap < best_gap) {
                best_gap = gap;
                best_num = num;
                best_nd = nd;
                bi = k;
            }
        }
    }
    *face = bi == INVALID ? INVALID : (bi | FACE_TRI);
    return bi == INVALID ? R_INF : R_FMAX(best_num, RF(0.0)) / best_nd;
}

static int tri_contains(const oscene_t* s, const volume_t* v, v3 q) {
    if (v->open) {
        return 0; /* an open solid holds no point */
    }
    uint32_t f;
    int l;
    const v3 u = mk((real)0.26726124f, (real)0.53452248f, (real)0.80178373f);
    const real t = tri_closest(s, v, q, u, RF(0.0), 2, &f, &l);
    return t < R_INF && l;
}

static int solid_contains(const oscene_t* s, const volume_t* v, v3 q) {
    return v->tri ? tri_contains(s, v, q) : mesh_contains(s, v, q);
}

static const plane_t* face_record(const oscene_t* s, uint32_t face) {
    return (face & FACE_TRI) ? &s->tris[face & ~FACE_TRI].face : &s->planes[face];
}

/* Nearest daughter entry by brute force; ties go to the lower index.  *plane
 * is the entered mesh face, INVALID for a box face. */
static real nearest_daughter(const oscene_t* s, uint32_t cur, v3 p, v3 d, real tmin, uint32_t* which, int* axis,
                             uint32_t* plane) {
    const v3 inv = mk(RF(1.0) / d.x, RF(1.0) / d.y, RF(1.0) / d.z);
    real best = R_INF;
    *which = INVALID;
    for (uint32_t j = 0; j < s->n_vol; ++j) {
        if (s->vol[j].parent != cur) {
            continue;
        }
        int a = 0;
        uint32_t f = INVALID;
        int l;
        const real t = s->vol[j].tri ? tri_closest(s, &s->vol[j], p, d, tmin, s->vol[j].open ? 2 : 1, &f, &l)
                       : s->vol[j].plane_count ? mesh_entry(s, &s->vol[j], p, d, tmin, &f)
                                               : box_entry(&s->vol[j], p, inv, tmin, &a);
        if (t < best || (t == best && t < R_INF && j < *which)) {
            best = t;
            *which = j;
            *axis = a;
            *plane = f;
        }
    }
    return best;
}

static uint32_t point_neighbor(const oscene_t* s, uint32_t cur, v3 q) {
    const uint32_t parent = s->vol[cur].parent;
    if (parent == INVALID) {
        return INVALID;
    }
    for (uint32_t j = 0; j < s->n_vol; ++j) {
        if (j != cur && s->vol[j].parent == parent && box_contains(&s->vol[j], q) &&
            (s->vol[j].plane_count == 0 || solid_contains(s, &s->vol[j], q))) {
            return j;
        }
    }
    return parent;
}

/* ------------------------------------------------------------------ */
/* Surface sampling (OptiXTransportDevice.cu:492-595)                  */
/* ------------------------------------------------------------------ */
static v3 or_facet(const surface_t* sf, v3 n, v3 incoming, const rng_t* r, const photon_t* ph) {
    n = or_normalize(n);
    uint32_t w[4];
    if (sf->model == 0) {
        const real polish = clamp01(sf->model_value);
        if (polish >= RF(1.0)) {
            return n;
        }
        for (uint32_t a = 0; a < 128u; ++a) {
            block4(r, ph->photon, ph->event, ph->step, P_FACET, a, w);
            const v3 smear = mk(RF(2.0) * u01(w[0]) - RF(1.0), RF(2.0) * u01(w[1]) - RF(1.0), RF(2.0) * u01(w[2]) - RF(1.0));
            if (vdot(smear, smear) > RF(1.0)) {
                continue;
            }
            const v3 facet = or_normalize(vadd(n, vscale(smear, RF(1.0) - polish)));
            if (vdot(incoming, facet) < RF(0.0)) {
                return facet;
            }
        }
        return n;
    }
    if (sf->model != 1) {
        return n;
    }
    const real sigma = R_FMAX(sf->model_value, RF(0.0));
    if (!(sigma > RF(0.0))) {
        return n;
    }
    const v3 helper = R_FABS(n.z) < RF(0.9) ? mk(RF(0.0), RF(0.0), RF(1.0)) : mk(RF(1.0), RF(0.0), RF(0.0));
    const v3 t = or_normalize(vcross(n, helper));
    const v3 b = or_normalize(vcross(n, t));
    const real s2 = sigma * sigma;
    for (uint32_t a = 0; a < 256u; ++a) {
        block4(r, ph->photon, ph->event, ph->step, P_FACET, a, w);
        const real quarter = RF(0.25) * u01(w[0]);
        const real alpha = RF(1.57079632679489662) * u01(w[0]);
        real sa, ca;
        r_sincos_turn(quarter, &sa, &ca);
        const real g = r_exp(RF(-0.5) * alpha * alpha / s2);
        if (u01(w[1]) > g * sa) {
            continue;
        }
        real sp, cp;
        r_sincos_turn(u01(w[2]), &sp, &cp);
        const v3 facet = or_normalize(vadd(vscale(n, ca), vadd(vscale(t, sa * cp), vscale(b, sa * sp))));
        if (vdot(incoming, facet) < RF(0.0)) {
            return facet;
        }
    }
    return n;
}

/* branch (when not NULL): 0 spike, 1 lobe, 2 backscatter, 3 Lambertian. */
static v3 or_unified(const surface_t* sf, v3 d, v3 n, const rng_t* r, const photon_t* ph, int* branch) {
    const int ground = sf->finish == 1;
    const v3 facet = or_facet(sf, n, d, r, ph);
    const real spike = clamp01(or_prop(&sf->spike, ph->energy, ground ? RF(0.0) : RF(1.0)));
    const real lobe = clamp01(or_prop(&sf->lobe, ph->energy, RF(0.0)));
    const real back = clamp01(or_prop(&sf->back, ph->energy, RF(0.0)));
    uint32_t w[4];
    block4(r, ph->photon, ph->event, ph->step, P_LOBE, 0, w);
    const real u = u01(w[0]);
    const int b = u < spike ? 0 : (u < spike + lobe ? 1 : (u < spike + lobe + back ? 2 : 3));
    if (branch != NULL) {
        *branch = b;
    }
    if (b == 0) {
        return or_reflect_dir(d, n);
    }
    if (b == 1) {
        return or_reflect_dir(d, facet);
    }
    if (b == 2) {
        return vscale(d, RF(-1.0));
    }
    return or_lambertian(n, u01(w[1]), u01(w[2]));
}

static v3 or_surface_reflection(const surface_t* sf, v3 d, v3 n, const rng_t* r, const photon_t* ph, int* branch) {
    if (sf->finish == 0) {
        return or_reflect_dir(d, n);
    }
    if (sf->model == 0) {
        return or_reflect_dir(d, or_facet(sf, n, d, r, ph));
    }
    return or_unified(sf, d, n, r, ph, branch);
}

/* Dipole scattering about the old polarisation p (Geant4 G4OpRayleigh's
 * distribution), sampled directly from one block: c = p.k' is the median of
 * three uniforms on (-1, 1] (density 3/4 (1 - c^2)), taken as u01 of the
 * median word; the azimuth about p is uniform; the polarisation sign is the
 * parity of the order of the three words. */
static void or_rayleigh(photon_t* ph, const rng_t* r) {
    const v3 op = ph->pol;
    uint32_t w[4];
    block4(r, ph->photon, ph->event, ph->step, P_RAYLEIGH, 0, w);
    const uint32_t lo = w[0] < w[1] ? w[0] : w[1];
    const uint32_t hi = w[0] < w[1] ? w[1] : w[0];
    const uint32_t hz = hi < w[2] ? hi : w[2];
    const uint32_t med = lo < hz ? hz : lo;
    const int flip = ((w[0] < w[1]) != (w[1] < w[2])) != (w[0] < w[2]);
    const real c = RF(2.0) * u01(med) - RF(1.0);
    const real s = R_SQRT(R_FMAX(RF(0.0), RF(1.0) - c * c));
    real sp, cp;
    r_sincos_turn(u01(w[3]), &sp, &cp);
    const v3 nd = or_rotate_uz(mk(s * cp, s * sp, c), op);
    v3 np = vsub(op, vscale(nd, vdot(nd, op)));
    const real l2 = vdot(np, np);
    if (!(l2 > RF(0.0))) {
        real sq, cq;
        r_sincos_turn((real)(w[3] & 0x00ffffffu) * RF(0x1p-24), &sq, &cq);
        np = or_rotate_uz(mk(cq, sq, RF(0.0)), nd);
    } else {
        np = vscale(np, RF(1.0) / R_SQRT(l2));
        if (flip) {
            np = vscale(np, RF(-1.0));
        }
    }
    ph->dir = nd;
    ph->pol = np;
}

/* ------------------------------------------------------------------ */
/* Photon generation                                                   */
/* ------------------------------------------------------------------ */
static void isotropic(real ct, real sp, real cp, real spsi, real cpsi, v3* dir, v3* pol) {
    const real st = R_SQRT(R_FMAX(RF(0.0), RF(1.0) - ct * ct));
    *dir = mk(st * cp, st * sp, ct);
    const v3 et = mk(ct * cp, ct * sp, -st);
    const v3 ep = mk(-sp, cp, RF(0.0));
    *pol = vadd(vscale(et, cpsi), vscale(ep, spsi));
}

static void gen_synthetic(photon_t* ph, const g4go_synthetic_source* src, uint32_t volume, const rng_t* r,
                          uint32_t event, uint32_t photon) {
    memset(ph, 0, sizeof(*ph));
    ph->photon = photon;
    ph->event = event;
    ph->volume = volume;
    ph->material = INVALID;
    uint32_t a[4], b[4], c[4];
    block4(r, photon, event, 0, P_SYNTH, 0, a);
    block4(r, photon, event, 0, P_SYNTH, 1, b);
    const real ct = RF(1.0) - RF(2.0) * u01(a[0]);
    real sp, cp, spsi, cpsi, sg, cg;
    r_sincos_turn(u01(a[1]), &sp, &cp);
    r_sincos_turn(u01(a[2]), &spsi, &cpsi);
    isotropic(ct, sp, cp, spsi, cpsi, &ph->dir, &ph->pol);
    const real radius = R_SQRT(RF(-2.0) * r_log(u01(a[3])));
    r_sincos_turn(u01(b[0]), &sg, &cg);
    const real wl = (real)src->wavelength_mean_nm + (real)src->wavelength_sigma_nm * (radius * cg);
    ph->energy = RF(1239.84198) / wl;
    ph->pos = mk((real)src->lo_mm[0] + ((real)src->hi_mm[0] - (real)src->lo_mm[0]) * u01(b[1]),
                 (real)src->lo_mm[1] + ((real)src->hi_mm[1] - (real)src->lo_mm[1]) * u01(b[2]),
                 (real)src->lo_mm[2] + ((real)src->hi_mm[2] - (real)src->lo_mm[2]) * u01(b[3]));
    block4(r, photon, event, 0, P_SYNTH, 2, c);
    ph->time = (double)(-(real)src->decay_time_ns * r_log(u01(c[0])));
}

static void gen_emission(photon_t* ph, const g4go_emission* em, uint32_t volume, const oscene_t* s, const rng_t* r,
                         uint32_t local) {
    memset(ph, 0, sizeof(*ph));
    ph->photon = em->first_photon_id + local;
    ph->event = em->event_id;
    ph->volume = volume;
    ph->material = INVALID;
    ph->pos = mk(em->position_mm[0], em->position_mm[1], em->position_mm[2]);
    ph->time = (double)em->time_ns;
    ph->dir = or_normalize(mk(em->direction[0], em->direction[1], em->direction[2]));
    ph->pol = mk(em->polarization[0], em->polarization[1], em->polarization[2]);
    ph->energy = em->energy_ev;
    if (em->type == 0) {
        return;
    }
    if (em->material_id >= s->n_mat) {
        ph->energy = NAN;
        return;
    }
    const material_t* m = &s->mat[em->material_id];
    uint32_t w0[4], w1[4], w[4];
    real frac = RF(1.0);
    if (em->type == 1) {
        const real beta = ((real)em->pre_velocity_mm_per_ns + RF(0.5) * (real)em->delta_velocity_mm_per_ns) / RF(299.792458);
        if (!(beta > RF(0.0)) || !(beta < RF(1.0)) || m->rindex.n == 0) {
            ph->energy = NAN;
            return;
        }
        const real emin = m->rindex.e[0], emax = m->rindex.e[m->rindex.n - 1];
        const real nmax = m->rindex_max;
        const real binv = RF(1.0) / beta;
        const real maxcos = binv / nmax;
        const real maxsin2 = RF(1.0) - maxcos * maxcos;
        const real de = emax - emin;
        if (!(nmax > RF(0.0)) || !finite_r(nmax) || !(maxcos > RF(0.0)) || !(maxcos < RF(1.0)) || !(de >= RF(0.0)) ||
            !finite_r(de) || !(maxsin2 > RF(0.0))) {
            ph->energy = NAN;
            return;
        }
        real energy = emin, ct = RF(0.0), s2 = RF(0.0);
        int ok = 0;
        for (uint32_t a = 0; a < 1024u; ++a) {
            block4(r, ph->photon, ph->event, 0, P_EMISSION, 16u + a, w);
            energy = emin + u01(w[0]) * de;
            const real n = or_prop(&m->rindex, energy, NAN);
            ct = binv / n;
            s2 = RF(1.0) - ct * ct;
            if (finite_r(n) && ct > RF(0.0) && ct < RF(1.0) && s2 > RF(0.0) && u01(w[1]) * maxsin2 <= s2) {
                ok = 1;
                break;
            }
        }
        if (!ok) {
            ph->energy = NAN;
            return;
        }
        block4(r, ph->photon, ph->event, 0, P_EMISSION, 0, w0);
        real sp, cp;
        r_sincos_turn(u01(w0[2]), &sp, &cp);
        const real st = R_SQRT(R_FMAX(s2, RF(0.0)));
        const v3 axis = or_normalize(mk(em->direction[0], em->direction[1], em->direction[2]));
        ph->dir = or_normalize(or_rotate_uz(mk(st * cp, st * sp, ct), axis));
        ph->pol = or_normalize(or_rotate_uz(mk(ct * cp, ct * sp, -st), axis));
        const real maxmean = R_FMAX((real)em->pre_mean_photon_count, (real)em->post_mean_photon_count);
        const real dmean = (real)em->pre_mean_photon_count - (real)em->post_mean_photon_count;
        for (uint32_t a = 0; a < 1024u; ++a) {
            block4(r, ph->photon, ph->event, 0, P_EMISSION, 2048u + a, w);
            frac = u01(w[0]);
            if (!(u01(w[1]) * maxmean > (real)em->pre_mean_photon_count - frac * dmean)) {
                break;
            }
        }
        ph->energy = energy;
    } else if (em->type == 2) {
        if (em->spectrum_id >= 3u) {
            ph->energy = NAN;
            return;
        }
        block4(r, ph->photon, ph->event, 0, P_EMISSION, 0, w0);
        block4(r, ph->photon, ph->event, 0, P_EMISSION, 1, w1);
        ph->energy = or_spectrum(&m->spectrum[em->spectrum_id], u01(w0[0]));
        const real ct = RF(1.0) - RF(2.0) * u01(w0[1]);
        real sp, cp, spsi, cpsi;
        r_sincos_turn(u01(w0[2]), &sp, &cp);
        r_sincos_turn(u01(w0[3]), &spsi, &cpsi);
        isotropic(ct, sp, cp, spsi, cpsi, &ph->dir, &ph->pol);
        frac = em->charge == 0.0f ? RF(1.0) : u01(w1[0]);
        const real decay = em->decay_time_ns;
        if (decay > RF(0.0) && finite_r(decay)) {
            if (!(em->rise_time_ns > 0.0f) || !isfinite(em->rise_time_ns)) {
                ph->time += (double)(-decay * r_log(R_FMAX(u01(w1[1]), RF(1.0e-7))));
            } else {
                real sample = RF(0.0);
                for (uint32_t a = 0; a < 1024u; ++a) {
                    block4(r, ph->photon, ph->event, 0, P_EMISSION, 2u + a, w);
                    sample = -decay * r_log(R_FMAX(RF(1.0e-7), RF(1.0) - u01(w[0])));
                    if (!(u01(w[1]) > RF(1.0) - r_exp(-sample / (real)em->rise_time_ns))) {
                        break;
                    }
                }
                ph->time += (double)sample;
            }
        }
    } else {
        ph->energy = NAN;
        return;
    }
    ph->pos = vadd(mk(em->position_mm[0], em->position_mm[1], em->position_mm[2]),
                   vscale(mk(em->step_delta_mm[0], em->step_delta_mm[1], em->step_delta_mm[2]), frac));
    const real vel = (real)em->pre_velocity_mm_per_ns + RF(0.5) * frac * (real)em->delta_velocity_mm_per_ns;
    if (vel > RF(0.0) && finite_r(vel)) {
        ph->time += (double)(frac * (real)em->step_length_mm / vel);
    }
}

static int start_photon(photon_t* ph) {
    const real l2 = vdot(ph->dir, ph->dir);
    if (!(ph->energy > RF(0.0)) || !finite_r(ph->energy) || !(l2 > RF(0.0)) || !finite_r(l2)) {
        return 0;
    }
    ph->dir = or_normalize(ph->dir);
    ph->pol = or_project(ph->pol, ph->dir);
    return 1;
}

/* ------------------------------------------------------------------ */
/* One step of the bounce loop                                         */
/* ------------------------------------------------------------------ */
enum { S_ALIVE = 0, S_DETECTED, S_ABS_BULK, S_ABS_SURFACE, S_ESCAPED, S_INVALID };

/* Per-step record for the likelihood-ratio replay. */
typedef struct {
    real segment;
    uint32_t material, surface;
    int event; /* 0 none, 1 Rayleigh, 2 interaction, 3 direct transmit, 4 detect */
} trace_t;

static int or_step(photon_t* ph, const oscene_t* s, const rng_t* r, real eps, int* zero, uint32_t* sensor, trace_t* tr) {
    *zero = 0;
    if (ph->volume >= s->n_vol) {
        return S_INVALID;
    }
    const volume_t* vol = &s->vol[ph->volume];
    if (vol->material != ph->material && !cache_material(ph, s, vol->material)) {
        return S_INVALID;
    }
    tr->segment = RF(0.0);
    tr->material = ph->material;
    tr->surface = INVALID;
    tr->event = 0;
    /* with mesh solids in the scene: faces by clipping, no snap onto a mesh
     * face, and a daughter entered up to eps beyond the exit wins (the
     * reference's coincident-surface rank, OptiXTransportDevice.cu:105-127) */
    const int meshes = s->n_planes != 0 || s->n_tris != 0;
    real dist;
    int axis = 0;
    uint32_t plane = INVALID;
    int passed_out = 0; /* in an open solid with none of its triangles ahead */
    if (vol->tri && vol->open) {
        /* any of its triangles beyond eps, from either side (the reference's
         * trace, tmin = boundaryEpsilonMm, OptiXTransportDevice.cu:742-746) */
        int l;
        dist = tri_closest(s, vol, ph->pos, ph->dir, eps, 2, &plane, &l);
        passed_out = !finite_r(dist);
    } else if (vol->tri) {
        int l;
        dist = tri_closest(s, vol, ph->pos, ph->dir, RF(0.0), 0, &plane, &l);
        if (!finite_r(dist)) {
            /* rounding left the photon a hair outside its solid at an edge,
             * moving out past it: it leaves through the nearest face it moves
             * away from */
            dist = tri_nearest_leaving(s, vol, ph->pos, ph->dir, &plane);
            if (!finite_r(dist)) {
                return S_INVALID;
            }
        }
    } else if (vol->plane_count) {
        mesh_exit(s, vol, ph->pos, ph->dir, &dist, &plane);
    } else {
        box_exit(vol, ph->pos, ph->dir, &dist, &axis);
    }
    uint32_t next = INVALID, surface = INVALID;
    int entering = 0;
    {
        uint32_t which, p2 = INVALID;
        int a2 = 0;
        const real t = nearest_daughter(s, ph->volume, ph->pos, ph->dir, eps, &which, &a2, &p2);
        if (which != INVALID && (meshes ? t <= dist + eps : t <= dist)) {
            dist = t;
            axis = a2;
            plane = p2;
            next = which;
            surface = or_find_surface(s, ph->volume, which);
            entering = 1;
        }
    }
    if (!finite_r(dist)) {
        if (passed_out && vol->parent != INVALID) {
            /* out of the open solid into its parent where the photon is */
            ph->volume = vol->parent;
            *zero = 1;
            return S_ALIVE;
        }
        return S_ESCAPED;
    }
    if (dist <= eps) {
        *zero = 1;
    }
    uint32_t w[4];
    block4(r, ph->photon, ph->event, ph->step, P_STEP, 0, w);
    if (finite_r(ph->ilen)) {
        const real dint = -r_log(u01(w[0])) * ph->ilen;
        if (dint < dist) {
            ph->time += (double)(dint * ph->inv_v);
            tr->segment = dint;
            if (u01(w[1]) <= ph->afrac) {
                return S_ABS_BULK;
            }
            ph->pos = vadd(ph->pos, vscale(ph->dir, dint));
            or_rayleigh(ph, r);
            tr->event = 1;
            return S_ALIVE;
        }
    }
    ph->time += (double)(dist * ph->inv_v);
    tr->segment = dist;
    v3 bpos = vadd(ph->pos, vscale(ph->dir, dist));
    v3 normal;
    const int mesh_face = plane != INVALID;
    if (mesh_face) {
        const plane_t* pl = face_record(s, plane);
        const v3 sn = mk(pl->s[0], pl->s[1], pl->s[2]);
        normal = vdot(ph->dir, sn) > RF(0.0) ? vscale(sn, RF(-1.0)) : sn;
    } else {
        const real dsign = vaxis(ph->dir, axis) > RF(0.0) ? RF(1.0) : RF(-1.0);
        const volume_t* fb = entering ? &s->vol[next] : vol;
        const real fplane = entering ? (dsign > RF(0.0) ? (real)fb->lo[axis] : (real)fb->hi[axis])
                                     : (dsign > RF(0.0) ? (real)fb->hi[axis] : (real)fb->lo[axis]);
        if (axis == 0) {
            bpos.x = fplane;
        } else if (axis == 1) {
            bpos.y = fplane;
        } else {
            bpos.z = fplane;
        }
        normal = axis == 0 ? mk(-dsign, RF(0.0), RF(0.0)) : (axis == 1 ? mk(RF(0.0), -dsign, RF(0.0)) : mk(RF(0.0), RF(0.0), -dsign));
    }
    ph->pos = bpos;
    if (!entering) {
        /* the point just outside the face (an open solid's triangle: on the
         * side of its normal the photon moves to) */
        v3 q;
        if (mesh_face) {
            const plane_t* fr = face_record(s, plane);
            const v3 pn = mk(fr->n[0], fr->n[1], fr->n[2]);
            q = vadd(bpos, vscale(pn, vol->open && vdot(pn, ph->dir) < RF(0.0) ? -eps : eps));
        } else {
            q = vadd(bpos, vscale(normal, -eps));
        }
        next = point_neighbor(s, ph->volume, q);
        if (next == INVALID) {
            return ph->volume == s->world ? S_ESCAPED : S_INVALID;
        }
        surface = or_find_surface(s, ph->volume, next);
    }
    const volume_t* nvol = &s->vol[next];
    if (nvol->material >= s->n_mat) {
        return S_INVALID;
    }
    const material_t* nmat = &s->mat[nvol->material];
    const real n1 = ph->rindex;
    const real e = ph->energy;
    int reflected = 0, detected = 0, absorbed = 0, transmitted = 0;
    v3 dir = ph->dir, pol = ph->pol;
    if (surface < s->n_surf) {
        const surface_t* sf = &s->surf[surface];
        tr->surface = surface;
        const real refl = clamp01(or_prop(&sf->refl, e, RF(0.0)));
        const real eff = clamp01(or_prop(&sf->eff, e, RF(0.0)));
        const real trans = clamp01(or_prop(&sf->trans, e, RF(0.0)));
        const int outcome = ORACLE_NAME(classify)(refl, trans, sf->refl.n != 0, sf->trans.n != 0, u01(w[2]));
        tr->event = outcome == O_INTERACT ? 2 : (outcome == O_TRANSMIT ? 3 : 0);
        if (outcome == O_ABSORB) {
            absorbed = 1;
            detected = ORACLE_NAME(absorption)(eff, nvol->sensor != INVALID || vol->sensor != INVALID, u01(w[3])) == O_DETECT;
        } else if (outcome == O_TRANSMIT) {
            transmitted = 1;
        } else if (sf->type == 0) {
            const real n2 = or_prop(&nmat->rindex, e, NAN);
            if (!(n1 > RF(0.0)) || !finite_r(n1) || !(n2 > RF(0.0)) || !finite_r(n2)) {
                absorbed = 1;
            } else {
                const v3 inormal = sf->finish == 0 ? or_normalize(normal) : or_facet(sf, normal, dir, r, ph);
                const real rough = or_prop(&sf->rough, e, RF(0.0));
                int rough_pass = 1;
                uint32_t wr[4] = {0, 0, 0, 0};
                if (rough > RF(0.0) && n1 > n2 && finite_r(rough) && finite_r(e) && e > RF(0.0)) {
                    const real cost1 = R_FMAX(RF(0.0), -vdot(dir, inormal));
                    const real wl = RF(1.239841984e-3) / e;
                    const real expo = RF(4.0) * RF(3.14159265358979323846) * rough * n1 * cost1 / wl;
                    block4(r, ph->photon, ph->event, ph->step, P_ROUGH, 0, wr);
                    rough_pass = u01(wr[0]) < r_exp(-expo * expo);
                }
                const fresnel_t f = or_fresnel(dir, pol, inormal, n1, n2);
                reflected = f.tir || (double)u01(w[3]) > f.t;
                if (reflected) {
                    if (rough_pass) {
                        if (sf->model == 1 && sf->finish != 0) {
                            dir = or_unified(sf, dir, normal, r, ph, NULL);
                            pol = or_reflect_pol(pol, normal, dir);
                        } else {
                            dir = f.rdir;
                            pol = f.rpol;
                        }
                    } else {
                        dir = or_lambertian(normal, u01(wr[1]), u01(wr[2]));
                        pol = or_project(pol, dir);
                    }
                } else {
                    dir = f.tdir;
                    pol = f.tpol;
                    transmitted = 1;
                }
            }
        } else {
            reflected = 1;
            if (sf->finish == 0 && mesh_face) {
                dir = or_reflect_dir(dir, normal);
                pol = or_reflect_pol(pol, normal, dir);
            } else if (sf->finish == 0) {
                /* specular reflection at the face normal to `axis`: d - 2 (d.n) n
                 * negates d's axis component and -p + 2 (p.n) n negates p's
                 * other two, exactly (sign flips) */
                dir = mk(axis == 0 ? -dir.x : dir.x, axis == 1 ? -dir.y : dir.y, axis == 2 ? -dir.z : dir.z);
                pol = mk(axis == 0 ? pol.x : -pol.x, axis == 1 ? pol.y : -pol.y, axis == 2 ? pol.z : -pol.z);
            } else {
                dir = or_surface_reflection(sf, dir, normal, r, ph, NULL);
                pol = or_reflect_pol(pol, normal, dir);
            }
        }
    } else {
        const real n2 = or_prop(&nmat->rindex, e, NAN);
        if (!(n1 > RF(0.0)) || !finite_r(n1) || !(n2 > RF(0.0)) || !finite_r(n2)) {
            absorbed = 1;
        } else {
            const fresnel_t f = or_fresnel(dir, pol, normal, n1, n2);
            reflected = (double)u01(w[3]) > f.t;
            if (reflected) {
                dir = f.rdir;
                pol = f.rpol;
            } else {
                dir = f.tdir;
                pol = f.tpol;
                transmitted = 1;
            }
        }
    }
    (void)reflected;
    if (detected) {
        *sensor = nvol->sensor != INVALID ? nvol->sensor : vol->sensor;
        if (*sensor == INVALID) {
            return S_INVALID;
        }
        tr->event = 4;
        return S_DETECTED;
    }
    if (absorbed) {
        return S_ABS_SURFACE;
    }
    ph->dir = dir;
    ph->pol = pol;
    if (transmitted) {
        ph->volume = next;
    }
    return S_ALIVE;
}

/* ------------------------------------------------------------------ */
/* Batch driver                                                        */
/* ------------------------------------------------------------------ */
typedef struct {
    uint64_t detected, absorbed, escaped, truncated, max_bounce, invalid, zero_step, steps, generated;
} ostats_t;

/* bin floor((t - 0) * 256 / 200), as the device computes it */
static int hist_slot(double t) {
    const double x = (t - 0.0) * (256.0 / 200.0);
    if (!(x >= 0.0)) {
        return 256;
    }
    if (x >= 256.0) {
        return 257;
    }
    return (int)x;
}

/* Transports one photon to termination; returns the final status (or -1 for
 * truncation) and fills the hit on detection. */
static int transport_one(photon_t* ph, const oscene_t* s, const rng_t* r, real eps, uint32_t max_steps,
                         ostats_t* st, g4go_hit* hit) {
    st->generated++;
    if (!start_photon(ph)) {
        st->invalid++;
        return S_INVALID;
    }
    for (;;) {
        int zero = 0;
        uint32_t sensor = INVALID;
        trace_t tr;
        int status = or_step(ph, s, r, eps, &zero, &sensor, &tr);
        ph->step++;
        st->steps++;
        st->zero_step += (uint64_t)zero;
        if (status == S_ALIVE && ph->step >= max_steps) {
            status = -1;
        }
        if (status == S_ALIVE) {
            continue;
        }
        if (ph->step > st->max_bounce) {
            st->max_bounce = ph->step;
        }
        if (status == S_DETECTED) {
            st->detected++;
            hit->position_mm[0] = (float)ph->pos.x;
            hit->position_mm[1] = (float)ph->pos.y;
            hit->position_mm[2] = (float)ph->pos.z;
            hit->direction[0] = (float)ph->dir.x;
            hit->direction[1] = (float)ph->dir.y;
            hit->direction[2] = (float)ph->dir.z;
            hit->energy_ev = (float)ph->energy;
            hit->sensor_id = sensor;
            hit->event_id = ph->event;
            hit->photon_id = ph->photon;
            hit->time_ns = ph->time;
        } else if (status == S_ABS_BULK || status == S_ABS_SURFACE) {
            st->absorbed++;
        } else if (status == S_ESCAPED) {
            st->escaped++;
        } else if (status == S_INVALID) {
            st->invalid++;
        } else {
            st->truncated++;
        }
        return status;
    }
}

/* Synthetic batch.  events: n_events (event_id, photon_count) pairs.
 * Hits are written in (event, photon) order; stats per event (9 u64 each:
 * detected, absorbed, escaped, truncated, max_bounce, invalid, zero_step,
 * steps, generated); hist: 258 u64 slots.  Returns the hit count, or -1 when
 * hit_capacity is too small. */
int64_t ORACLE_NAME(run_synthetic)(const g4go_scene* scene, const g4go_synthetic_source* src, uint64_t seed,
                                   float eps, uint32_t max_steps, const uint32_t* event_ids,
                                   const uint32_t* photon_counts, uint32_t n_events, uint32_t photon_begin,
                                   uint32_t photon_end, g4go_hit* hits, uint64_t hit_capacity, uint64_t* stats,
                                   uint64_t* hist) {
    oscene_t* s = build(scene);
    const uint32_t volume = vol_index(s, src->volume_id);
    const rng_t r = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    uint64_t n_hits = 0;
    for (uint32_t e = 0; e < n_events; ++e) {
        ostats_t st;
        memset(&st, 0, sizeof(st));
        const uint32_t end = photon_counts[e] < photon_end ? photon_counts[e] : photon_end;
        for (uint32_t k = photon_begin; k < end; ++k) {
            photon_t ph;
            gen_synthetic(&ph, src, volume, &r, event_ids[e], k);
            g4go_hit hit;
            const int status = transport_one(&ph, s, &r, (real)eps, max_steps, &st, &hit);
            if (status == S_DETECTED) {
                if (n_hits >= hit_capacity) {
                    destroy(s);
                    return -1;
                }
                hits[n_hits++] = hit;
                hist[hist_slot(hit.time_ns)]++;
            }
        }
        uint64_t* o = stats + 9 * (uint64_t)e;
        o[0] += st.detected;
        o[1] += st.absorbed;
        o[2] += st.escaped;
        o[3] += st.truncated;
        o[4] = st.max_bounce > o[4] ? st.max_bounce : o[4];
        o[5] += st.invalid;
        o[6] += st.zero_step;
        o[7] += st.steps;
        o[8] += st.generated;
    }
    destroy(s);
    return (int64_t)n_hits;
}

/* Emission batch (reference semantics of EnqueueEmissions + __raygen__rg):
 * emission_event_indices map each record to an event slot of `stats`. */
int64_t ORACLE_NAME(run_emissions)(const g4go_scene* scene, const g4go_emission* emissions, uint64_t n_emissions,
                                   const uint32_t* emission_event_indices, uint32_t n_events, uint64_t seed,
                                   float eps, uint32_t max_steps, g4go_hit* hits, uint64_t hit_capacity,
                                   uint64_t* stats, uint64_t* hist) {
    oscene_t* s = build(scene);
    const rng_t r = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    uint64_t n_hits = 0;
    for (uint64_t i = 0; i < n_emissions; ++i) {
        const g4go_emission* em = &emissions[i];
        const uint32_t ev = emission_event_indices[i];
        if (ev >= n_events) {
            destroy(s);
            return -2;
        }
        const uint32_t volume = vol_index(s, em->volume_id);
        ostats_t st;
        memset(&st, 0, sizeof(st));
        for (uint32_t k = 0; k < em->photon_count; ++k) {
            photon_t ph;
            gen_emission(&ph, em, volume, s, &r, k);
            g4go_hit hit;
            const int status = transport_one(&ph, s, &r, (real)eps, max_steps, &st, &hit);
            if (status == S_DETECTED) {
                if (n_hits >= hit_capacity) {
                    destroy(s);
                    return -1;
                }
                hits[n_hits++] = hit;
                hist[hist_slot(hit.time_ns)]++;
            }
        }
        uint64_t* o = stats + 9 * (uint64_t)ev;
        o[0] += st.detected;
        o[1] += st.absorbed;
        o[2] += st.escaped;
        o[3] += st.truncated;
        o[4] = st.max_bounce > o[4] ? st.max_bounce : o[4];
        o[5] += st.invalid;
        o[6] += st.zero_step;
        o[7] += st.steps;
        o[8] += st.generated;
    }
    destroy(s);
    return (int64_t)n_hits;
}

/* ------------------------------------------------------------------ */
/* Exports for the closed-form tests                                   */
/* ------------------------------------------------------------------ */
void ORACLE_NAME(rayleigh)(const real dir[3], const real pol[3], uint64_t seed, uint32_t photon, uint32_t event,
                           uint32_t step, real out_dir[3], real out_pol[3]) {
    photon_t ph;
    memset(&ph, 0, sizeof(ph));
    ph.dir = mk(dir[0], dir[1], dir[2]);
    ph.pol = mk(pol[0], pol[1], pol[2]);
    ph.photon = photon;
    ph.event = event;
    ph.step = step;
    const rng_t r = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    or_rayleigh(&ph, &r);
    out_dir[0] = ph.dir.x;
    out_dir[1] = ph.dir.y;
    out_dir[2] = ph.dir.z;
    out_pol[0] = ph.pol.x;
    out_pol[1] = ph.pol.y;
    out_pol[2] = ph.pol.z;
}

void ORACLE_NAME(box_exit)(const float lo[3], const float hi[3], const real p[3], const real d[3], real* dist,
                           int* axis) {
    volume_t v;
    memset(&v, 0, sizeof(v));
    for (int k = 0; k < 3; ++k) {
        v.lo[k] = lo[k];
        v.hi[k] = hi[k];
    }
    box_exit(&v, mk(p[0], p[1], p[2]), mk(d[0], d[1], d[2]), dist, axis);
}

/* Mesh solid queries on the volume `volume_id` of a scene, for n rays
 * (p, d: [n][3]): entry = 0 the exit distance from inside, 1 the entry
 * distance from outside (tmin as in the transport step); face[i] = the face
 * (index among the volume's planes, INVALID for none). */
void ORACLE_NAME(mesh_trace)(const g4go_scene* scene, uint32_t volume_id, const real* p, const real* d, uint64_t n,
                             real tmin, int entry, real* dist, uint32_t* face) {
    oscene_t* s = build(scene);
    const uint32_t vi = vol_index(s, volume_id);
    const volume_t* v = &s->vol[vi];
    for (uint64_t i = 0; i < n; ++i) {
        const v3 pp = mk(p[3 * i], p[3 * i + 1], p[3 * i + 2]);
        const v3 dd = mk(d[3 * i], d[3 * i + 1], d[3 * i + 2]);
        uint32_t pl = INVALID;
        int l;
        if (v->tri) {
            dist[i] = tri_closest(s, v, pp, dd, entry ? tmin : RF(0.0), entry ? 1 : 0, &pl, &l);
            pl = pl == INVALID ? INVALID : (pl & ~FACE_TRI);
        } else if (entry) {
            dist[i] = mesh_entry(s, v, pp, dd, tmin, &pl);
        } else {
            mesh_exit(s, v, pp, dd, &dist[i], &pl);
        }
        face[i] = pl == INVALID ? INVALID : pl - v->plane_begin;
    }
    destroy(s);
}

/* Solid kind of volume `volume_id`: 0 box, 1 convex mesh (planes), 2
 * triangles, 3 open triangle solid; and whether point q lies inside it. */
int ORACLE_NAME(solid_kind)(const g4go_scene* scene, uint32_t volume_id) {
    oscene_t* s = build(scene);
    const volume_t* v = &s->vol[vol_index(s, volume_id)];
    const int kind = v->tri ? (v->open ? 3 : 2) : (v->plane_count ? 1 : 0);
    destroy(s);
    return kind;
}

void ORACLE_NAME(solid_contains)(const g4go_scene* scene, uint32_t volume_id, const real* q, uint64_t n, int* out) {
    oscene_t* s = build(scene);
    const volume_t* v = &s->vol[vol_index(s, volume_id)];
    for (uint64_t i = 0; i < n; ++i) {
        const v3 p = mk(q[3 * i], q[3 * i + 1], q[3 * i + 2]);
        out[i] = v->plane_count ? solid_contains(s, v, p) : box_contains(v, p);
    }
    destroy(s);
}

/* The face planes of volume `volume_id` (n[3], h, s[3] each) into out
 * (room for max planes); returns the plane count. */
uint32_t ORACLE_NAME(mesh_planes)(const g4go_scene* scene, uint32_t volume_id, real* out, uint32_t max) {
    oscene_t* s = build(scene);
    const volume_t* v = &s->vol[vol_index(s, volume_id)];
    const uint32_t n = v->plane_count;
    for (uint32_t k = 0; k < n && k < max; ++k) {
        memcpy(out + 7 * k, v->tri ? &s->tris[v->plane_begin + k].face : &s->planes[v->plane_begin + k],
               7 * sizeof(real));
    }
    destroy(s);
    return n;
}

void ORACLE_NAME(project_polarization)(const real pol[3], const real dir[3], real out[3]) {
    const v3 r = or_project(mk(pol[0], pol[1], pol[2]), mk(dir[0], dir[1], dir[2]));
    out[0] = r.x;
    out[1] = r.y;
    out[2] = r.z;
}

void ORACLE_NAME(reflect_direction)(const real dir[3], const real n[3], real out[3]) {
    const v3 r = or_reflect_dir(mk(dir[0], dir[1], dir[2]), mk(n[0], n[1], n[2]));
    out[0] = r.x;
    out[1] = r.y;
    out[2] = r.z;
}

void ORACLE_NAME(reflect_polarization)(const real pol[3], const real n[3], const real dir[3], real out[3]) {
    const v3 r = or_reflect_pol(mk(pol[0], pol[1], pol[2]), mk(n[0], n[1], n[2]), mk(dir[0], dir[1], dir[2]));
    out[0] = r.x;
    out[1] = r.y;
    out[2] = r.z;
}

void ORACLE_NAME(normalize)(const real a[3], real out[3]) {
    const v3 r = or_normalize(mk(a[0], a[1], a[2]));
    out[0] = r.x;
    out[1] = r.y;
    out[2] = r.z;
}

/* ------------------------------------------------------------------ */
/* Path-replay likelihood-ratio gradient (the adjoint of the backend)  */
/* ------------------------------------------------------------------ */
static void or_prop_node(const prop_t* p, real e, uint32_t* i, real* w) {
    if (p->n <= 1 || e <= p->e[0]) {
        *i = 0;
        *w = RF(0.0);
        return;
    }
    const uint32_t last = p->n - 1;
    if (e >= p->e[last]) {
        *i = last;
        *w = RF(0.0);
        return;
    }
    uint32_t k = 0;
    while (k + 1 < p->n && e > p->e[k + 1]) {
        ++k;
    }
    *i = k;
    *w = (e - p->e[k]) / (p->e[k + 1] - p->e[k]);
}

static void or_add_table(double* grad, const prop_t* p, uint32_t offset, real e, double g) {
    uint32_t i;
    real w;
    or_prop_node(p, e, &i, &w);
    grad[offset + i] += g * (1.0 - (double)w);
    if (w != RF(0.0)) {
        grad[offset + i + 1] += g * (double)w;
    }
}

/* A replayed path's sums, per material (bulk length, Rayleigh scatters) and
 * per parameterised surface (reflections), kept sparse: the arrays span the
 * scene and the lists name the entries this path touched, in first-touch
 * order, so that they are cleared in time proportional to the path. */
typedef struct {
    double* length;
    uint32_t* rayleigh;
    uint32_t* reflections;
    uint32_t *mats, *surfs;
    uint32_t n_mats, n_surfs;
} path_sums_t;

static void path_add_segment(path_sums_t* p, uint32_t m, real len, int rayleigh) {
    if (p->length[m] == 0.0 && p->rayleigh[m] == 0u) {
        p->mats[p->n_mats++] = m;
    }
    p->length[m] += (double)len;
    p->rayleigh[m] += rayleigh ? 1u : 0u;
}

static void path_add_reflection(path_sums_t* p, uint32_t sf) {
    if (p->reflections[sf]++ == 0u) {
        p->surfs[p->n_surfs++] = sf;
    }
}

static void path_clear(path_sums_t* p) {
    for (uint32_t k = 0; k < p->n_mats; ++k) {
        p->length[p->mats[k]] = 0.0;
        p->rayleigh[p->mats[k]] = 0u;
    }
    for (uint32_t k = 0; k < p->n_surfs; ++k) {
        p->reflections[p->surfs[k]] = 0u;
    }
    p->n_mats = p->n_surfs = 0;
}

/* Whether a replayed path ends at its recorded detection.  The FP32 replay
 * repeats the forward arithmetic and must match bit for bit; the FP64 replay of
 * an FP32 run follows the same decisions with other roundings, so it matches
 * within 1e-6 relative in time and 1e-3 mm in position. */
static int same_end(const photon_t* ph, int status, const g4go_hit* hit) {
    if (status != S_DETECTED) return 0;
#ifdef ORACLE_DOUBLE
    return fabs(ph->time - hit->time_ns) <= 1e-6 * (1.0 + fabs(hit->time_ns)) &&
           fabs(ph->pos.x - (double)hit->position_mm[0]) <= 1e-3 && fabs(ph->pos.y - (double)hit->position_mm[1]) <= 1e-3 &&
           fabs(ph->pos.z - (double)hit->position_mm[2]) <= 1e-3;
#else
    return ph->time == hit->time_ns && (float)ph->pos.x == hit->position_mm[0] && (float)ph->pos.y == hit->position_mm[1] &&
           (float)ph->pos.z == hit->position_mm[2];
#endif
}

/* Regenerates the photon of a recorded hit; returns 0 when the hit names no
 * photon of the batch. */
typedef int (*regen_fn)(photon_t* ph, const void* ctx, const oscene_t* s, const rng_t* r, const g4go_hit* hit);

typedef struct {
    const g4go_synthetic_source* src;
    uint32_t volume;
} synth_ctx_t;

static int regen_synthetic(photon_t* ph, const void* c, const oscene_t* s, const rng_t* r, const g4go_hit* hit) {
    const synth_ctx_t* x = (const synth_ctx_t*)c;
    (void)s;
    gen_synthetic(ph, x->src, x->volume, r, hit->event_id, hit->photon_id);
    return 1;
}

typedef struct {
    uint32_t event_id, first, index, pad;
} em_key_t;

typedef struct {
    const g4go_emission* em;
    const em_key_t* keys;  /* records sorted by (event_id, first_photon_id) */
    uint64_t n;
} em_ctx_t;

static int em_key_cmp(const void* a, const void* b) {
    const em_key_t* x = (const em_key_t*)a;
    const em_key_t* y = (const em_key_t*)b;
    if (x->event_id != y->event_id) return x->event_id < y->event_id ? -1 : 1;
    if (x->first != y->first) return x->first < y->first ? -1 : 1;
    return x->index < y->index ? -1 : (x->index > y->index);
}

/* The record of event hit->event_id whose photon range holds hit->photon_id
 * (photon IDs are assigned per record from first_photon_id,
 * OptiXTransportDevice.cu:359). */
static int regen_emission(photon_t* ph, const void* c, const oscene_t* s, const rng_t* r, const g4go_hit* hit) {
    const em_ctx_t* x = (const em_ctx_t*)c;
    uint64_t lo = 0, hi = x->n;
    while (lo < hi) { /* first key > (event, photon) */
        const uint64_t mid = (lo + hi) / 2;
        const em_key_t* k = &x->keys[mid];
        if (k->event_id < hit->event_id || (k->event_id == hit->event_id && k->first <= hit->photon_id)) {
            lo = mid + 1;
        } else {
            hi = mid;
        }
    }
    if (lo == 0) return 0;
    const em_key_t* k = &x->keys[lo - 1];
    const g4go_emission* em = &x->em[k->index];
    if (k->event_id != hit->event_id || hit->photon_id - k->first >= em->photon_count) return 0;
    gen_emission(ph, em, vol_index(s, em->volume_id), s, r, hit->photon_id - k->first);
    return 1;
}

/* tables: n_tables (kind, index) pairs bound to theta in order (kind 0 absorption,
 * 1 Rayleigh of a material; 2 reflectivity, 3 efficiency of a surface).
 * counters: replayed, mismatched, replayed paths beyond two materials or two
 * parameterised surfaces, steps.  Returns 0, or -1 when the
 * layout does not fit n_params. */
static int adjoint_core(oscene_t* s, regen_fn regen, const void* ctx, uint64_t seed, float eps, uint32_t max_steps,
                        const g4go_hit* hits, uint64_t n_hits, const double* slot_weight, const uint32_t* tables,
                        uint32_t n_tables, double* grad, uint64_t n_params, uint64_t* counters) {
    uint32_t* mabs = (uint32_t*)malloc(sizeof(uint32_t) * (s->n_mat + 1));
    uint32_t* mray = (uint32_t*)malloc(sizeof(uint32_t) * (s->n_mat + 1));
    uint32_t* srefl = (uint32_t*)malloc(sizeof(uint32_t) * (s->n_surf + 1));
    uint32_t* seff = (uint32_t*)malloc(sizeof(uint32_t) * (s->n_surf + 1));
    for (uint32_t i = 0; i <= s->n_mat; ++i) mabs[i] = mray[i] = INVALID;
    for (uint32_t i = 0; i <= s->n_surf; ++i) srefl[i] = seff[i] = INVALID;
    uint32_t offset = 0;
    for (uint32_t t = 0; t < n_tables; ++t) {
        const uint32_t kind = tables[2 * t], idx = tables[2 * t + 1];
        const prop_t* p = NULL;
        if (kind == 0 && idx < s->n_mat) { mabs[idx] = offset; p = &s->mat[idx].abslen; }
        if (kind == 1 && idx < s->n_mat) { mray[idx] = offset; p = &s->mat[idx].rayleigh; }
        if (kind == 2 && idx < s->n_surf) { srefl[idx] = offset; p = &s->surf[idx].refl; }
        if (kind == 3 && idx < s->n_surf) { seff[idx] = offset; p = &s->surf[idx].eff; }
        if (!p) { offset = 0xffffffffu; break; }
        offset += p->n;
    }
    if (offset != n_params) {
        free(mabs); free(mray); free(srefl); free(seff);
        return -1;
    }
    const rng_t r = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    path_sums_t ps;
    ps.length = (double*)calloc(s->n_mat + 1, sizeof(double));
    ps.rayleigh = (uint32_t*)calloc(s->n_mat + 1, sizeof(uint32_t));
    ps.reflections = (uint32_t*)calloc(s->n_surf + 1, sizeof(uint32_t));
    ps.mats = (uint32_t*)malloc(sizeof(uint32_t) * (s->n_mat + 1));
    ps.surfs = (uint32_t*)malloc(sizeof(uint32_t) * (s->n_surf + 1));
    ps.n_mats = ps.n_surfs = 0;
    for (uint64_t h = 0; h < n_hits; ++h) {
        const g4go_hit* hit = &hits[h];
        photon_t ph;
        if (!regen(&ph, ctx, s, &r, hit) || !start_photon(&ph)) { counters[1]++; continue; }
        path_clear(&ps);
        int status = S_ALIVE, zero = 0;
        uint32_t sensor = INVALID, det_surface = INVALID;
        for (;;) {
            trace_t tr;
            status = or_step(&ph, s, &r, (real)eps, &zero, &sensor, &tr);
            ph.step++;
            counters[3]++;
            if ((tr.segment > RF(0.0) || tr.event == 1) && tr.material < s->n_mat) {
                path_add_segment(&ps, tr.material, tr.segment, tr.event == 1);
            }
            if (tr.event == 2 && tr.surface < s->n_surf && srefl[tr.surface] != INVALID) {
                path_add_reflection(&ps, tr.surface);
            }
            if (status == S_DETECTED) det_surface = tr.surface;
            if (status != S_ALIVE || ph.step >= max_steps) break;
        }
        if (!same_end(&ph, status, hit)) {
            counters[1]++;
            continue;
        }
        counters[0]++;
        /* paths beyond two materials or two parameterised surfaces (the device
         * scores those on a second replay) */
        if (ps.n_mats > 2 || ps.n_surfs > 2) counters[2]++;
        const real e = ph.energy;
        const double wgt = slot_weight[hist_slot(hit->time_ns)];
        for (uint32_t k = 0; k < ps.n_mats; ++k) {
            const uint32_t mi = ps.mats[k];
            const material_t* mt = &s->mat[mi];
            if (mabs[mi] != INVALID && mt->abslen.n) {
                const real la = or_prop(&mt->abslen, e, R_INF);
                if (finite_r(la) && la > RF(0.0)) {
                    const double l = (double)la;
                    or_add_table(grad, &mt->abslen, mabs[mi], e, wgt * ps.length[mi] / (l * l));
                }
            }
            if (mray[mi] != INVALID && mt->rayleigh.n) {
                const real lr = or_prop(&mt->rayleigh, e, R_INF);
                if (finite_r(lr) && lr > RF(0.0)) {
                    const double l = (double)lr;
                    or_add_table(grad, &mt->rayleigh, mray[mi], e,
                                 wgt * (ps.length[mi] / (l * l) - (double)ps.rayleigh[mi] / l));
                }
            }
        }
        for (uint32_t k = 0; k < ps.n_surfs; ++k) {
            const uint32_t si = ps.surfs[k];
            const surface_t* sf = &s->surf[si];
            const real rr = or_prop(&sf->refl, e, RF(0.0));
            if (rr > RF(0.0) && rr < RF(1.0)) {
                or_add_table(grad, &sf->refl, srefl[si], e, wgt * (double)ps.reflections[si] / (double)rr);
            }
        }
        if (det_surface < s->n_surf) {
            const surface_t* sf = &s->surf[det_surface];
            if (seff[det_surface] != INVALID && sf->eff.n) {
                const real ef = or_prop(&sf->eff, e, RF(0.0));
                if (ef > RF(0.0) && ef < RF(1.0)) {
                    or_add_table(grad, &sf->eff, seff[det_surface], e, wgt / (double)ef);
                }
            }
            if (srefl[det_surface] != INVALID && sf->refl.n) {
                const real rr = or_prop(&sf->refl, e, RF(0.0));
                const real tt = sf->trans.n ? clamp01(or_prop(&sf->trans, e, RF(0.0))) : RF(0.0);
                const double ab = 1.0 - (double)rr - (double)tt;
                if (rr > RF(0.0) && rr < RF(1.0) && ab > 0.0) {
                    or_add_table(grad, &sf->refl, srefl[det_surface], e, -wgt / ab);
                }
            }
        }
    }
    free(ps.length); free(ps.rayleigh); free(ps.reflections); free(ps.mats); free(ps.surfs);
    free(mabs); free(mray); free(srefl); free(seff);
    return 0;
}

int ORACLE_NAME(adjoint_synthetic)(const g4go_scene* scene, const g4go_synthetic_source* src, uint64_t seed, float eps,
                                   uint32_t max_steps, const g4go_hit* hits, uint64_t n_hits,
                                   const double* slot_weight, const uint32_t* tables, uint32_t n_tables, double* grad,
                                   uint64_t n_params, uint64_t* counters) {
    oscene_t* s = build(scene);
    const synth_ctx_t ctx = {src, vol_index(s, src->volume_id)};
    const int rc = adjoint_core(s, regen_synthetic, &ctx, seed, eps, max_steps, hits, n_hits, slot_weight, tables,
                                n_tables, grad, n_params, counters);
    destroy(s);
    return rc;
}

/* The replay gradient of an emission-record batch: each hit's photon is
 * regenerated from the record that owns its photon ID. */
int ORACLE_NAME(adjoint_emissions)(const g4go_scene* scene, const g4go_emission* emissions, uint64_t n_emissions,
                                   uint64_t seed, float eps, uint32_t max_steps, const g4go_hit* hits, uint64_t n_hits,
                                   const double* slot_weight, const uint32_t* tables, uint32_t n_tables, double* grad,
                                   uint64_t n_params, uint64_t* counters) {
    oscene_t* s = build(scene);
    em_key_t* keys = (em_key_t*)malloc(sizeof(em_key_t) * (n_emissions + 1));
    for (uint64_t i = 0; i < n_emissions; ++i) {
        keys[i].event_id = emissions[i].event_id;
        keys[i].first = emissions[i].first_photon_id;
        keys[i].index = (uint32_t)i;
        keys[i].pad = 0;
    }
    qsort(keys, n_emissions, sizeof(em_key_t), em_key_cmp);
    const em_ctx_t ctx = {emissions, keys, n_emissions};
    const int rc = adjoint_core(s, regen_emission, &ctx, seed, eps, max_steps, hits, n_hits, slot_weight, tables,
                                n_tables, grad, n_params, counters);
    free(keys);
    destroy(s);
    return rc;
}

/* ------------------------------------------------------------------ */
/* Exports for the distribution tests of generation and surfaces      */
/* ------------------------------------------------------------------ */
/* Photon state as 12 doubles: pos[3], dir[3], pol[3], energy, time, valid
 * (valid = 0 when start_photon rejects the photon, OptiXTransportDevice.cu:693-702). */
static void put_state(photon_t* ph, double* o) {
    const int ok = start_photon(ph);
    o[0] = ph->pos.x; o[1] = ph->pos.y; o[2] = ph->pos.z;
    o[3] = ph->dir.x; o[4] = ph->dir.y; o[5] = ph->dir.z;
    o[6] = ph->pol.x; o[7] = ph->pol.y; o[8] = ph->pol.z;
    o[9] = ph->energy;
    o[10] = ph->time;
    o[11] = ok;
}

/* Photons first .. first + n - 1 of one emission record (GeneratePhoton,
 * OptiXTransportDevice.cu:346-481), as generated before transport. */
void ORACLE_NAME(generate_emission)(const g4go_scene* scene, const g4go_emission* em, uint64_t seed, uint32_t first,
                                    uint32_t n, double* out) {
    oscene_t* s = build(scene);
    const rng_t r = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    const uint32_t volume = vol_index(s, em->volume_id);
    for (uint32_t k = 0; k < n; ++k) {
        photon_t ph;
        gen_emission(&ph, em, volume, s, &r, first + k);
        put_state(&ph, out + 12 * (uint64_t)k);
    }
    destroy(s);
}

/* Photons first .. first + n - 1 of one synthetic event. */
void ORACLE_NAME(generate_synthetic)(const g4go_scene* scene, const g4go_synthetic_source* src, uint64_t seed,
                                     uint32_t event, uint32_t first, uint32_t n, double* out) {
    oscene_t* s = build(scene);
    const rng_t r = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    const uint32_t volume = vol_index(s, src->volume_id);
    for (uint32_t k = 0; k < n; ++k) {
        photon_t ph;
        gen_synthetic(&ph, src, volume, &r, event, first + k);
        put_state(&ph, out + 12 * (uint64_t)k);
    }
    destroy(s);
}

/* Reflection off surface `surface` (its index in the scene's surface list)
 * for photons photon_first .. + n - 1 of `event` at step `step`, arriving
 * along d at a face whose normal nrm faces the photon
 * (SampleSurfaceReflection, OptiXTransportDevice.cu:582-595): out_dir the
 * reflected direction, out_facet the facet normal SampleFacetNormal draws
 * (:492-556), branch the unified lobe (0 spike, 1 lobe, 2 backscatter, 3
 * Lambertian; -1 when no lobe is chosen). */
void ORACLE_NAME(surface_reflect)(const g4go_scene* scene, uint32_t surface, const real d[3], const real nrm[3],
                                  real energy, uint64_t seed, uint32_t event, uint32_t step, uint32_t photon_first,
                                  uint32_t n, real* out_dir, real* out_facet, int* branch) {
    oscene_t* s = build(scene);
    const rng_t r = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    const surface_t* sf = &s->surf[surface];
    const v3 dd = mk(d[0], d[1], d[2]);
    const v3 nn = mk(nrm[0], nrm[1], nrm[2]);
    for (uint32_t k = 0; k < n; ++k) {
        photon_t ph;
        memset(&ph, 0, sizeof(ph));
        ph.photon = photon_first + k;
        ph.event = event;
        ph.step = step;
        ph.energy = energy;
        int b = -1;
        const v3 o = or_surface_reflection(sf, dd, nn, &r, &ph, &b);
        const v3 f = or_facet(sf, nn, dd, &r, &ph);
        out_dir[3 * k] = o.x; out_dir[3 * k + 1] = o.y; out_dir[3 * k + 2] = o.z;
        out_facet[3 * k] = f.x; out_facet[3 * k + 1] = f.y; out_facet[3 * k + 2] = f.z;
        branch[k] = b;
    }
    destroy(s);
}

/* One transport step (the bounce loop body, OptiXTransportDevice.cu:708-931)
 * from a given state, for photons photon_first .. + n - 1 of `event` at step
 * `step`, all starting from state[12] (pos, dir, pol, energy, time, -) in
 * volume `volume_id`.  out: the state after the step (12 doubles; [11] = the
 * volume index after the step), status: S_* (0 alive, 1 detected, 2 bulk
 * absorbed, 3 surface absorbed, 4 escaped, 5 invalid). */
void ORACLE_NAME(step_from)(const g4go_scene* scene, uint32_t volume_id, const double* state, uint64_t seed,
                            uint32_t event, uint32_t step, uint32_t photon_first, uint32_t n, float eps, double* out,
                            int* status) {
    oscene_t* s = build(scene);
    const rng_t r = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    const uint32_t volume = vol_index(s, volume_id);
    for (uint32_t k = 0; k < n; ++k) {
        photon_t ph;
        memset(&ph, 0, sizeof(ph));
        ph.pos = mk((real)state[0], (real)state[1], (real)state[2]);
        ph.dir = mk((real)state[3], (real)state[4], (real)state[5]);
        ph.pol = mk((real)state[6], (real)state[7], (real)state[8]);
        ph.energy = (real)state[9];
        ph.time = state[10];
        ph.volume = volume;
        ph.material = INVALID;
        ph.photon = photon_first + k;
        ph.event = event;
        ph.step = step;
        int zero = 0;
        uint32_t sensor = INVALID;
        trace_t tr;
        status[k] = or_step(&ph, s, &r, (real)eps, &zero, &sensor, &tr);
        double* o = out + 12 * (uint64_t)k;
        o[0] = ph.pos.x; o[1] = ph.pos.y; o[2] = ph.pos.z;
        o[3] = ph.dir.x; o[4] = ph.dir.y; o[5] = ph.dir.z;
        o[6] = ph.pol.x; o[7] = ph.pol.y; o[8] = ph.pol.z;
        o[9] = ph.energy;
        o[10] = ph.time;
        o[11] = ph.volume;
    }
    destroy(s);
}
