// Host side of the C ABI: scene flattening, batch validation and staging,
// launches and readback.
//
// Mirrors OptiXTransportHost::Impl (src/optical/optix/OptiXTransportHost.cpp):
// one in-flight batch slot (:523-525), validation before any device work
// (:572-666), pinned staging, stream-ordered upload/launch/readback, and a
// completion call that waits on the batch and assembles the output
// (:779-883).  Host-side checks all run before the first HIP call, so the
// error paths are testable without a GPU.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <limits>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <unordered_set>
#include <memory>
#include <mutex>
#include <vector>

#include "../../include/g4go_hip.h"
#include "kernels.h"

using namespace g4go;

static_assert(sizeof(g4go_emission) == sizeof(Emission), "emission layout");
static_assert(offsetof(g4go_emission, event_id) == offsetof(Emission, event_id), "emission layout");
static_assert(offsetof(g4go_emission, type) == offsetof(Emission, type), "emission layout");
static_assert(sizeof(g4go_hit) == sizeof(Hit), "hit layout");
static_assert(offsetof(g4go_hit, time_ns) == offsetof(Hit, time), "hit layout");

namespace {

struct Failure {
    g4go_status status;
    std::string message;
};

[[noreturn]] void fail(g4go_status status, const std::string& message) { throw Failure{status, message}; }

void hip_check(hipError_t err, const char* what) {
    if (err != hipSuccess) {
        fail(G4GO_ERROR_RUNTIME, std::string(what) + " failed: " + hipGetErrorString(err));
    }
}

thread_local std::string g_create_error;

// ---------------------------------------------------------------------------
// Device scratch of the allocation-free entry point
// (g4go_transport_synthetic_noalloc): a fixed array in every GPU's memory,
// laid out as
//   [0]            batch header {n_sources, chunk, n_events, 0}
//   [256]          counters (kCounterCount u64: steps, photons, dropped, queue head)
//   [512]          item offsets (kPrefixMaxSources + 1 u32)
//   [4864]         photon offsets (kPrefixMaxSources + 1 u32)
//   [9216]         kEventReplicas x kPrefixMaxSources per-event accumulation copies
// Host code never passes `noalloc_scratch` itself to a kernel (in host code the
// name is a host-side stand-in, not a device address): it uses the device
// address noalloc_scratch_address() looked up once per device.
// ---------------------------------------------------------------------------
constexpr size_t scratch_round(size_t x) { return (x + 255) & ~static_cast<size_t>(255); }
constexpr size_t kScratchHeader = 0;
constexpr size_t kScratchCounters = 256;
constexpr size_t kScratchItems = 512;
constexpr size_t kScratchPhotons = scratch_round(kScratchItems + (kPrefixMaxSources + 1) * sizeof(uint32_t));
constexpr size_t kScratchReplicas = scratch_round(kScratchPhotons + (kPrefixMaxSources + 1) * sizeof(uint32_t));
constexpr size_t kScratchBytes =
    kScratchReplicas + static_cast<size_t>(kEventReplicas) * kPrefixMaxSources * sizeof(EventStats);
static_assert(kScratchBytes == G4GO_NOALLOC_SCRATCH_BYTES, "scratch size of the C ABI");
static_assert(kPrefixMaxSources == G4GO_NOALLOC_MAX_EVENTS, "event limit of the C ABI");

static __device__ __align__(256) unsigned char noalloc_scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
void* g_scratch_address[kMaxDevices] = {};

// Device address of noalloc_scratch on the current device `device`, looked up
// once per device (each GPU holds its own copy of the array).
void* noalloc_scratch_address(int device) {
    if (device < 0 || device >= kMaxDevices) {
        fail(G4GO_ERROR_RUNTIME, "device index beyond the scratch table");
    }
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (g_scratch_address[device] == nullptr) {
        void* ptr = nullptr;
        hip_check(hipGetSymbolAddress(&ptr, HIP_SYMBOL(noalloc_scratch)), "hipGetSymbolAddress scratch");
        if (ptr == nullptr) {
            fail(G4GO_ERROR_RUNTIME, "hipGetSymbolAddress returned no scratch address");
        }
        g_scratch_address[device] = ptr;
    }
    return g_scratch_address[device];
}

// ---------------------------------------------------------------------------
// Host scene (flattened Scene, see scene.h)
// ---------------------------------------------------------------------------
struct HostScene {
    std::vector<float> pool;
    std::vector<Material> materials;
    std::vector<Surface> surfaces;
    std::vector<Volume> volumes;
    std::vector<uint32_t> daughters;
    std::vector<Binding> bindings;
    std::vector<BvhNode> bvh;
    std::vector<Plane> planes;  // faces of the convex mesh solids
    std::vector<std::vector<double>> face_points;  // per plane: its triangles' vertices (xyz), host only
    std::vector<Tri> tris;                         // triangles of the non-convex mesh solids
    std::vector<std::vector<double>> tri_points;   // per triangle: its vertices (xyz), host only
    std::vector<uint32_t> face_sibs;               // sibling lists of kFaceDynamic faces (Plane::sib_first)
    uint32_t world = kInvalid;
    uint32_t hot_offset = 0, hot_count = 0;  // hot energy grid (mark_hot_grid)
    std::unordered_map<uint32_t, uint32_t> index_of;  // public volume ID -> dense index
    std::vector<uint32_t> volume_ids;                 // dense index -> public volume ID
};

bool finite_all(const float* v, uint32_t n) {
    for (uint32_t i = 0; i < n; ++i) {
        if (!std::isfinite(v[i])) {
            return false;
        }
    }
    return true;
}

Prop add_property(HostScene& s, const g4go_property& p, const std::string& what) {
    Prop out{0, 0, 0, 0};
    if (p.count == 0) {
        return out;
    }
    if (p.energy_ev == nullptr || p.values == nullptr) {
        fail(G4GO_ERROR_INVALID_ARGUMENT, what + ": null property table");
    }
    if (!finite_all(p.energy_ev, p.count) || !finite_all(p.values, p.count)) {
        fail(G4GO_ERROR_INVALID_ARGUMENT, what + ": non-finite property value");
    }
    for (uint32_t i = 1; i < p.count; ++i) {
        if (p.energy_ev[i] < p.energy_ev[i - 1]) {
            fail(G4GO_ERROR_INVALID_ARGUMENT, what + ": property energies must be ascending");
        }
    }
    out.offset = static_cast<uint32_t>(s.pool.size());
    out.count = p.count;
    bool constant = true;
    for (uint32_t i = 1; i < p.count; ++i) {
        constant = constant && p.values[i] == p.values[0];
    }
    out.flags = constant ? kPropConstant : 0u;  // PropertyTable::Constant (Scene.cpp:17-21)
    out.value0 = p.values[0];
    s.pool.insert(s.pool.end(), p.energy_ev, p.energy_ev + p.count);
    s.pool.insert(s.pool.end(), p.values, p.values + p.count);
    return out;
}

// Normalised trapezoid CDF of a scintillation spectrum
// (UploadSpectrumProperty, OptiXTransportHost.cpp:371-410).
Prop add_spectrum(HostScene& s, const g4go_property& p, const std::string& what) {
    Prop out{0, 0, 0, 0};
    if (p.count == 0 || p.energy_ev == nullptr || p.values == nullptr) {
        return out;
    }
    std::vector<float> cdf(p.count, 0.0f);
    if (p.count == 1) {
        cdf[0] = 1.0f;
    } else {
        for (uint32_t i = 1; i < p.count; ++i) {
            const float de = p.energy_ev[i] - p.energy_ev[i - 1];
            const float v0 = std::max(p.values[i - 1], 0.0f);
            const float v1 = std::max(p.values[i], 0.0f);
            if (!(de >= 0.0f) || !std::isfinite(de) || !std::isfinite(v0) || !std::isfinite(v1)) {
                return out;
            }
            cdf[i] = cdf[i - 1] + 0.5f * de * (v0 + v1);
        }
        const float integral = cdf.back();
        if (!(integral > 0.0f) || !std::isfinite(integral)) {
            return out;
        }
        for (float& v : cdf) {
            v /= integral;
        }
        cdf.back() = 1.0f;
    }
    (void)what;
    out.offset = static_cast<uint32_t>(s.pool.size());
    out.count = p.count;
    out.flags = 0;
    out.value0 = cdf[0];
    s.pool.insert(s.pool.end(), p.energy_ev, p.energy_ev + p.count);
    s.pool.insert(s.pool.end(), cdf.begin(), cdf.end());
    return out;
}

// Scene::FindBoundarySurface semantics (Scene.cpp:120-150) on dense indices.
uint32_t host_find_surface(const HostScene& s, uint32_t from, uint32_t to) {
    for (const Binding& b : s.bindings) {
        if (b.from == from && b.to == to) {
            return b.surface;
        }
    }
    if (from >= s.volumes.size() || to >= s.volumes.size()) {
        return kInvalid;
    }
    const Volume& f = s.volumes[from];
    const Volume& t = s.volumes[to];
    if (t.parent == from) {
        return t.skin != kInvalid ? t.skin : f.skin;
    }
    return f.skin != kInvalid ? f.skin : t.skin;
}

void build_bvh(HostScene& s, Volume& vol, uint32_t begin, uint32_t count) {
    // Median split on the longest axis of the centroid bounds; leaves of <= 4.
    struct Task {
        uint32_t node, first, count, depth;
    };
    auto bounds = [&](uint32_t first, uint32_t n, float* lo, float* hi) {
        for (int k = 0; k < 3; ++k) {
            lo[k] = std::numeric_limits<float>::infinity();
            hi[k] = -std::numeric_limits<float>::infinity();
        }
        for (uint32_t i = first; i < first + n; ++i) {
            const Volume& d = s.volumes[s.daughters[i]];
            for (int k = 0; k < 3; ++k) {
                lo[k] = std::min(lo[k], d.lo[k]);
                hi[k] = std::max(hi[k], d.hi[k]);
            }
        }
    };
    vol.bvh_root = static_cast<uint32_t>(s.bvh.size());
    s.bvh.push_back(BvhNode{});
    std::vector<Task> stack{{vol.bvh_root, begin, count, 0u}};
    while (!stack.empty()) {
        const Task t = stack.back();
        stack.pop_back();
        if (t.depth > static_cast<uint32_t>(kBvhMaxDepth)) {
            fail(G4GO_ERROR_INVALID_ARGUMENT, "daughter BVH deeper than the device traversal stack");
        }
        BvhNode node{};
        bounds(t.first, t.count, node.lo, node.hi);
        if (t.count <= 4) {
            node.left = t.first;
            node.right_or_count = t.count | kBvhLeaf;
            s.bvh[t.node] = node;
            continue;
        }
        float clo[3], chi[3];
        for (int k = 0; k < 3; ++k) {
            clo[k] = std::numeric_limits<float>::infinity();
            chi[k] = -std::numeric_limits<float>::infinity();
        }
        for (uint32_t i = t.first; i < t.first + t.count; ++i) {
            const Volume& d = s.volumes[s.daughters[i]];
            for (int k = 0; k < 3; ++k) {
                const float c = 0.5f * (d.lo[k] + d.hi[k]);
                clo[k] = std::min(clo[k], c);
                chi[k] = std::max(chi[k], c);
            }
        }
        int axis = 0;
        for (int k = 1; k < 3; ++k) {
            if (chi[k] - clo[k] > chi[axis] - clo[axis]) {
                axis = k;
            }
        }
        auto first = s.daughters.begin() + t.first;
        auto mid = first + t.count / 2;
        std::nth_element(first, mid, first + t.count, [&](uint32_t a, uint32_t b) {
            const float ca = s.volumes[a].lo[axis] + s.volumes[a].hi[axis];
            const float cb = s.volumes[b].lo[axis] + s.volumes[b].hi[axis];
            return ca < cb || (ca == cb && a < b);
        });
        const uint32_t half = t.count / 2;
        node.left = static_cast<uint32_t>(s.bvh.size());
        node.right_or_count = node.left + 1;
        s.bvh.push_back(BvhNode{});
        s.bvh.push_back(BvhNode{});
        s.bvh[t.node] = node;
        stack.push_back({node.left, t.first, half, t.depth + 1});
        stack.push_back({node.right_or_count, t.first + half, t.count - half, t.depth + 1});
    }
}

// The energy grid shared by the most non-constant tables of materials and
// surfaces (ties: the first seen) becomes the hot grid; every table on it is
// flagged kPropHotGrid.  Spectrum CDFs are sampled, not interpolated, and stay
// out.  Tables on one grid have bit-identical energy arrays.
void mark_hot_grid(HostScene& s) {
    std::vector<Prop*> props;
    for (Material& m : s.materials) {
        for (Prop* p : {&m.rindex, &m.group_velocity, &m.abs_length, &m.rayleigh_length}) {
            props.push_back(p);
        }
    }
    for (Surface& f : s.surfaces) {
        for (Prop* p : {&f.reflectivity, &f.efficiency, &f.transmittance, &f.lobe, &f.spike, &f.backscatter,
                        &f.roughness}) {
            props.push_back(p);
        }
    }
    auto same_grid = [&](const Prop& a, const Prop& b) {
        return a.count == b.count &&
               std::memcmp(s.pool.data() + a.offset, s.pool.data() + b.offset, a.count * sizeof(float)) == 0;
    };
    const Prop* best = nullptr;
    size_t best_uses = 0;
    for (const Prop* a : props) {
        if (a->count < 2) {
            continue;
        }
        size_t uses = 0;
        for (const Prop* b : props) {
            uses += b->count >= 2 && (b->flags & kPropConstant) == 0 && same_grid(*a, *b) ? 1u : 0u;
        }
        if (uses > best_uses) {
            best = a;
            best_uses = uses;
        }
    }
    s.hot_offset = 0;
    s.hot_count = 0;
    if (best == nullptr || best_uses == 0) {
        return;
    }
    const Prop grid = *best;
    s.hot_offset = grid.offset;
    s.hot_count = grid.count;
    for (Prop* p : props) {
        if (p->count >= 2 && same_grid(grid, *p)) {
            p->flags |= kPropHotGrid;
        }
    }
}

// Solid of a triangle mesh (world frame): face planes when it is closed and
// convex, else its triangles, plus the bounding box.  All arithmetic is FP64
// in a fixed order, the results rounded once to FP32:
//   * each triangle's plane: unit normal u = (v1 - v0) x (v2 - v0) / |...|,
//     offset h = u.v0, turned so that the vertex centroid c lies inside
//     (u.c < h);
//   * triangles whose planes agree within 1e-9 (normal components, and
//     offsets relative to the mesh size) and whose boundary normals agree
//     form one face; faces keep the first triangle's values, in triangle
//     order;
//   * the mesh is closed when every edge (vertices welded by equal
//     coordinates) belongs to exactly two triangles; it is convex when no
//     vertex lies more than 1e-6 of the mesh size outside a face plane and c
//     lies more than that inside every face plane.
// A closed convex mesh becomes planes (clipping tests).  Any other mesh
// becomes triangles with keys = their index in the mesh.  A closed one's
// winding is made consistent (each edge traversed once each way) when it is
// not: from triangle 0, over shared welded edges, each triangle is reversed
// when it traverses an edge in the same direction as the neighbour it is
// reached from; then all are turned outwards when the signed volume is
// negative.  A mesh that is open, flat, not orientable, or in several parts
// and not wound consistently is an open solid (kSolidOpen): its triangles as
// given, hit from either side, holding no point.  The reference accepts any
// triangle mesh (OptiXTransportHost.cpp:1118-1200) and ignores the winding.
void mesh_solid(const g4go_volume_mesh& m, HostScene& s, Volume& dv) {
    const std::string who = "volume " + std::to_string(m.volume_id) + " mesh";
    if (m.vertices_mm == nullptr || m.indices == nullptr || m.n_vertices < 3 || m.n_triangles < 1) {
        fail(G4GO_ERROR_INVALID_ARGUMENT, who + ": needs at least 3 vertices and 1 triangle");
    }
    if (m.reserved != 0) {
        fail(G4GO_ERROR_INVALID_ARGUMENT, who + ": reserved field must be 0");
    }
    const uint32_t nv = m.n_vertices, nt = m.n_triangles;
    const double* v = m.vertices_mm;
    for (uint64_t i = 0; i < 3ull * nv; ++i) {
        if (!std::isfinite(v[i])) {
            fail(G4GO_ERROR_INVALID_ARGUMENT, who + ": non-finite vertex");
        }
    }
    for (uint64_t i = 0; i < 3ull * nt; ++i) {
        if (m.indices[i] >= nv) {
            fail(G4GO_ERROR_INVALID_ARGUMENT, who + ": index exceeds vertex count");
        }
    }
    double c[3] = {0.0, 0.0, 0.0};
    for (uint32_t i = 0; i < nv; ++i) {
        for (int k = 0; k < 3; ++k) {
            c[k] += v[3 * i + k];
        }
    }
    for (int k = 0; k < 3; ++k) {
        c[k] /= static_cast<double>(nv);
    }
    double size = 0.0;
    double vlo[3] = {v[0], v[1], v[2]}, vhi[3] = {v[0], v[1], v[2]};
    for (uint32_t i = 0; i < nv; ++i) {
        for (int k = 0; k < 3; ++k) {
            size = std::max(size, std::fabs(v[3 * i + k] - c[k]));
            vlo[k] = std::min(vlo[k], v[3 * i + k]);
            vhi[k] = std::max(vhi[k], v[3 * i + k]);
        }
    }
    const double tol = 1e-6 * size;
    // closed: every edge of the welded mesh in exactly two triangles;
    // consistently wound: traversed once each way
    bool closed = true, consistent = true;
    struct MeshEdge {
        uint64_t key;  // welded (min, max) vertex pair
        uint32_t tri;
        uint32_t dir;  // 1: traversed max -> min
    };
    std::vector<MeshEdge> edges;  // sorted by (key, tri): the two triangles of an edge side by side
    {
        std::vector<uint32_t> order(nv);
        for (uint32_t i = 0; i < nv; ++i) {
            order[i] = i;
        }
        const auto less = [&](uint32_t a, uint32_t b) {
            for (int k = 0; k < 3; ++k) {
                if (v[3 * a + k] != v[3 * b + k]) {
                    return v[3 * a + k] < v[3 * b + k];
                }
            }
            return a < b;
        };
        std::sort(order.begin(), order.end(), less);
        std::vector<uint32_t> weld(nv);
        for (uint32_t i = 0; i < nv; ++i) {
            const uint32_t a = order[i];
            const bool same = i > 0 && v[3 * a] == v[3 * order[i - 1]] && v[3 * a + 1] == v[3 * order[i - 1] + 1] &&
                              v[3 * a + 2] == v[3 * order[i - 1] + 2];
            weld[a] = same ? weld[order[i - 1]] : a;
        }
        edges.reserve(3ull * nt);
        for (uint32_t t = 0; t < nt; ++t) {
            for (int e = 0; e < 3; ++e) {
                const uint32_t a = weld[m.indices[3 * t + e]], b = weld[m.indices[3 * t + (e + 1) % 3]];
                if (a == b) {
                    fail(G4GO_ERROR_INVALID_ARGUMENT, who + ": degenerate triangle");
                }
                edges.push_back({(static_cast<uint64_t>(std::min(a, b)) << 32) | std::max(a, b), t, a > b ? 1u : 0u});
            }
        }
        std::sort(edges.begin(), edges.end(), [](const MeshEdge& x, const MeshEdge& y) {
            return x.key != y.key ? x.key < y.key : x.tri < y.tri;
        });
        for (size_t i = 0; i < edges.size();) {
            size_t j = i;
            while (j < edges.size() && edges[j].key == edges[i].key) {
                ++j;
            }
            if (j - i != 2) {
                closed = false;
            } else {
                consistent = consistent && edges[i].dir != edges[i + 1].dir;
            }
            i = j;
        }
    }
    struct Face64 {
        double u[3], h, s[3];
        std::vector<double> points;  // its triangles' vertices
    };
    // boundary normal of triangle t: the caller's, else u
    const auto shade_of = [&](uint32_t t, const double* u, double* out) {
        if (m.triangle_normals == nullptr) {
            for (int k = 0; k < 3; ++k) {
                out[k] = u[k];
            }
            return;
        }
        const double* n = m.triangle_normals + 3 * t;
        const double nl = std::sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
        if (!(nl > 0.0) || !std::isfinite(nl)) {
            fail(G4GO_ERROR_INVALID_ARGUMENT, who + ": invalid triangle normal");
        }
        for (int k = 0; k < 3; ++k) {
            out[k] = n[k] / nl;
        }
    };
    std::vector<Face64> faces;
    bool too_many = false;
    for (uint32_t t = 0; t < nt; ++t) {
        const double* p0 = v + 3 * m.indices[3 * t];
        const double* p1 = v + 3 * m.indices[3 * t + 1];
        const double* p2 = v + 3 * m.indices[3 * t + 2];
        const double a[3] = {p1[0] - p0[0], p1[1] - p0[1], p1[2] - p0[2]};
        const double b[3] = {p2[0] - p0[0], p2[1] - p0[1], p2[2] - p0[2]};
        const double g[3] = {a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]};
        const double len = std::sqrt(g[0] * g[0] + g[1] * g[1] + g[2] * g[2]);
        if (!(len > 0.0)) {
            fail(G4GO_ERROR_INVALID_ARGUMENT, who + ": degenerate triangle");
        }
        Face64 f;
        for (int k = 0; k < 3; ++k) {
            f.u[k] = g[k] / len;
        }
        f.h = f.u[0] * p0[0] + f.u[1] * p0[1] + f.u[2] * p0[2];
        if (f.u[0] * c[0] + f.u[1] * c[1] + f.u[2] * c[2] > f.h) {
            for (int k = 0; k < 3; ++k) {
                f.u[k] = -f.u[k];
            }
            f.h = -f.h;
        }
        shade_of(t, f.u, f.s);
        Face64* into = nullptr;
        for (Face64& o : faces) {
            bool same = std::fabs(o.h - f.h) <= 1e-9 * size;
            for (int k = 0; k < 3; ++k) {
                same = same && std::fabs(o.u[k] - f.u[k]) <= 1e-9 && std::fabs(o.s[k] - f.s[k]) <= 1e-9;
            }
            if (same) {
                into = &o;
                break;
            }
        }
        if (into == nullptr) {
            if (faces.size() == kMaxMeshPlanes) {
                too_many = true;  // not a convex solid this backend holds as planes
                break;
            }
            faces.push_back(f);
            into = &faces.back();
        }
        into->points.insert(into->points.end(), {p0[0], p0[1], p0[2], p1[0], p1[1], p1[2], p2[0], p2[1], p2[2]});
    }
    bool convex = closed && !too_many;
    for (size_t j = 0; convex && j < faces.size(); ++j) {
        const Face64& f = faces[j];
        convex = f.h - (f.u[0] * c[0] + f.u[1] * c[1] + f.u[2] * c[2]) > tol;
        for (uint32_t i = 0; convex && i < nv; ++i) {
            convex = !(f.u[0] * v[3 * i] + f.u[1] * v[3 * i + 1] + f.u[2] * v[3 * i + 2] - f.h > tol);
        }
    }
    if (convex) {
        dv.plane_begin = static_cast<uint32_t>(s.planes.size());
        for (const Face64& f : faces) {
            Plane pl{};
            for (int k = 0; k < 3; ++k) {
                pl.n[k] = static_cast<float>(f.u[k]);
                pl.shade[k] = static_cast<float>(f.s[k]);
            }
            pl.h = static_cast<float>(f.h);
            pl.across = kFaceDynamic;
            pl.surface_parent = pl.surface_across = kInvalid;
            s.planes.push_back(pl);
            s.face_points.push_back(f.points);
        }
        dv.plane_count = static_cast<uint32_t>(faces.size());
    } else {
        // An open mesh, or one whose winding cannot be made consistent, or a
        // flat one, is an open solid: its triangles as given (kSolidOpen).
        bool open = !closed;
        // reversed[t]: triangle t's winding is turned to make the mesh consistent
        std::vector<uint8_t> reversed(nt, 0);
        if (!open && !consistent) {
            std::vector<uint32_t> first(nt * 3u);  // per triangle edge: its entry in `edges`
            std::vector<uint32_t> seen(nt, 0);
            for (uint32_t i = 0; i < edges.size(); ++i) {
                first[3u * edges[i].tri + seen[edges[i].tri]++] = i;
            }
            std::vector<uint8_t> done(nt, 0);
            std::vector<uint32_t> queue{0u};
            done[0] = 1;
            for (size_t q = 0; q < queue.size(); ++q) {
                const uint32_t t = queue[q];
                for (int e = 0; e < 3; ++e) {
                    const uint32_t i = first[3u * t + e];
                    const MeshEdge& mine = edges[i];
                    const MeshEdge& other = edges[(i & 1u) == 0u ? i + 1u : i - 1u];
                    // opposite directions once both windings are applied
                    const uint8_t want = static_cast<uint8_t>((mine.dir ^ reversed[t] ^ 1u) ^ other.dir);
                    if (!done[other.tri]) {
                        done[other.tri] = 1;
                        reversed[other.tri] = want;
                        queue.push_back(other.tri);
                    } else if (reversed[other.tri] != want) {
                        open = true;  // not orientable
                    }
                }
            }
            open = open || queue.size() != nt;  // in several parts, not wound consistently
            if (open) {
                std::fill(reversed.begin(), reversed.end(), 0);
            }
        }
        // triangle t's corners in its consistent winding
        const auto corner = [&](uint32_t t, int j) {
            const int k = reversed[t] != 0u && j != 0 ? 3 - j : j;
            return m.indices[3 * t + k];
        };
        if (nt > kMaxSolidTriangles) {
            fail(G4GO_ERROR_INVALID_ARGUMENT,
                 who + ": more than " + std::to_string(kMaxSolidTriangles) + " triangles in a non-convex mesh");
        }
        // signed volume about the centroid: negative when the winding points inwards
        double vol6 = 0.0;
        for (uint32_t t = 0; t < nt; ++t) {
            double q[3][3];
            for (int j = 0; j < 3; ++j) {
                for (int k = 0; k < 3; ++k) {
                    q[j][k] = v[3 * corner(t, j) + k] - c[k];
                }
            }
            vol6 += q[0][0] * (q[1][1] * q[2][2] - q[1][2] * q[2][1]) + q[0][1] * (q[1][2] * q[2][0] - q[1][0] * q[2][2]) +
                    q[0][2] * (q[1][0] * q[2][1] - q[1][1] * q[2][0]);
        }
        open = open || !(std::fabs(vol6) > tol * size * size);  // flat: no enclosed volume
        const bool flip = !open && vol6 < 0.0;
        dv.plane_begin = static_cast<uint32_t>(s.tris.size());
        for (uint32_t t = 0; t < nt; ++t) {
            const double* p0 = v + 3 * corner(t, 0);
            const double* p1 = v + 3 * corner(t, flip ? 2 : 1);
            const double* p2 = v + 3 * corner(t, flip ? 1 : 2);
            const double a[3] = {p1[0] - p0[0], p1[1] - p0[1], p1[2] - p0[2]};
            const double b[3] = {p2[0] - p0[0], p2[1] - p0[1], p2[2] - p0[2]};
            const double g[3] = {a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]};
            const double len = std::sqrt(g[0] * g[0] + g[1] * g[1] + g[2] * g[2]);
            double u[3], sh[3];
            for (int k = 0; k < 3; ++k) {
                u[k] = g[k] / len;
            }
            shade_of(t, u, sh);
            Tri tr{};
            for (int k = 0; k < 3; ++k) {
                tr.face.n[k] = static_cast<float>(u[k]);
                tr.face.shade[k] = static_cast<float>(sh[k]);
                tr.v0[k] = static_cast<float>(p0[k]);
                tr.v1[k] = static_cast<float>(p1[k]);
                tr.v2[k] = static_cast<float>(p2[k]);
            }
            tr.face.h = static_cast<float>(u[0] * p0[0] + u[1] * p0[1] + u[2] * p0[2]);
            tr.face.across = kFaceDynamic;
            tr.face.surface_parent = tr.face.surface_across = kInvalid;
            tr.key = t;
            s.tris.push_back(tr);
            s.tri_points.push_back({p0[0], p0[1], p0[2], p1[0], p1[1], p1[2], p2[0], p2[1], p2[2]});
        }
        dv.plane_count = nt | kSolidTriangles | (open ? kSolidOpen : 0u);
    }
    // bounding box rounded outwards
    for (int k = 0; k < 3; ++k) {
        dv.lo[k] = static_cast<float>(vlo[k]);
        if (static_cast<double>(dv.lo[k]) > vlo[k]) {
            dv.lo[k] = std::nextafter(dv.lo[k], -INFINITY);
        }
        dv.hi[k] = static_cast<float>(vhi[k]);
        if (static_cast<double>(dv.hi[k]) < vhi[k]) {
            dv.hi[k] = std::nextafter(dv.hi[k], INFINITY);
        }
    }
}

// BVH over the triangles [begin, begin + count) of a triangle solid (median
// split of the centroids on the longest axis, leaves of <= 4; the triangles
// and their points are reordered in place).  Node boxes are widened by 1e-6
// of their coordinates so that a hit on a triangle is never culled by the
// rounding of the slab test.  Returns the root node.
uint32_t build_tri_bvh(HostScene& s, uint32_t begin, uint32_t count) {
    struct Task {
        uint32_t node, first, count, depth;
    };
    const auto tri_lo = [&](const Tri& t, int k) { return std::min({t.v0[k], t.v1[k], t.v2[k]}); };
    const auto tri_hi = [&](const Tri& t, int k) { return std::max({t.v0[k], t.v1[k], t.v2[k]}); };
    const uint32_t root = static_cast<uint32_t>(s.bvh.size());
    s.bvh.push_back(BvhNode{});
    std::vector<Task> stack{{root, begin, count, 0u}};
    std::vector<uint32_t> perm;
    while (!stack.empty()) {
        const Task t = stack.back();
        stack.pop_back();
        if (t.depth > static_cast<uint32_t>(kBvhMaxDepth)) {
            fail(G4GO_ERROR_INVALID_ARGUMENT, "triangle BVH deeper than the device traversal stack");
        }
        BvhNode node{};
        float clo[3], chi[3];
        for (int k = 0; k < 3; ++k) {
            node.lo[k] = clo[k] = std::numeric_limits<float>::infinity();
            node.hi[k] = chi[k] = -std::numeric_limits<float>::infinity();
        }
        for (uint32_t i = t.first; i < t.first + t.count; ++i) {
            for (int k = 0; k < 3; ++k) {
                node.lo[k] = std::min(node.lo[k], tri_lo(s.tris[i], k));
                node.hi[k] = std::max(node.hi[k], tri_hi(s.tris[i], k));
                const float cc = tri_lo(s.tris[i], k) + tri_hi(s.tris[i], k);
                clo[k] = std::min(clo[k], cc);
                chi[k] = std::max(chi[k], cc);
            }
        }
        for (int k = 0; k < 3; ++k) {
            const float pad = 1e-6f * std::max(std::fabs(node.lo[k]), std::fabs(node.hi[k])) + 1e-6f;
            node.lo[k] -= pad;
            node.hi[k] += pad;
        }
        if (t.count <= 4) {
            node.left = t.first;
            node.right_or_count = t.count | kBvhLeaf;
            s.bvh[t.node] = node;
            continue;
        }
        int axis = 0;
        for (int k = 1; k < 3; ++k) {
            if (chi[k] - clo[k] > chi[axis] - clo[axis]) {
                axis = k;
            }
        }
        perm.resize(t.count);
        for (uint32_t i = 0; i < t.count; ++i) {
            perm[i] = t.first + i;
        }
        const uint32_t half = t.count / 2;
        std::nth_element(perm.begin(), perm.begin() + half, perm.end(), [&](uint32_t a, uint32_t b) {
            const float ca = tri_lo(s.tris[a], axis) + tri_hi(s.tris[a], axis);
            const float cb = tri_lo(s.tris[b], axis) + tri_hi(s.tris[b], axis);
            return ca < cb || (ca == cb && s.tris[a].key < s.tris[b].key);
        });
        std::vector<Tri> tris(t.count);
        std::vector<std::vector<double>> pts(t.count);
        for (uint32_t i = 0; i < t.count; ++i) {
            tris[i] = s.tris[perm[i]];
            pts[i] = std::move(s.tri_points[perm[i]]);
        }
        for (uint32_t i = 0; i < t.count; ++i) {
            s.tris[t.first + i] = tris[i];
            s.tri_points[t.first + i] = std::move(pts[i]);
        }
        node.left = static_cast<uint32_t>(s.bvh.size());
        node.right_or_count = node.left + 1;
        s.bvh.push_back(BvhNode{});
        s.bvh.push_back(BvhNode{});
        s.bvh[t.node] = node;
        stack.push_back({node.left, t.first, half, t.depth + 1});
        stack.push_back({node.right_or_count, t.first + half, t.count - half, t.depth + 1});
    }
    return root;
}

HostScene build_scene(const g4go_scene& in, float tolerance) {
    HostScene s;
    if (in.n_volumes == 0 || in.volumes == nullptr) {
        fail(G4GO_ERROR_INVALID_ARGUMENT, "scene has no volumes");
    }
    if ((in.n_materials != 0 && in.materials == nullptr) || (in.n_surfaces != 0 && in.surfaces == nullptr) ||
        (in.n_bindings != 0 && in.bindings == nullptr)) {
        fail(G4GO_ERROR_INVALID_ARGUMENT, "scene array pointer is null");
    }
    for (uint32_t i = 0; i < in.n_materials; ++i) {
        const g4go_material& m = in.materials[i];
        const std::string tag = "material " + std::to_string(i);
        Material dm{};
        dm.rindex = add_property(s, m.rindex, tag + " RINDEX");
        dm.group_velocity = add_property(s, m.group_velocity_mm_per_ns, tag + " GROUPVEL");
        dm.abs_length = add_property(s, m.abs_length_mm, tag + " ABSLENGTH");
        dm.rayleigh_length = add_property(s, m.rayleigh_length_mm, tag + " RAYLEIGH");
        for (int k = 0; k < 3; ++k) {
            dm.spectrum[k] = add_spectrum(s, m.scintillation_spectrum[k], tag + " spectrum");
        }
        dm.rindex_max = 1.0f;
        if (m.rindex.count != 0) {
            dm.rindex_max = *std::max_element(m.rindex.values, m.rindex.values + m.rindex.count);
        }
        s.materials.push_back(dm);
    }
    for (uint32_t i = 0; i < in.n_surfaces; ++i) {
        const g4go_surface& f = in.surfaces[i];
        const std::string tag = "surface " + std::to_string(i);
        if (f.type > 1 || f.model > 1 || f.finish > 1) {
            fail(G4GO_ERROR_INVALID_ARGUMENT, tag + ": unknown type, model or finish");
        }
        Surface ds{};
        ds.type = f.type;
        ds.model = f.model;
        ds.finish = f.finish;
        ds.model_value = f.model_value;
        ds.reflectivity = add_property(s, f.reflectivity, tag + " REFLECTIVITY");
        ds.efficiency = add_property(s, f.efficiency, tag + " EFFICIENCY");
        ds.transmittance = add_property(s, f.transmittance, tag + " TRANSMITTANCE");
        ds.lobe = add_property(s, f.specular_lobe, tag + " SPECULARLOBECONSTANT");
        ds.spike = add_property(s, f.specular_spike, tag + " SPECULARSPIKECONSTANT");
        ds.backscatter = add_property(s, f.backscatter, tag + " BACKSCATTERCONSTANT");
        ds.roughness = add_property(s, f.surface_roughness, tag + " SURFACEROUGHNESS");
        s.surfaces.push_back(ds);
    }
    // volumes: dense indices in the given order
    s.volume_ids.resize(in.n_volumes);
    for (uint32_t i = 0; i < in.n_volumes; ++i) {
        const uint32_t id = in.volumes[i].volume_id;
        if (id == kInvalid || !s.index_of.emplace(id, i).second) {
            fail(G4GO_ERROR_INVALID_ARGUMENT, "duplicate or invalid volume ID " + std::to_string(id));
        }
        s.volume_ids[i] = id;
    }
    const auto map_id = [&](uint32_t id, const char* what) -> uint32_t {
        const auto it = s.index_of.find(id);
        if (it == s.index_of.end()) {
            fail(G4GO_ERROR_INVALID_ARGUMENT, std::string("unknown ") + what + " volume ID " + std::to_string(id));
        }
        return it->second;
    };
    s.world = map_id(in.world_volume_id, "world");
    // mesh solids, by volume
    if (in.n_meshes != 0 && in.meshes == nullptr) {
        fail(G4GO_ERROR_INVALID_ARGUMENT, "scene array pointer is null");
    }
    std::vector<const g4go_volume_mesh*> mesh_of(in.n_volumes, nullptr);
    for (uint32_t i = 0; i < in.n_meshes; ++i) {
        const uint32_t vi = map_id(in.meshes[i].volume_id, "mesh");
        if (mesh_of[vi] != nullptr) {
            fail(G4GO_ERROR_INVALID_ARGUMENT, "volume " + std::to_string(in.meshes[i].volume_id) + " has two meshes");
        }
        mesh_of[vi] = &in.meshes[i];
    }
    for (uint32_t i = 0; i < in.n_volumes; ++i) {
        const g4go_volume& v = in.volumes[i];
        Volume dv{};
        if (mesh_of[i] != nullptr) {
            mesh_solid(*mesh_of[i], s, dv);
        } else {
            for (int k = 0; k < 3; ++k) {
                if (!std::isfinite(v.center_mm[k]) || !(v.half_mm[k] > 0.0f) || !std::isfinite(v.half_mm[k])) {
                    fail(G4GO_ERROR_INVALID_ARGUMENT, "volume " + std::to_string(v.volume_id) + ": invalid box");
                }
                dv.lo[k] = static_cast<float>(static_cast<double>(v.center_mm[k]) - v.half_mm[k]);
                dv.hi[k] = static_cast<float>(static_cast<double>(v.center_mm[k]) + v.half_mm[k]);
            }
        }
        if (v.material_id >= in.n_materials) {
            fail(G4GO_ERROR_INVALID_ARGUMENT, "volume " + std::to_string(v.volume_id) + ": unknown material");
        }
        if (v.skin_surface_id != kInvalid && v.skin_surface_id >= in.n_surfaces) {
            fail(G4GO_ERROR_INVALID_ARGUMENT, "volume " + std::to_string(v.volume_id) + ": unknown skin surface");
        }
        dv.material = v.material_id;
        dv.parent = i == s.world ? kInvalid : map_id(v.parent_volume_id, "parent");
        if (i == s.world && v.parent_volume_id != kInvalid) {
            fail(G4GO_ERROR_INVALID_ARGUMENT, "the world volume must not have a parent");
        }
        if (i != s.world && dv.parent == i) {
            fail(G4GO_ERROR_INVALID_ARGUMENT, "volume is its own parent");
        }
        dv.skin = v.skin_surface_id;
        dv.sensor = v.sensor_id;
        dv.entry_surface = kInvalid;
        dv.bvh_root = kInvalid;
        s.volumes.push_back(dv);
    }
    // every volume must reach the world
    for (uint32_t i = 0; i < in.n_volumes; ++i) {
        uint32_t v = i, hops = 0;
        while (v != s.world) {
            v = s.volumes[v].parent;
            if (v == kInvalid || ++hops > in.n_volumes) {
                fail(G4GO_ERROR_INVALID_ARGUMENT, "volume hierarchy does not reach the world");
            }
        }
    }
    // daughter lists
    for (uint32_t p = 0; p < in.n_volumes; ++p) {
        s.volumes[p].daughter_begin = static_cast<uint32_t>(s.daughters.size());
        for (uint32_t c = 0; c < in.n_volumes; ++c) {
            if (s.volumes[c].parent == p) {
                s.daughters.push_back(c);
            }
        }
        s.volumes[p].daughter_count = static_cast<uint32_t>(s.daughters.size()) - s.volumes[p].daughter_begin;
    }
    // border surfaces, sorted by (from, to) (OptiXTransportHost.cpp:1263-1277)
    for (uint32_t i = 0; i < in.n_bindings; ++i) {
        const g4go_surface_binding& b = in.bindings[i];
        if (b.surface_id >= in.n_surfaces) {
            fail(G4GO_ERROR_INVALID_ARGUMENT, "surface binding references an unknown surface");
        }
        s.bindings.push_back(Binding{map_id(b.from_volume_id, "boundary source"),
                                     map_id(b.to_volume_id, "boundary target"), b.surface_id, 0});
    }
    // stable: among duplicate (from, to) pairs the first given wins, as in Scene::FindBoundarySurface
    std::stable_sort(s.bindings.begin(), s.bindings.end(),
                     [](const Binding& a, const Binding& b) { return a.from < b.from || (a.from == b.from && a.to < b.to); });
    // entry surfaces and face neighbours
    for (uint32_t i = 0; i < in.n_volumes; ++i) {
        Volume& v = s.volumes[i];
        if (v.parent != kInvalid) {
            v.entry_surface = host_find_surface(s, v.parent, i);
        }
        for (int axis = 0; axis < 3; ++axis) {
            for (int side = 0; side < 2; ++side) {
                const int face = 2 * axis + side;
                v.faces[face].neighbor = kInvalid;
                v.faces[face].surface = kInvalid;
                if (v.plane_count != 0) {
                    v.faces[face].neighbor = kFaceDynamic;  // unused: a mesh solid's faces are its planes
                    continue;
                }
                if (v.parent == kInvalid) {
                    continue;  // leaving the world: escape
                }
                const float plane = side ? v.hi[axis] : v.lo[axis];
                const int a1 = (axis + 1) % 3, a2 = (axis + 2) % 3;
                std::vector<uint32_t> touching;
                bool covered = false;
                bool mesh_near = false;
                const Volume& par = s.volumes[v.parent];
                for (uint32_t k = 0; k < par.daughter_count; ++k) {
                    const uint32_t j = s.daughters[par.daughter_begin + k];
                    if (j == i) {
                        continue;
                    }
                    const Volume& o = s.volumes[j];
                    if (o.plane_count != 0) {
                        // a mesh sibling whose bounding box reaches the face: found at run time
                        mesh_near = mesh_near ||
                                    (o.lo[axis] <= plane + tolerance && o.hi[axis] >= plane - tolerance &&
                                     o.lo[a1] <= v.hi[a1] + tolerance && o.hi[a1] >= v.lo[a1] - tolerance &&
                                     o.lo[a2] <= v.hi[a2] + tolerance && o.hi[a2] >= v.lo[a2] - tolerance);
                        continue;
                    }
                    const float oplane = side ? o.lo[axis] : o.hi[axis];
                    if (std::fabs(oplane - plane) > tolerance) {
                        continue;
                    }
                    const bool overlap = o.lo[a1] < v.hi[a1] - tolerance && o.hi[a1] > v.lo[a1] + tolerance &&
                                         o.lo[a2] < v.hi[a2] - tolerance && o.hi[a2] > v.lo[a2] + tolerance;
                    if (!overlap) {
                        continue;
                    }
                    touching.push_back(j);
                    covered = o.lo[a1] <= v.lo[a1] + tolerance && o.hi[a1] >= v.hi[a1] - tolerance &&
                              o.lo[a2] <= v.lo[a2] + tolerance && o.hi[a2] >= v.hi[a2] - tolerance;
                }
                uint32_t next;
                if (mesh_near) {
                    v.faces[face].neighbor = kFaceDynamic;
                    continue;
                }
                if (touching.empty()) {
                    next = v.parent;
                } else if (touching.size() == 1 && covered) {
                    next = touching[0];
                } else {
                    v.faces[face].neighbor = kFaceDynamic;
                    continue;
                }
                v.faces[face].neighbor = next;
                v.faces[face].surface = host_find_surface(s, i, next);
            }
        }
    }
    // faces that the lean box step serves (Volume::box_step_faces)
    for (Volume& v : s.volumes) {
        if (v.plane_count != 0) {
            continue;  // a mesh solid: the word holds plane_begin
        }
        v.box_step_faces = 0;
        if (v.daughter_count != 0) {
            continue;
        }
        for (int face = 0; face < 6; ++face) {
            const Face& f = v.faces[face];
            if (f.neighbor == kInvalid || f.neighbor == kFaceDynamic || f.surface >= s.surfaces.size() ||
                s.volumes[f.neighbor].material >= in.n_materials) {
                continue;
            }
            const Surface& sf = s.surfaces[f.surface];
            if (sf.type == kDielectricMetal && sf.finish == kPolished) {
                v.box_step_faces |= 1u << face;
            }
        }
    }
    // daughter BVHs for volumes with many daughters
    for (uint32_t i = 0; i < in.n_volumes; ++i) {
        Volume& v = s.volumes[i];
        if (v.daughter_count > 8) {
            build_bvh(s, v, v.daughter_begin, v.daughter_count);
        }
    }
    // triangle BVHs (the root sits in the otherwise unused first face record)
    // for solids of more than 64 triangles; smaller ones are tested in a plain
    // loop, which beats the walk's node tests there (30% on a 20-triangle solid)
    for (uint32_t i = 0; i < in.n_volumes; ++i) {
        Volume& v = s.volumes[i];
        if ((v.plane_count & kSolidTriangles) != 0) {
            const uint32_t n = v.plane_count & kSolidCountMask;
            v.faces[0].neighbor = n > 64 ? build_tri_bvh(s, v.plane_begin, n) : kInvalid;
        }
    }
    // what lies across each mesh face: the siblings whose boxes meet the box
    // of the face's points offset by the boundary tolerance along the outward
    // normal, widened by a margin far above rounding (a superset of the
    // siblings the run-time point query can find there)
    for (uint32_t i = 0; i < in.n_volumes; ++i) {
        const Volume& v = s.volumes[i];
        const bool tri = (v.plane_count & kSolidTriangles) != 0;
        const uint32_t n_faces = v.plane_count & kSolidCountMask;
        for (uint32_t k = v.plane_begin; k < v.plane_begin + n_faces; ++k) {
            Plane& pl = tri ? s.tris[k].face : s.planes[k];
            if (v.parent == kInvalid) {
                pl.across = kInvalid;  // leaving the world: escape
                continue;
            }
            const std::vector<double>& pts = tri ? s.tri_points[k] : s.face_points[k];
            double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
            double scale = 0.0;
            for (size_t j = 0; j < pts.size(); j += 3) {
                for (int a = 0; a < 3; ++a) {
                    const double x = pts[j + a] + 2.0 * tolerance * pl.n[a];
                    lo[a] = std::min(lo[a], x);
                    hi[a] = std::max(hi[a], x);
                    scale = std::max(scale, std::fabs(x));
                }
            }
            const double margin = 1e-3 + 1e-5 * scale + 2.0 * tolerance;
            std::vector<uint32_t> meets;  // daughter-list slots, ascending
            const Volume& par = s.volumes[v.parent];
            for (uint32_t slot = par.daughter_begin; slot < par.daughter_begin + par.daughter_count; ++slot) {
                const uint32_t j = s.daughters[slot];
                if (j == i) {
                    continue;
                }
                const Volume& o = s.volumes[j];
                bool meet = true;
                for (int a = 0; a < 3; ++a) {
                    meet = meet && o.lo[a] <= hi[a] + margin && o.hi[a] >= lo[a] - margin;
                }
                if (meet) {
                    meets.push_back(slot);
                }
            }
            pl.surface_parent = host_find_surface(s, i, v.parent);
            pl.sib_first = pl.sib_count = 0;
            if (meets.empty()) {
                pl.across = kInvalid;
            } else if (meets.size() == 1) {
                pl.across = meets[0];
                pl.surface_across = host_find_surface(s, i, s.daughters[meets[0]]);
            } else {
                pl.across = kFaceDynamic;
                if (meets.size() <= kMaxFaceSiblings) {
                    pl.sib_first = static_cast<uint32_t>(s.face_sibs.size());
                    pl.sib_count = static_cast<uint32_t>(meets.size());
                    s.face_sibs.insert(s.face_sibs.end(), meets.begin(), meets.end());
                }
            }
        }
    }
    s.face_points.clear();
    s.tri_points.clear();
    if (s.pool.empty()) {
        s.pool.push_back(0.0f);
    }
    mark_hot_grid(s);
    return s;
}

// ---------------------------------------------------------------------------
// Device buffers
// ---------------------------------------------------------------------------
struct DevBuf {
    void* ptr = nullptr;
    size_t cap = 0;
    void ensure(size_t bytes, const char* what) {
        if (bytes <= cap) {
            return;
        }
        if (ptr != nullptr) {
            (void)hipFree(ptr);
            ptr = nullptr;
            cap = 0;
        }
        hip_check(hipMalloc(&ptr, std::max<size_t>(bytes, 256)), what);
        cap = std::max<size_t>(bytes, 256);
    }
    void release() {
        if (ptr != nullptr) {
            (void)hipFree(ptr);
        }
        ptr = nullptr;
        cap = 0;
    }
    template <class T>
    T* as() const {
        return static_cast<T*>(ptr);
    }
};

struct PinBuf {
    void* ptr = nullptr;
    size_t cap = 0;
    void ensure(size_t bytes, const char* what) {
        if (bytes <= cap) {
            return;
        }
        if (ptr != nullptr) {
            (void)hipHostFree(ptr);
            ptr = nullptr;
            cap = 0;
        }
        hip_check(hipHostMalloc(&ptr, std::max<size_t>(bytes, 256), hipHostMallocDefault), what);
        cap = std::max<size_t>(bytes, 256);
    }
    void release() {
        if (ptr != nullptr) {
            (void)hipHostFree(ptr);
        }
        ptr = nullptr;
        cap = 0;
    }
    template <class T>
    T* as() const {
        return static_cast<T*>(ptr);
    }
};

size_t align256(size_t x) { return (x + 255) & ~static_cast<size_t>(255); }

}  // namespace

// One batch slot: its own stream, events, buffers and a copy of the scene.
struct Core {
    g4go_config cfg{};
    g4go_launch launch{};
    std::string error;

    bool hip_ready = false;
    int device = 0;
    int cu_count = 256;
    int lds_per_cu = 0;  // bytes of LDS per CU (0 until the device is set up: the smallest scene budget)
    hipStream_t stream = nullptr;
    hipEvent_t ev_start = nullptr, ev_kernel0 = nullptr, ev_kernel1 = nullptr, ev_meta = nullptr;
    // per-stage timers: H2D copies end, launch-order pass end, hit readback
    hipEvent_t ev_h2d = nullptr, ev_order = nullptr, ev_hits0 = nullptr, ev_hits1 = nullptr;
    // the caller's stream (g4go_set_stream): each batch waits for the work
    // queued on it before the batch was enqueued
    hipStream_t user_stream = nullptr;
    hipEvent_t ev_user = nullptr;
    char* scratch = nullptr;  // device address of this device's noalloc_scratch

    HostScene scene;
    bool scene_built = false;
    bool scene_uploaded = false;
    DevBuf d_scene;
    SceneView view{};

    // batch slot
    bool busy = false;
    int source = kSourceSynthetic;
    uint64_t n_photons = 0;
    uint32_t n_items = 0;
    uint64_t device_hits = 0;  // detections of the last completed batch left in d_hits
    uint32_t chunk = 64;
    std::vector<uint32_t> event_ids;
    // validated records of an emission batch (reused between batches)
    std::vector<Emission> staged;
    std::vector<uint32_t> staged_counts, staged_events;
    std::vector<uint64_t> captured;
    std::vector<uint32_t> captured_u32() const { return std::vector<uint32_t>(captured.begin(), captured.end()); }
    std::chrono::steady_clock::time_point started;
    uint64_t h2d_bytes = 0;

    DevBuf d_inputs, d_hits, d_event_stats, d_meta;
    DevBuf d_starts;  // start states of the pre-generation pass (kStartQuads float4 per photon)
    // launch index of each hit, and the launch-order copies (cfg.sort_hits)
    DevBuf d_hit_launch, d_hits_ordered, d_launch_ordered, d_order;
    bool ordered = false;  // the last batch's final hits are the launch-order copies
    const Hit* final_hits() const { return ordered ? d_hits_ordered.as<Hit>() : d_hits.as<Hit>(); }
    const uint32_t* final_launch() const { return ordered ? d_launch_ordered.as<uint32_t>() : d_hit_launch.as<uint32_t>(); }
    // differentiation
    struct ParamTable {
        uint32_t kind, index, offset, count;
    };
    std::vector<ParamTable> param_tables;
    uint32_t n_params = 0;
    std::vector<uint32_t> param_map;  // [abs(n_mat)][rayleigh(n_mat)][reflectivity(n_surf)][efficiency(n_surf)]
    DevBuf d_param_map, d_adjoint;
    KernelParams last_params{};
    int last_source = -1;
    bool last_valid = false;
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    KernelParams graph_params{};
    int graph_grid = 0, graph_block = 0, graph_source = -1;
    PinBuf h_stage, h_meta, h_hits;

    uint64_t out_n_hits = 0;  // hits of the last completed batch in h_hits
    std::vector<g4go_event_stats> out_events;
    uint64_t out_hist[kHistSlots] = {};

    // d_meta layout: [counters (kCounterCount u64)] [hit_count u32, pad u32] [histogram u64 x kHistSlots]
    static constexpr size_t kMetaCounters = 0;
    static constexpr size_t kMetaHitCount = kCounterCount * 8;
    static constexpr size_t kMetaHist = kMetaHitCount + 8;
    static constexpr size_t kMetaBytes = kMetaHist + kHistSlots * 8;

    void init_hip() {
        if (hip_ready) {
            return;
        }
        if (cfg.device >= 0) {
            hip_check(hipSetDevice(cfg.device), "hipSetDevice");
        }
        hip_check(hipGetDevice(&device), "hipGetDevice");
        hip_check(hipDeviceGetAttribute(&cu_count, hipDeviceAttributeMultiprocessorCount, device),
                  "hipDeviceGetAttribute");
        hip_check(hipDeviceGetAttribute(&lds_per_cu, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor, device),
                  "hipDeviceGetAttribute");
        hip_check(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking), "hipStreamCreate");
        hip_check(hipEventCreate(&ev_start), "hipEventCreate");
        hip_check(hipEventCreate(&ev_kernel0), "hipEventCreate");
        hip_check(hipEventCreate(&ev_kernel1), "hipEventCreate");
        hip_check(hipEventCreate(&ev_meta), "hipEventCreate");
        for (hipEvent_t* e : {&ev_h2d, &ev_order, &ev_hits0, &ev_hits1}) {
            hip_check(hipEventCreate(e), "hipEventCreate");
        }
        hip_check(hipEventCreateWithFlags(&ev_user, hipEventDisableTiming), "hipEventCreate");
        scratch = static_cast<char*>(noalloc_scratch_address(device));
        hip_ready = true;
    }

    void upload_scene() {
        if (scene_uploaded) {
            return;
        }
        const HostScene& s = scene;
        size_t off = 0;
        const size_t o_pool = off;
        off = align256(off + s.pool.size() * sizeof(float));
        const size_t o_mat = off;
        off = align256(off + std::max<size_t>(1, s.materials.size()) * sizeof(Material));
        const size_t o_surf = off;
        off = align256(off + std::max<size_t>(1, s.surfaces.size()) * sizeof(Surface));
        const size_t o_vol = off;
        off = align256(off + s.volumes.size() * sizeof(Volume));
        const size_t o_bind = off;
        off = align256(off + std::max<size_t>(1, s.bindings.size()) * sizeof(Binding));
        const size_t o_bvh = off;
        off = align256(off + std::max<size_t>(1, s.bvh.size()) * sizeof(BvhNode));
        const size_t o_dbox = off;
        off = align256(off + std::max<size_t>(1, s.daughters.size()) * sizeof(DaughterBox));
        const size_t o_planes = off;  // planes, then triangles, then the face sibling lists
        const size_t sib_quads = (s.face_sibs.size() + 3) / 4;
        off = align256(off + std::max<size_t>(1, s.planes.size() * sizeof(Plane) + s.tris.size() * sizeof(Tri) +
                                                     sib_quads * 16));
        std::vector<char> blob(off, 0);
        // daughter boxes in daughter-list order (after the BVH build reordered the list)
        for (size_t k = 0; k < s.daughters.size(); ++k) {
            const Volume& d = s.volumes[s.daughters[k]];
            uint32_t mesh = 0;
            if (d.plane_count != 0) {
                const bool encodable = (d.plane_count & kSolidTriangles) == 0 &&
                                       d.plane_begin < (1u << (32 - kMeshBeginShift)) && d.plane_count < kMeshInVolume;
                mesh = encodable ? d.plane_count | d.plane_begin << kMeshBeginShift : kMeshInVolume;
            }
            DaughterBox b{{d.lo[0], d.lo[1], d.lo[2]}, s.daughters[k], {d.hi[0], d.hi[1], d.hi[2]}, mesh};
            std::memcpy(blob.data() + o_dbox + k * sizeof(DaughterBox), &b, sizeof(b));
        }
        if (!s.planes.empty()) std::memcpy(blob.data() + o_planes, s.planes.data(), s.planes.size() * sizeof(Plane));
        if (!s.tris.empty()) {
            std::memcpy(blob.data() + o_planes + s.planes.size() * sizeof(Plane), s.tris.data(),
                        s.tris.size() * sizeof(Tri));
        }
        if (!s.face_sibs.empty()) {
            std::memcpy(blob.data() + o_planes + s.planes.size() * sizeof(Plane) + s.tris.size() * sizeof(Tri),
                        s.face_sibs.data(), s.face_sibs.size() * sizeof(uint32_t));
        }
        std::memcpy(blob.data() + o_pool, s.pool.data(), s.pool.size() * sizeof(float));
        if (!s.materials.empty()) std::memcpy(blob.data() + o_mat, s.materials.data(), s.materials.size() * sizeof(Material));
        if (!s.surfaces.empty()) std::memcpy(blob.data() + o_surf, s.surfaces.data(), s.surfaces.size() * sizeof(Surface));
        std::memcpy(blob.data() + o_vol, s.volumes.data(), s.volumes.size() * sizeof(Volume));
        if (!s.bindings.empty()) std::memcpy(blob.data() + o_bind, s.bindings.data(), s.bindings.size() * sizeof(Binding));
        if (!s.bvh.empty()) std::memcpy(blob.data() + o_bvh, s.bvh.data(), s.bvh.size() * sizeof(BvhNode));
        d_scene.ensure(blob.size(), "hipMalloc scene");
        hip_check(hipMemcpy(d_scene.ptr, blob.data(), blob.size(), hipMemcpyHostToDevice), "hipMemcpy scene");
        char* base = d_scene.as<char>();
        view = SceneView{};
        view.pool = reinterpret_cast<const float*>(base + o_pool);
        view.n_pool = static_cast<uint32_t>(s.pool.size());
        view.materials = reinterpret_cast<const Material*>(base + o_mat);
        view.surfaces = reinterpret_cast<const Surface*>(base + o_surf);
        view.volumes = reinterpret_cast<const Volume*>(base + o_vol);
        view.bindings = reinterpret_cast<const Binding*>(base + o_bind);
        view.bvh = reinterpret_cast<const BvhNode*>(base + o_bvh);
        view.dbox = reinterpret_cast<const DaughterBox*>(base + o_dbox);
        view.planes = reinterpret_cast<const Plane*>(base + o_planes);
        view.n_planes = static_cast<uint32_t>(s.planes.size());
        view.n_tris = static_cast<uint32_t>(s.tris.size());
        view.n_sib_quads = static_cast<uint32_t>(sib_quads);
        for (const Volume& v : s.volumes) {
            view.max_daughters = std::max(view.max_daughters, v.daughter_count);
            view.open_solids += (v.plane_count & kSolidOpen) != 0u ? 1u : 0u;
        }
        view.n_daughters = static_cast<uint32_t>(s.daughters.size());
        view.hot_offset = s.hot_offset;
        view.hot_count = s.hot_count;
        view.n_materials = static_cast<uint32_t>(s.materials.size());
        view.n_surfaces = static_cast<uint32_t>(s.surfaces.size());
        view.n_volumes = static_cast<uint32_t>(s.volumes.size());
        view.n_bindings = static_cast<uint32_t>(s.bindings.size());
        view.n_bvh_nodes = static_cast<uint32_t>(s.bvh.size());
        view.world = s.world;
        scene_uploaded = true;
    }

    void release() {
        if (stream != nullptr) {
            (void)hipStreamSynchronize(stream);
        }
        drop_graph();
        d_param_map.release();
        d_hit_launch.release();
        d_hits_ordered.release();
        d_launch_ordered.release();
        d_order.release();
        d_adjoint.release();
        d_scene.release();
        d_inputs.release();
        d_starts.release();
        d_hits.release();
        d_event_stats.release();
        d_meta.release();
        h_stage.release();
        h_meta.release();
        h_hits.release();
        for (hipEvent_t* e : {&ev_start, &ev_kernel0, &ev_kernel1, &ev_meta, &ev_user, &ev_h2d, &ev_order, &ev_hits0,
                              &ev_hits1}) {
            if (*e != nullptr) {
                (void)hipEventDestroy(*e);
                *e = nullptr;
            }
        }
        if (stream != nullptr) {
            (void)hipStreamDestroy(stream);
            stream = nullptr;
        }
    }

    KernelParams base_params() const {
        KernelParams p;
        std::memset(&p, 0, sizeof(p));  // padding included: captured parameters are compared bytewise
        p.scene = view;
        p.rng = Rng{static_cast<uint32_t>(cfg.seed), static_cast<uint32_t>(cfg.seed >> 32)};
        p.cfg = StepConfig{cfg.boundary_tolerance_mm};
        p.max_steps = cfg.max_bounces_per_photon;
        p.refill_threshold = std::max<uint32_t>(1, std::min<uint32_t>(64, launch.refill_threshold));
        // Every block stages scene parts in LDS, each when it still fits the
        // launch form's budget, in the order a step needs them: the volume
        // and surface records and the mesh planes (read every step), the BVH
        // nodes and daughter boxes (read by photons in a volume with a
        // daughter BVH), the property pool (read mostly when a photon
        // starts), the material records (read when it changes volume).
        const uint64_t budget = lds_scene_budget(launch.block_size, static_cast<size_t>(lds_per_cu));
        uint64_t used = 0;
        const auto fits = [&](uint64_t bytes) {
            if (bytes == 0 || used + bytes > budget) {
                return false;
            }
            used += bytes;
            return true;
        };
        if (fits(static_cast<uint64_t>(view.n_volumes) * sizeof(Volume))) p.lds_records |= kLdsVolumes;
        if (fits(static_cast<uint64_t>(view.n_surfaces) * sizeof(Surface))) p.lds_records |= kLdsSurfaces;
        if (fits(mesh_record_bytes(view))) p.lds_records |= kLdsPlanes;
        // (a mesh face's sibling test reads the daughter boxes too)
        const uint64_t entries = static_cast<uint64_t>(view.n_bvh_nodes) + view.n_daughters;
        if ((view.n_bvh_nodes != 0 || has_meshes(view)) && fits(entries * 32u)) {
            p.lds_entries = static_cast<uint32_t>(entries);
        }
        // The pool and the material records only within the smallest budget:
        // past it they would take the LDS of the ready-photon ring, which
        // serves a scene better (the 8x8 array: 6% slower with its 17 KB pool
        // staged in 512-thread blocks).
        const auto fits_small = [&](uint64_t bytes) { return used + bytes <= kLdsSceneBytes && fits(bytes); };
        if (fits_small((static_cast<uint64_t>(view.n_pool) + 3u) / 4u * 16u)) p.lds_pool_floats = view.n_pool;
        if (fits_small(static_cast<uint64_t>(view.n_materials) * sizeof(Material))) p.lds_records |= kLdsMaterials;
        return p;
    }

    // Device input layout, sized by capacities so that every pointer stays
    // fixed while a batch fits (a captured graph can then be replayed):
    // [header 256 B: n_sources, chunk, n_events][emissions][event ids]
    // [item offsets][source photon counts][source event indices].
    struct Layout {
        uint64_t cap_em = 0, cap_ev = 0, cap_src = 0;
        size_t o_em = 0, o_ev = 0, o_off = 0, o_cnt = 0, o_sev = 0, o_poff = 0, bytes = 0;
    };
    Layout layout;

    void ensure_layout(uint64_t n_em, uint64_t n_ev, uint64_t n_src) {
        if (n_em <= layout.cap_em && n_ev <= layout.cap_ev && n_src <= layout.cap_src && d_inputs.ptr != nullptr) {
            return;
        }
        Layout l;
        l.cap_em = std::max<uint64_t>({n_em, layout.cap_em, 1});
        l.cap_ev = std::max<uint64_t>({n_ev, layout.cap_ev, 1});
        l.cap_src = std::max<uint64_t>({n_src, layout.cap_src, 1});
        l.cap_em += l.cap_em / 2;
        l.cap_ev += l.cap_ev / 2;
        l.cap_src += l.cap_src / 2;
        // the small arrays first and the records last, so that one copy of
        // [0, o_em + record bytes) moves a batch's inputs
        l.o_ev = 256;
        l.o_off = l.o_ev + align256(l.cap_ev * sizeof(uint32_t));
        l.o_cnt = l.o_off + align256((l.cap_src + 1) * sizeof(uint32_t));
        l.o_sev = l.o_cnt + align256(l.cap_src * sizeof(uint32_t));
        l.o_poff = l.o_sev + align256(l.cap_src * sizeof(uint32_t));
        l.o_em = l.o_poff + align256((l.cap_src + 1) * sizeof(uint32_t));
        l.bytes = l.o_em + align256(l.cap_em * sizeof(Emission));
        d_inputs.release();
        d_inputs.ensure(l.bytes, "hipMalloc batch inputs");
        h_stage.ensure(l.bytes, "hipHostMalloc staging");
        layout = l;
    }

    void drop_graph() {
        if (graph_exec != nullptr) {
            (void)hipGraphExecDestroy(graph_exec);
            graph_exec = nullptr;
        }
        if (graph != nullptr) {
            (void)hipGraphDestroy(graph);
            graph = nullptr;
        }
    }

    // Stage the batch, allocate outputs, launch (plain, plain per event, or
    // persistent, the latter replayed from a captured graph when enabled)
    // and queue the metadata readback.  `emissions` is null for the
    // synthetic source (one source per event).
    void run_batch(const Emission* emissions, uint64_t n_emissions, const std::vector<uint32_t>& source_counts,
                   const std::vector<uint32_t>& source_event, const SyntheticSource* synth) {
        init_hip();
        if (user_stream != nullptr) {
            hip_check(hipEventRecord(ev_user, user_stream), "hipEventRecord caller stream");
            hip_check(hipStreamWaitEvent(stream, ev_user, 0), "hipStreamWaitEvent caller stream");
        }
        upload_scene();
        const uint64_t n_events = event_ids.size();
        const uint64_t n_sources = source_counts.size();
        std::vector<uint32_t> offsets(n_sources + 1, 0), photon_offsets(n_sources + 1, 0);
        for (uint64_t i = 0; i < n_sources; ++i) {
            offsets[i + 1] = offsets[i] + (source_counts[i] + chunk - 1) / chunk;
            photon_offsets[i + 1] = photon_offsets[i] + source_counts[i];  // < 2^32: checked at enqueue
        }
        n_items = offsets[n_sources];
        ensure_layout(n_emissions, n_events, n_sources);
        const Layout& l = layout;
        d_hits.ensure(std::max<uint64_t>(n_photons, 1) * sizeof(Hit), "hipMalloc hits");
        d_hit_launch.ensure(std::max<uint64_t>(n_photons, 1) * sizeof(uint32_t), "hipMalloc hit launch indices");
        // kEventReplicas accumulation copies, then the reduced array
        d_event_stats.ensure((kEventReplicas + 1) * l.cap_ev * sizeof(EventStats), "hipMalloc event stats");
        const size_t replica_bytes = kEventReplicas * l.cap_ev * sizeof(EventStats);
        EventStats* const d_events_out = d_event_stats.as<EventStats>() + kEventReplicas * l.cap_ev;
        d_meta.ensure(kMetaBytes, "hipMalloc counters");
        h_meta.ensure(kMetaBytes + n_events * sizeof(EventStats), "hipHostMalloc counters");

        char* st = h_stage.as<char>();
        auto* header = reinterpret_cast<uint32_t*>(st);
        header[0] = static_cast<uint32_t>(n_sources);
        header[1] = chunk;
        header[2] = static_cast<uint32_t>(n_events);
        header[3] = 0;
        const size_t em_bytes = n_emissions * sizeof(Emission);
        const size_t ev_bytes = n_events * sizeof(uint32_t);
        const size_t off_bytes = (n_sources + 1) * sizeof(uint32_t);
        const size_t cnt_bytes = n_sources * sizeof(uint32_t);
        const size_t sev_bytes = source_event.size() * sizeof(uint32_t);
        if (em_bytes) std::memcpy(st + l.o_em, emissions, em_bytes);
        if (ev_bytes) std::memcpy(st + l.o_ev, event_ids.data(), ev_bytes);
        std::memcpy(st + l.o_off, offsets.data(), off_bytes);
        if (cnt_bytes) std::memcpy(st + l.o_cnt, source_counts.data(), cnt_bytes);
        if (sev_bytes) std::memcpy(st + l.o_sev, source_event.data(), sev_bytes);
        const size_t poff_bytes = (n_sources + 1) * sizeof(uint32_t);
        std::memcpy(st + l.o_poff, photon_offsets.data(), poff_bytes);
        char* in = d_inputs.as<char>();
        hip_check(hipEventRecord(ev_start, stream), "hipEventRecord");
        // one copy (the unused tails of the small arrays ride along)
        h2d_bytes = l.o_em + em_bytes;
        hip_check(hipMemcpyAsync(in, st, h2d_bytes, hipMemcpyHostToDevice, stream), "hipMemcpyAsync batch inputs");
        hip_check(hipEventRecord(ev_h2d, stream), "hipEventRecord");

        KernelParams p = base_params();
        if (synth != nullptr) {
            p.synth = *synth;
        }
        p.header = reinterpret_cast<const uint32_t*>(in);
        p.emissions = reinterpret_cast<const Emission*>(in + l.o_em);
        p.event_ids = reinterpret_cast<const uint32_t*>(in + l.o_ev);
        p.item_offsets = reinterpret_cast<const uint32_t*>(in + l.o_off);
        p.source_counts = reinterpret_cast<const uint32_t*>(in + l.o_cnt);
        p.source_event = reinterpret_cast<const uint32_t*>(in + l.o_sev);
        p.source_photon_offset = reinterpret_cast<const uint32_t*>(in + l.o_poff);
        p.hits = d_hits.as<Hit>();
        p.hit_capacity = static_cast<uint32_t>(std::min<size_t>(d_hits.cap / sizeof(Hit), 0xffffffffu));
        p.hit_count = reinterpret_cast<uint32_t*>(d_meta.as<char>() + kMetaHitCount);
        // launch indices order the hits (sort_hits) and let the adjoint find an
        // emission hit's record; a synthetic batch without ordering needs neither
        p.hit_launch = cfg.sort_hits != 0 || source == kSourceEmission ? d_hit_launch.as<uint32_t>() : nullptr;
        p.histogram = reinterpret_cast<unsigned long long*>(d_meta.as<char>() + kMetaHist);
        p.event_stats = d_event_stats.as<EventStats>();
        p.event_stride = static_cast<uint32_t>(l.cap_ev);
        p.counters = reinterpret_cast<unsigned long long*>(d_meta.as<char>() + kMetaCounters);
        last_params = p;
        last_source = source;
        last_valid = true;

        const int block = static_cast<int>(launch.block_size);
        const uint32_t waves_per_block = launch.block_size / 64;
        const auto zero_outputs = [&] {
            if (launch_zero(d_meta.ptr, kMetaBytes, stream) != 0 ||
                launch_zero(d_event_stats.ptr, replica_bytes, stream) != 0) {
                fail(G4GO_ERROR_RUNTIME, "zero kernel launch failed");
            }
        };
        const auto reduce_events = [&] {
            if (launch_reduce_event_stats(p.event_stats, p.event_stride, p.header, d_events_out, stream) != 0) {
                fail(G4GO_ERROR_RUNTIME, "event statistics reduction launch failed");
            }
        };
        if (launch.persistent) {
            const int grid = static_cast<int>(std::max<uint32_t>(1, launch.blocks_per_cu) * static_cast<uint32_t>(cu_count));
            // pre-generation: the start states of every photon by a full-wave pass
            // before the transport (not for triangle solids, whose kernels generate)
            if (launch.pregenerate != 0 && view.n_tris == 0) {
                d_starts.ensure(std::max<uint64_t>(n_photons, 1) * kStartQuads * sizeof(float4), "hipMalloc start states");
                p.starts = d_starts.as<float4>();
                last_params.starts = nullptr;  // the adjoint replays by generating
            }
            const int gen_grid = 8 * cu_count;
            const auto generate = [&] {
                return p.starts == nullptr || launch_generate(p, source, const_cast<float4*>(p.starts), gen_grid, stream) == 0;
            };
            if (launch.graph) {
                // zero kernels + generation + transport kernel, captured once per parameter set
                if (graph_exec == nullptr || std::memcmp(&graph_params, &p, sizeof(p)) != 0 || graph_grid != grid ||
                    graph_block != block || graph_source != source) {
                    drop_graph();
                    hip_check(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
                    bool ok = launch_zero(d_meta.ptr, kMetaBytes, stream) == 0 &&
                              launch_zero(d_event_stats.ptr, replica_bytes, stream) == 0 && generate() &&
                              launch_transport(p, source, true, block, grid, stream) == 0 &&
                              launch_reduce_event_stats(p.event_stats, p.event_stride, p.header, d_events_out,
                                                        stream) == 0;
                    const hipError_t end = hipStreamEndCapture(stream, &graph);
                    if (!ok || end != hipSuccess) {
                        drop_graph();
                        fail(G4GO_ERROR_RUNTIME, "graph capture of the transport launch failed");
                    }
                    hip_check(hipGraphInstantiate(&graph_exec, graph, nullptr, nullptr, 0), "hipGraphInstantiate");
                    graph_params = p;
                    graph_grid = grid;
                    graph_block = block;
                    graph_source = source;
                }
                hip_check(hipEventRecord(ev_kernel0, stream), "hipEventRecord");
                hip_check(hipGraphLaunch(graph_exec, stream), "hipGraphLaunch");
            } else {
                zero_outputs();
                hip_check(hipEventRecord(ev_kernel0, stream), "hipEventRecord");
                if (!generate()) {
                    fail(G4GO_ERROR_RUNTIME, "generation kernel launch failed");
                }
                if (launch_transport(p, source, true, block, grid, stream) != 0) {
                    fail(G4GO_ERROR_RUNTIME, "transport kernel launch failed");
                }
                reduce_events();
            }
        } else {
            zero_outputs();
            hip_check(hipEventRecord(ev_kernel0, stream), "hipEventRecord");
            if (!launch.per_event) {
                const int grid = static_cast<int>((static_cast<uint64_t>(n_items) + waves_per_block - 1) / waves_per_block);
                if (grid > 0 && launch_transport(p, source, false, block, grid, stream) != 0) {
                    fail(G4GO_ERROR_RUNTIME, "transport kernel launch failed");
                }
            } else {
                // one launch per event: runs of consecutive sources of the same event
                uint64_t s0 = 0;
                while (s0 < n_sources) {
                    uint64_t s1 = s0 + 1;
                    const uint32_t ev = source_event.empty() ? static_cast<uint32_t>(s0) : source_event[s0];
                    while (!source_event.empty() && s1 < n_sources && source_event[s1] == ev) {
                        ++s1;
                    }
                    KernelParams q = p;
                    q.item_begin = offsets[s0];
                    q.item_end = offsets[s1];
                    const uint32_t items = q.item_end - q.item_begin;
                    if (items != 0) {
                        const int grid = static_cast<int>((items + waves_per_block - 1) / waves_per_block);
                        if (launch_transport(q, source, false, block, grid, stream) != 0) {
                            fail(G4GO_ERROR_RUNTIME, "transport kernel launch failed");
                        }
                    }
                    s0 = s1;
                }
            }
            reduce_events();
        }
        hip_check(hipEventRecord(ev_kernel1, stream), "hipEventRecord");
        ordered = cfg.sort_hits != 0 && n_photons != 0;
        if (ordered) {
            // hits in the launch order of their photons (a stable compaction)
            const uint64_t words = (n_photons + 31) / 32;
            const uint64_t blocks = (words + kOrderWordsPerBlock - 1) / kOrderWordsPerBlock;
            d_hits_ordered.ensure(std::max<uint64_t>(n_photons, 1) * sizeof(Hit), "hipMalloc ordered hits");
            d_launch_ordered.ensure(std::max<uint64_t>(n_photons, 1) * sizeof(uint32_t), "hipMalloc ordered launch");
            d_order.ensure((2 * words + blocks) * sizeof(uint32_t), "hipMalloc order scratch");
            OrderParams o{};
            o.hits = d_hits.as<Hit>();
            o.launch = d_hit_launch.as<uint32_t>();
            o.hit_count = p.hit_count;
            o.capacity = static_cast<uint32_t>(std::min<uint64_t>(n_photons, p.hit_capacity));
            o.n_photons = static_cast<uint32_t>(n_photons);
            o.mask = d_order.as<uint32_t>();
            o.word_prefix = o.mask + words;
            o.block_prefix = o.word_prefix + words;
            o.out_hits = d_hits_ordered.as<Hit>();
            o.out_launch = d_launch_ordered.as<uint32_t>();
            if (launch_order_hits(o, stream) != 0) {
                fail(G4GO_ERROR_RUNTIME, "hit ordering launch failed");
            }
        }
        hip_check(hipEventRecord(ev_order, stream), "hipEventRecord");
        hip_check(hipMemcpyAsync(h_meta.ptr, d_meta.ptr, kMetaBytes, hipMemcpyDeviceToHost, stream),
                  "hipMemcpyAsync counters");
        if (n_events) {
            hip_check(hipMemcpyAsync(h_meta.as<char>() + kMetaBytes, d_events_out, n_events * sizeof(EventStats),
                                     hipMemcpyDeviceToHost, stream),
                      "hipMemcpyAsync event stats");
        }
        hip_check(hipEventRecord(ev_meta, stream), "hipEventRecord");
    }
};

namespace {

template <class F>
g4go_status guarded(Core* h, F&& f) {
    try {
        f();
        return G4GO_OK;
    } catch (const Failure& e) {
        if (h != nullptr) {
            h->error = e.message;
        }
        return e.status;
    } catch (const std::bad_alloc&) {
        if (h != nullptr) {
            h->error = "host allocation failed";
        }
        return G4GO_ERROR_RUNTIME;
    } catch (const std::exception& e) {
        if (h != nullptr) {
            h->error = e.what();
        }
        return G4GO_ERROR_RUNTIME;
    }
}

void check_launch(const g4go_launch& l) {
    if (l.block_size != 64 && l.block_size != 128 && l.block_size != 256 && l.block_size != 512) {
        fail(G4GO_ERROR_INVALID_ARGUMENT, "block_size must be 64, 128, 256 or 512");
    }
    if (l.persistent && (l.chunk == 0 || l.chunk > (1u << 20))) {
        fail(G4GO_ERROR_INVALID_ARGUMENT, "chunk must be in [1, 2^20]");
    }
    if (l.refill_threshold == 0 || l.refill_threshold > 64) {
        fail(G4GO_ERROR_INVALID_ARGUMENT, "refill_threshold must be in [1, 64]");
    }
    if (l.per_event > 1 || l.graph > 1 || (l.per_event && l.persistent) || (l.graph && !l.persistent)) {
        fail(G4GO_ERROR_INVALID_ARGUMENT, "per_event applies to the plain form and graph to the persistent form");
    }
    if (l.persistent && (l.blocks_per_cu == 0 || l.blocks_per_cu > 64)) {
        fail(G4GO_ERROR_INVALID_ARGUMENT, "blocks_per_cu must be in [1, 64]");
    }
    if (l.pregenerate > 1) {
        fail(G4GO_ERROR_INVALID_ARGUMENT, "pregenerate must be 0 or 1");
    }
}

}  // namespace

extern "C" {

const char* g4go_version(void) { return "g4go-hip 0.1 (gfx950)"; }

static g4go_status core_set_launch(Core* h, const g4go_launch* launch) {
    if (h == nullptr || launch == nullptr) {
        return G4GO_ERROR_INVALID_ARGUMENT;
    }
    return guarded(h, [&] {
        check_launch(*launch);
        if (h->busy) {
            fail(G4GO_ERROR_LOGIC, "launch form cannot change with a pending batch");
        }
        h->launch = *launch;
    });
}

static g4go_status core_get_launch(const Core* h, g4go_launch* launch) {
    if (h == nullptr || launch == nullptr) {
        return G4GO_ERROR_INVALID_ARGUMENT;
    }
    *launch = h->launch;
    return G4GO_OK;
}

static g4go_status core_prepare_scene(Core* h, const g4go_scene* scene) {
    if (h == nullptr || scene == nullptr) {
        return G4GO_ERROR_INVALID_ARGUMENT;
    }
    return guarded(h, [&] {
        if (h->busy) {
            fail(G4GO_ERROR_LOGIC, "scene cannot be rebuilt with pending batches");
        }
        HostScene s = build_scene(*scene, 0.5f * h->cfg.boundary_tolerance_mm);
        h->scene = std::move(s);
        h->scene_built = true;
        h->scene_uploaded = false;
        h->param_tables.clear();
        h->n_params = 0;
        h->param_map.clear();
        h->last_valid = false;
    });
}

static g4go_status core_enqueue_emissions(Core* h, const g4go_emission* emissions, uint64_t n_emissions,
                                   const uint32_t* event_ids, uint64_t n_events,
                                   const uint32_t* emission_event_indices) {
    if (h == nullptr) {
        return G4GO_ERROR_INVALID_ARGUMENT;
    }
    return guarded(h, [&] {
        if (!h->scene_built) {
            fail(G4GO_ERROR_LOGIC, "no scene has been prepared");
        }
        if (h->busy) {
            fail(G4GO_ERROR_LOGIC, "transport queue is full");
        }
        if ((n_emissions != 0 && (emissions == nullptr || emission_event_indices == nullptr)) ||
            (n_events != 0 && event_ids == nullptr)) {
            fail(G4GO_ERROR_INVALID_ARGUMENT, "null batch array");
        }
        if (n_events > 0xffffffffull) {
            fail(G4GO_ERROR_OVERFLOW, "event batch exceeds the supported event count");
        }
        std::unordered_map<uint32_t, uint32_t> index_of_event;
        index_of_event.reserve(n_events);
        for (uint64_t i = 0; i < n_events; ++i) {
            if (!index_of_event.emplace(event_ids[i], static_cast<uint32_t>(i)).second) {
                fail(G4GO_ERROR_INVALID_ARGUMENT, "event batch contains duplicate event IDs");
            }
        }
        std::vector<uint64_t> captured(n_events, 0);
        for (uint64_t i = 0; i < n_emissions; ++i) {
            const uint32_t ei = emission_event_indices[i];
            if (ei >= n_events || emissions[i].event_id != event_ids[ei]) {
                fail(G4GO_ERROR_INVALID_ARGUMENT, "emission event index does not match event ID");
            }
            captured[ei] += emissions[i].photon_count;
        }
        if (n_emissions > 0xffffffffull) {
            fail(G4GO_ERROR_OVERFLOW, "emission batch exceeds the supported launch size");
        }
        uint64_t photons = 0;
        for (uint64_t i = 0; i < n_emissions; ++i) {
            photons += emissions[i].photon_count;
        }
        if (photons > 0xffffffffull) {
            fail(G4GO_ERROR_OVERFLOW, "emission batch exceeds the supported photon count");
        }
        // the handle's staging vectors, kept between batches
        std::vector<Emission>& staged = h->staged;
        std::vector<uint32_t>& counts = h->staged_counts;
        std::vector<uint32_t>& source_event = h->staged_events;
        staged.resize(photons != 0 ? n_emissions : 0);
        counts.resize(staged.size());
        source_event.resize(staged.size());
        // a batch without photons is not validated further (OptiXTransportHost.cpp:609-627); otherwise
        // every record's volume and material must exist, whatever its type (:650-660)
        uint32_t last_id = kInvalid, last_index = kInvalid;  // consecutive records mostly share a volume
        for (uint64_t i = 0; i < staged.size(); ++i) {
            std::memcpy(&staged[i], &emissions[i], sizeof(Emission));
            const uint32_t id = emissions[i].volume_id;
            if (id != last_id || last_index == kInvalid) {
                const auto it = h->scene.index_of.find(id);
                if (it == h->scene.index_of.end()) {
                    fail(G4GO_ERROR_INVALID_ARGUMENT, "emission references an unknown scene volume ID");
                }
                last_id = id;
                last_index = it->second;
            }
            if (emissions[i].material_id >= h->scene.materials.size()) {
                fail(G4GO_ERROR_INVALID_ARGUMENT, "emission references an unknown scene material ID");
            }
            staged[i].volume = last_index;
            counts[i] = emissions[i].photon_count;
            source_event[i] = emission_event_indices[i];
        }
        h->source = kSourceEmission;
        h->event_ids.assign(event_ids, event_ids + n_events);
        h->captured = std::move(captured);
        h->n_photons = photons;
        h->chunk = h->launch.persistent ? h->launch.chunk : 64u;
        h->started = std::chrono::steady_clock::now();
        h->last_valid = false;
        if (photons != 0) {
            h->run_batch(staged.data(), n_emissions, counts, source_event, nullptr);
        }
        h->busy = true;
    });
}

static g4go_status core_enqueue_synthetic(Core* h, const g4go_synthetic_source* source, const uint32_t* event_ids,
                                   const uint32_t* photon_counts, uint64_t n_events) {
    if (h == nullptr) {
        return G4GO_ERROR_INVALID_ARGUMENT;
    }
    return guarded(h, [&] {
        if (!h->scene_built) {
            fail(G4GO_ERROR_LOGIC, "no scene has been prepared");
        }
        if (h->busy) {
            fail(G4GO_ERROR_LOGIC, "transport queue is full");
        }
        if (source == nullptr || (n_events != 0 && (event_ids == nullptr || photon_counts == nullptr))) {
            fail(G4GO_ERROR_INVALID_ARGUMENT, "null synthetic batch argument");
        }
        if (n_events > 0xffffffffull) {
            fail(G4GO_ERROR_OVERFLOW, "event batch exceeds the supported event count");
        }
        const auto it = h->scene.index_of.find(source->volume_id);
        if (it == h->scene.index_of.end()) {
            fail(G4GO_ERROR_INVALID_ARGUMENT, "synthetic source references an unknown volume ID");
        }
        for (int k = 0; k < 3; ++k) {
            if (!std::isfinite(source->lo_mm[k]) || !std::isfinite(source->hi_mm[k]) || source->hi_mm[k] < source->lo_mm[k]) {
                fail(G4GO_ERROR_INVALID_ARGUMENT, "synthetic source box is invalid");
            }
        }
        if (!(source->wavelength_mean_nm > 0.0f) || !(source->wavelength_sigma_nm >= 0.0f) ||
            !(source->decay_time_ns >= 0.0f)) {
            fail(G4GO_ERROR_INVALID_ARGUMENT, "synthetic source distribution is invalid");
        }
        std::unordered_set<uint32_t> seen;
        uint64_t photons = 0;
        for (uint64_t i = 0; i < n_events; ++i) {
            if (!seen.insert(event_ids[i]).second) {
                fail(G4GO_ERROR_INVALID_ARGUMENT, "event batch contains duplicate event IDs");
            }
            photons += photon_counts[i];
        }
        if (photons > 0xffffffffull) {
            fail(G4GO_ERROR_OVERFLOW, "synthetic batch exceeds the supported photon count");
        }
        SyntheticSource synth{};
        for (int k = 0; k < 3; ++k) {
            synth.lo[k] = source->lo_mm[k];
            synth.hi[k] = source->hi_mm[k];
        }
        synth.wavelength_mean_nm = source->wavelength_mean_nm;
        synth.wavelength_sigma_nm = source->wavelength_sigma_nm;
        synth.decay_ns = source->decay_time_ns;
        synth.volume = it->second;
        h->source = kSourceSynthetic;
        h->event_ids.assign(event_ids, event_ids + n_events);
        h->captured.assign(photon_counts, photon_counts + n_events);
        h->n_photons = photons;
        h->chunk = h->launch.persistent ? h->launch.chunk : 64u;
        h->started = std::chrono::steady_clock::now();
        h->last_valid = false;
        if (photons != 0) {
            h->run_batch(nullptr, 0, h->captured_u32(), {}, &synth);
        }
        h->busy = true;
    });
}

// The allocation-free synthetic batch (g4go_transport_synthetic_noalloc):
// the scratch layout is described at noalloc_scratch.  Validation mirrors
// core_enqueue_synthetic as far as host data allows (event IDs and counts are
// in device memory).
static g4go_status core_transport_synthetic_noalloc(Core* h, const g4go_synthetic_source* source,
                                                    const uint32_t* event_ids, const uint32_t* photon_counts,
                                                    uint32_t n_events, const g4go_device_outputs* outputs,
                                                    void* stream_handle) {
    if (h == nullptr) {
        return G4GO_ERROR_INVALID_ARGUMENT;
    }
    return guarded(h, [&] {
        if (source == nullptr || outputs == nullptr) {
            fail(G4GO_ERROR_INVALID_ARGUMENT, "null synthetic source or outputs");
        }
        if (n_events > kPrefixMaxSources) {
            fail(G4GO_ERROR_INVALID_ARGUMENT, "the allocation-free entry point takes at most " +
                                                  std::to_string(kPrefixMaxSources) + " events per call");
        }
        if ((n_events != 0 && (event_ids == nullptr || photon_counts == nullptr)) ||
            (outputs->hit_capacity != 0 && outputs->hits == nullptr) || outputs->hit_count == nullptr ||
            outputs->histogram == nullptr || (n_events != 0 && outputs->event_counters == nullptr) ||
            outputs->batch_counters == nullptr) {
            fail(G4GO_ERROR_INVALID_ARGUMENT, "null device array");
        }
        if (outputs->hit_capacity > 0xffffffffull) {
            fail(G4GO_ERROR_INVALID_ARGUMENT, "hit capacity must be below 2^32");
        }
        if (!h->scene_built) {
            fail(G4GO_ERROR_LOGIC, "no scene has been prepared");
        }
        if (!h->hip_ready || !h->scene_uploaded) {
            fail(G4GO_ERROR_LOGIC,
                 "the scene is not resident on the device: transport one batch with the allocating entry points first");
        }
        const auto it = h->scene.index_of.find(source->volume_id);
        if (it == h->scene.index_of.end()) {
            fail(G4GO_ERROR_INVALID_ARGUMENT, "synthetic source references an unknown volume ID");
        }
        for (int k = 0; k < 3; ++k) {
            if (!std::isfinite(source->lo_mm[k]) || !std::isfinite(source->hi_mm[k]) || source->hi_mm[k] < source->lo_mm[k]) {
                fail(G4GO_ERROR_INVALID_ARGUMENT, "synthetic source box is invalid");
            }
        }
        if (!(source->wavelength_mean_nm > 0.0f) || !(source->wavelength_sigma_nm >= 0.0f) ||
            !(source->decay_time_ns >= 0.0f)) {
            fail(G4GO_ERROR_INVALID_ARGUMENT, "synthetic source distribution is invalid");
        }
        hipStream_t stream = stream_handle != nullptr ? static_cast<hipStream_t>(stream_handle) : h->stream;
        char* const scratch = h->scratch;
        auto* header = reinterpret_cast<uint32_t*>(scratch + kScratchHeader);
        auto* counters = reinterpret_cast<unsigned long long*>(scratch + kScratchCounters);
        auto* items = reinterpret_cast<uint32_t*>(scratch + kScratchItems);
        auto* photon_offsets = reinterpret_cast<uint32_t*>(scratch + kScratchPhotons);
        auto* replicas = reinterpret_cast<EventStats*>(scratch + kScratchReplicas);
        auto* batch_counters = reinterpret_cast<unsigned long long*>(outputs->batch_counters);
        // everything that accumulates starts at zero: fill kernels, no memset
        const size_t replica_bytes = static_cast<size_t>(kEventReplicas) * n_events * sizeof(EventStats);
        if (launch_zero(counters, kCounterCount * sizeof(unsigned long long), stream) != 0 ||
            launch_zero(replicas, replica_bytes, stream) != 0 ||
            launch_zero(outputs->hit_count, sizeof(uint32_t), stream) != 0 ||
            launch_zero(outputs->histogram, kHistSlots * sizeof(uint64_t), stream) != 0) {
            fail(G4GO_ERROR_RUNTIME, "zero kernel launch failed");
        }
        if (n_events == 0) {
            if (launch_export_counters(counters, batch_counters, stream) != 0) {
                fail(G4GO_ERROR_RUNTIME, "counter export launch failed");
            }
            return;
        }
        const uint32_t chunk = h->launch.chunk >= 1u && h->launch.chunk <= (1u << 20) ? h->launch.chunk : 256u;
        if (launch_synthetic_prefix(photon_counts, n_events, chunk, header, items, photon_offsets, stream) != 0) {
            fail(G4GO_ERROR_RUNTIME, "batch prefix launch failed");
        }
        KernelParams p = h->base_params();
        p.synth.lo[0] = source->lo_mm[0];
        p.synth.lo[1] = source->lo_mm[1];
        p.synth.lo[2] = source->lo_mm[2];
        p.synth.hi[0] = source->hi_mm[0];
        p.synth.hi[1] = source->hi_mm[1];
        p.synth.hi[2] = source->hi_mm[2];
        p.synth.wavelength_mean_nm = source->wavelength_mean_nm;
        p.synth.wavelength_sigma_nm = source->wavelength_sigma_nm;
        p.synth.decay_ns = source->decay_time_ns;
        p.synth.volume = it->second;
        p.header = header;
        p.event_ids = event_ids;
        p.item_offsets = items;
        p.source_counts = photon_counts;
        p.source_event = nullptr;
        p.source_photon_offset = photon_offsets;
        p.hits = reinterpret_cast<Hit*>(outputs->hits);
        p.hit_capacity = static_cast<uint32_t>(outputs->hit_capacity);
        p.hit_count = outputs->hit_count;
        p.hit_launch = nullptr;
        p.histogram = reinterpret_cast<unsigned long long*>(outputs->histogram);
        p.event_stats = replicas;
        p.event_stride = n_events;
        p.counters = counters;
        const int block = static_cast<int>(h->launch.block_size);
        const int grid =
            static_cast<int>(std::max<uint32_t>(1, h->launch.blocks_per_cu) * static_cast<uint32_t>(h->cu_count));
        if (launch_transport(p, kSourceSynthetic, true, block, grid, stream) != 0) {
            fail(G4GO_ERROR_RUNTIME, "transport kernel launch failed");
        }
        if (launch_reduce_event_stats(replicas, n_events, header, reinterpret_cast<EventStats*>(outputs->event_counters),
                                      stream) != 0 ||
            launch_export_counters(counters, batch_counters, stream) != 0) {
            fail(G4GO_ERROR_RUNTIME, "counter reduction launch failed");
        }
    });
}

static g4go_status complete_batch(Core* h, g4go_batch_result* out, bool copy_hits,
                                  const void** hits_device) {
    if (h == nullptr || out == nullptr) {
        return G4GO_ERROR_INVALID_ARGUMENT;
    }
    if (hits_device != nullptr) {
        *hits_device = nullptr;
    }
    return guarded(h, [&] {
        if (!h->busy) {
            fail(G4GO_ERROR_LOGIC, "transport queue is empty");
        }
        h->busy = false;
        h->device_hits = 0;
        h->out_n_hits = 0;
        const uint64_t n_events = h->event_ids.size();
        g4go_batch_stats bs{};
        h->out_events.assign(n_events, g4go_event_stats{});
        std::fill(std::begin(h->out_hist), std::end(h->out_hist), 0ull);
        for (uint64_t i = 0; i < n_events; ++i) {
            h->out_events[i].event_id = h->event_ids[i];
            h->out_events[i].captured = h->captured[i];
            bs.captured += h->captured[i];
        }
        if (h->n_photons != 0) {
            hip_check(hipEventSynchronize(h->ev_meta), "hipEventSynchronize");
            const char* meta = h->h_meta.as<char>();
            const auto* counters = reinterpret_cast<const uint64_t*>(meta + Core::kMetaCounters);
            const uint32_t hit_count = *reinterpret_cast<const uint32_t*>(meta + Core::kMetaHitCount);
            const auto* hist = reinterpret_cast<const uint64_t*>(meta + Core::kMetaHist);
            const auto* es = reinterpret_cast<const EventStats*>(meta + Core::kMetaBytes);
            const uint64_t n_hits = std::min<uint64_t>(hit_count, h->n_photons);
            float kernel_ms = 0.0f, h2d_ms = 0.0f, memset_ms = 0.0f, order_ms = 0.0f, meta_ms = 0.0f, hits_ms = 0.0f;
            hip_check(hipEventElapsedTime(&kernel_ms, h->ev_kernel0, h->ev_kernel1), "hipEventElapsedTime");
            hip_check(hipEventElapsedTime(&h2d_ms, h->ev_start, h->ev_h2d), "hipEventElapsedTime");
            hip_check(hipEventElapsedTime(&memset_ms, h->ev_h2d, h->ev_kernel0), "hipEventElapsedTime");
            hip_check(hipEventElapsedTime(&order_ms, h->ev_kernel1, h->ev_order), "hipEventElapsedTime");
            hip_check(hipEventElapsedTime(&meta_ms, h->ev_order, h->ev_meta), "hipEventElapsedTime");
            const auto d2h0 = std::chrono::steady_clock::now();
            h->device_hits = n_hits;
            if (!copy_hits) {
                if (hits_device != nullptr) {
                    *hits_device = h->final_hits();
                }
            } else if (n_hits != 0) {
                h->h_hits.ensure(n_hits * sizeof(Hit), "hipHostMalloc hits");
                hip_check(hipEventRecord(h->ev_hits0, h->stream), "hipEventRecord");
                hip_check(hipMemcpyAsync(h->h_hits.ptr, h->final_hits(), n_hits * sizeof(Hit), hipMemcpyDeviceToHost,
                                         h->stream),
                          "hipMemcpyAsync hits");
                hip_check(hipEventRecord(h->ev_hits1, h->stream), "hipEventRecord");
                hip_check(hipStreamSynchronize(h->stream), "hipStreamSynchronize");
                hip_check(hipEventElapsedTime(&hits_ms, h->ev_hits0, h->ev_hits1), "hipEventElapsedTime");
                h->out_n_hits = n_hits;
            }
            const auto d2h1 = std::chrono::steady_clock::now();
            std::memcpy(h->out_hist, hist, sizeof(h->out_hist));
            for (uint64_t i = 0; i < n_events; ++i) {
                g4go_event_stats& e = h->out_events[i];
                e.detected = es[i].detected;
                e.absorbed = es[i].absorbed;
                e.escaped = es[i].escaped;
                e.truncated = es[i].truncated;
                e.max_bounce = es[i].max_bounce;
                e.invalid = es[i].invalid;
                e.zero_step = es[i].zero_step;
                bs.detected += e.detected;
                bs.absorbed += e.absorbed;
                bs.escaped += e.escaped;
                bs.truncated += e.truncated;
                bs.invalid += e.invalid;
                bs.zero_step += e.zero_step;
                bs.max_bounce = std::max(bs.max_bounce, e.max_bounce);
            }
            bs.generated = counters[kCounterPhotons];
            bs.total_steps = counters[kCounterSteps];
            bs.dropped_hits = counters[kCounterDropped];
            bs.kernel_ms = kernel_ms;
            bs.host_to_device_ms = h2d_ms;
            bs.device_to_host_ms = std::chrono::duration<double, std::milli>(d2h1 - d2h0).count();
            bs.host_to_device_bytes = h->h2d_bytes;
            bs.device_memset_ms = memset_ms;
            bs.device_order_ms = order_ms;
            bs.metadata_to_host_ms = meta_ms;
            bs.metadata_to_host_bytes = Core::kMetaBytes + n_events * sizeof(EventStats);
            bs.hits_to_host_ms = hits_ms;
            bs.hits_to_host_bytes = h->out_n_hits * sizeof(Hit);
            bs.device_to_host_bytes = bs.metadata_to_host_bytes + bs.hits_to_host_bytes;
        }
        bs.transport_time_ms =
            std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - h->started).count();
        out->hits = copy_hits && h->out_n_hits != 0 ? h->h_hits.as<g4go_hit>() : nullptr;
        out->n_hits = copy_hits ? h->out_n_hits : h->device_hits;
        out->events = h->out_events.data();
        out->n_events = h->out_events.size();
        out->histogram = h->out_hist;
        out->stats = bs;
    });
}

static g4go_status core_complete(Core* h, g4go_batch_result* out) { return complete_batch(h, out, true, nullptr); }

static g4go_status core_complete_device(Core* h, g4go_batch_result* out, const void** hits_device) {
    return complete_batch(h, out, false, hits_device);
}

static g4go_status core_device_info(Core* h, char* name, uint64_t name_len, int32_t* cu_count) {
    if (h == nullptr || name == nullptr || name_len == 0 || cu_count == nullptr) {
        return G4GO_ERROR_INVALID_ARGUMENT;
    }
    return guarded(h, [&] {
        h->init_hip();
        hipDeviceProp_t prop{};
        hip_check(hipGetDeviceProperties(&prop, h->device), "hipGetDeviceProperties");
        const std::string text = std::string(prop.gcnArchName) + " " + prop.name;
        std::strncpy(name, text.c_str(), name_len - 1);
        name[name_len - 1] = '\0';
        *cu_count = h->cu_count;
    });
}

namespace {

Prop* bound_prop(HostScene& s, uint32_t kind, uint32_t index) {
    if (kind <= 1) {
        if (index >= s.materials.size()) {
            return nullptr;
        }
        return kind == 0 ? &s.materials[index].abs_length : &s.materials[index].rayleigh_length;
    }
    if (kind <= 3 && index < s.surfaces.size()) {
        return kind == 2 ? &s.surfaces[index].reflectivity : &s.surfaces[index].efficiency;
    }
    return nullptr;
}

}  // namespace

static g4go_status core_set_parameter_layout(Core* h, const g4go_param_table* tables, uint32_t n_tables,
                                      uint32_t* n_params) {
    if (h == nullptr || (n_tables != 0 && tables == nullptr) || n_params == nullptr) {
        return G4GO_ERROR_INVALID_ARGUMENT;
    }
    return guarded(h, [&] {
        if (!h->scene_built) {
            fail(G4GO_ERROR_LOGIC, "no scene has been prepared");
        }
        if (h->busy) {
            fail(G4GO_ERROR_LOGIC, "the parameter layout cannot change with a pending batch");
        }
        const uint32_t n_mat = static_cast<uint32_t>(h->scene.materials.size());
        const uint32_t n_surf = static_cast<uint32_t>(h->scene.surfaces.size());
        std::vector<uint32_t> map(2 * n_mat + 2 * n_surf, kInvalid);
        std::vector<Core::ParamTable> out;
        uint32_t offset = 0;
        for (uint32_t t = 0; t < n_tables; ++t) {
            const Prop* prop = bound_prop(h->scene, tables[t].kind, tables[t].index);
            if (prop == nullptr) {
                fail(G4GO_ERROR_INVALID_ARGUMENT, "parameter table " + std::to_string(t) + " names no scene table");
            }
            if (prop->count == 0) {
                fail(G4GO_ERROR_INVALID_ARGUMENT, "parameter table " + std::to_string(t) + " is an absent property");
            }
            const uint32_t slot = tables[t].kind <= 1 ? tables[t].kind * n_mat + tables[t].index
                                                      : 2 * n_mat + (tables[t].kind - 2) * n_surf + tables[t].index;
            if (map[slot] != kInvalid) {
                fail(G4GO_ERROR_INVALID_ARGUMENT, "parameter table bound twice");
            }
            map[slot] = offset;
            out.push_back({tables[t].kind, tables[t].index, offset, prop->count});
            offset += prop->count;
        }
        h->param_tables = std::move(out);
        h->param_map = std::move(map);
        h->n_params = offset;
        *n_params = offset;
    });
}

static g4go_status core_get_parameters(Core* h, double* theta, uint64_t n) {
    if (h == nullptr || (n != 0 && theta == nullptr)) {
        return G4GO_ERROR_INVALID_ARGUMENT;
    }
    return guarded(h, [&] {
        if (n != h->n_params) {
            fail(G4GO_ERROR_INVALID_ARGUMENT, "theta length does not match the parameter layout");
        }
        for (const auto& t : h->param_tables) {
            const Prop* prop = bound_prop(h->scene, t.kind, t.index);
            const float* values = h->scene.pool.data() + prop->offset + prop->count;
            for (uint32_t k = 0; k < t.count; ++k) {
                theta[t.offset + k] = values[k];
            }
        }
    });
}

static g4go_status core_set_parameters(Core* h, const double* theta, uint64_t n) {
    if (h == nullptr || (n != 0 && theta == nullptr)) {
        return G4GO_ERROR_INVALID_ARGUMENT;
    }
    return guarded(h, [&] {
        if (n != h->n_params) {
            fail(G4GO_ERROR_INVALID_ARGUMENT, "theta length does not match the parameter layout");
        }
        if (h->busy) {
            fail(G4GO_ERROR_LOGIC, "parameters cannot change with a pending batch");
        }
        for (uint64_t k = 0; k < n; ++k) {
            if (!std::isfinite(theta[k])) {
                fail(G4GO_ERROR_INVALID_ARGUMENT, "theta holds a non-finite value");
            }
        }
        for (const auto& t : h->param_tables) {
            Prop* prop = bound_prop(h->scene, t.kind, t.index);
            float* values = h->scene.pool.data() + prop->offset + prop->count;
            bool constant = true;
            for (uint32_t k = 0; k < t.count; ++k) {
                values[k] = static_cast<float>(theta[t.offset + k]);
                constant = constant && values[k] == values[0];
            }
            prop->flags = (prop->flags & ~kPropConstant) | (constant ? kPropConstant : 0u);
            prop->value0 = values[0];
        }
        h->scene_uploaded = false;
        h->last_valid = false;
    });
}

static g4go_status core_adjoint(Core* h, const double* slot_weight, double* grad, uint64_t n,
                         g4go_adjoint_stats* stats) {
    if (h == nullptr || slot_weight == nullptr || (n != 0 && grad == nullptr)) {
        return G4GO_ERROR_INVALID_ARGUMENT;
    }
    return guarded(h, [&] {
        if (n != h->n_params || h->param_tables.empty()) {
            fail(G4GO_ERROR_INVALID_ARGUMENT, "grad length does not match the parameter layout");
        }
        if (h->busy || !h->last_valid) {
            fail(G4GO_ERROR_LOGIC, "the adjoint replays the last completed batch: complete it first");
        }
        for (int k = 0; k < kHistSlots; ++k) {
            if (!std::isfinite(slot_weight[k])) {
                fail(G4GO_ERROR_INVALID_ARGUMENT, "slot weights must be finite");
            }
        }
        h->init_hip();
        const size_t map_bytes = h->param_map.size() * sizeof(uint32_t);
        // [map][slot weights 258 f64][grad n f64][counters 8 u64: 4 reported, then the replay's queue head]
        const size_t o_w = align256(map_bytes), o_g = o_w + align256(kHistSlots * sizeof(double));
        const size_t o_c = o_g + align256(n * sizeof(double)), bytes = o_c + 64;
        h->d_adjoint.ensure(bytes, "hipMalloc adjoint");
        char* d = h->d_adjoint.as<char>();
        hip_check(hipMemcpyAsync(d, h->param_map.data(), map_bytes, hipMemcpyHostToDevice, h->stream), "hipMemcpyAsync");
        hip_check(hipMemcpyAsync(d + o_w, slot_weight, kHistSlots * sizeof(double), hipMemcpyHostToDevice, h->stream),
                  "hipMemcpyAsync");
        if (launch_zero(d + o_g, o_c + 64 - o_g, h->stream) != 0) {
            fail(G4GO_ERROR_RUNTIME, "zero kernel launch failed");
        }
        AdjointParams a{};
        a.fwd = h->last_params;
        a.hits = h->final_hits();
        a.hit_launch = h->final_launch();
        a.n_hits = static_cast<uint32_t>(h->device_hits);
        a.slot_weight = reinterpret_cast<const double*>(d + o_w);
        a.grad = reinterpret_cast<double*>(d + o_g);
        const uint32_t n_mat = static_cast<uint32_t>(h->scene.materials.size());
        const uint32_t n_surf = static_cast<uint32_t>(h->scene.surfaces.size());
        const auto* map = reinterpret_cast<const uint32_t*>(d);
        a.map = ParamMap{map, map + n_mat, map + 2 * n_mat, map + 2 * n_mat + n_surf, h->n_params};
        a.counters = reinterpret_cast<unsigned long long*>(d + o_c);
        // persistent replay: four 256-thread blocks per CU (four waves per
        // SIMD), fewer for a short hit list
        const int block = 256;
        const int grid = static_cast<int>(std::max<uint64_t>(
            1, std::min<uint64_t>((a.n_hits + block - 1) / block, static_cast<uint64_t>(h->cu_count) * 4)));
        hip_check(hipEventRecord(h->ev_kernel0, h->stream), "hipEventRecord");
        if (a.n_hits != 0 && launch_adjoint(a, h->last_source, block, grid, h->stream) != 0) {
            fail(G4GO_ERROR_RUNTIME, "adjoint kernel launch failed");
        }
        hip_check(hipEventRecord(h->ev_kernel1, h->stream), "hipEventRecord");
        uint64_t counters[4] = {0, 0, 0, 0};
        hip_check(hipMemcpyAsync(grad, d + o_g, n * sizeof(double), hipMemcpyDeviceToHost, h->stream), "hipMemcpyAsync");
        hip_check(hipMemcpyAsync(counters, d + o_c, sizeof(counters), hipMemcpyDeviceToHost, h->stream), "hipMemcpyAsync");
        hip_check(hipStreamSynchronize(h->stream), "hipStreamSynchronize");
        float ms = 0.0f;
        hip_check(hipEventElapsedTime(&ms, h->ev_kernel0, h->ev_kernel1), "hipEventElapsedTime");
        if (stats != nullptr) {
            *stats = g4go_adjoint_stats{counters[0], counters[1], counters[2], counters[3], ms};
        }
    });
}

static g4go_status core_trace_synthetic(Core* h, const g4go_synthetic_source* source, uint32_t event_id,
                                 uint32_t photon_id, uint32_t max_steps, float* out, uint32_t* n_steps) {
    if (h == nullptr || source == nullptr || out == nullptr || n_steps == nullptr || max_steps == 0 ||
        max_steps > 65536) {
        return G4GO_ERROR_INVALID_ARGUMENT;
    }
    return guarded(h, [&] {
        if (!h->scene_built) {
            fail(G4GO_ERROR_LOGIC, "no scene has been prepared");
        }
        const auto it = h->scene.index_of.find(source->volume_id);
        if (it == h->scene.index_of.end()) {
            fail(G4GO_ERROR_INVALID_ARGUMENT, "synthetic source references an unknown volume ID");
        }
        h->init_hip();
        h->upload_scene();
        KernelParams p = h->base_params();
        for (int k = 0; k < 3; ++k) {
            p.synth.lo[k] = source->lo_mm[k];
            p.synth.hi[k] = source->hi_mm[k];
        }
        p.synth.wavelength_mean_nm = source->wavelength_mean_nm;
        p.synth.wavelength_sigma_nm = source->wavelength_sigma_nm;
        p.synth.decay_ns = source->decay_time_ns;
        p.synth.volume = it->second;
        DevBuf dout, dn;
        dout.ensure(static_cast<size_t>(max_steps) * 16 * sizeof(float), "hipMalloc");
        dn.ensure(sizeof(uint32_t), "hipMalloc");
        if (launch_trace(p, event_id, photon_id, max_steps, dout.as<float>(), dn.as<uint32_t>(), h->stream) != 0) {
            fail(G4GO_ERROR_RUNTIME, "trace kernel launch failed");
        }
        hip_check(hipStreamSynchronize(h->stream), "hipStreamSynchronize");
        hip_check(hipMemcpy(n_steps, dn.ptr, sizeof(uint32_t), hipMemcpyDeviceToHost), "hipMemcpy");
        hip_check(hipMemcpy(out, dout.ptr, static_cast<size_t>(*n_steps) * 16 * sizeof(float), hipMemcpyDeviceToHost),
                  "hipMemcpy");
        dout.release();
        dn.release();
    });
}

static g4go_status core_eval_math(Core* h, int which, const float* in, float* out, uint64_t n) {
    if (h == nullptr || in == nullptr || out == nullptr || which < 0 || which > 2 || n > (1u << 28)) {
        return G4GO_ERROR_INVALID_ARGUMENT;
    }
    return guarded(h, [&] {
        h->init_hip();
        const size_t out_n = which == 2 ? 2 * n : n;
        DevBuf din, dout;
        din.ensure(n * sizeof(float), "hipMalloc");
        dout.ensure(out_n * sizeof(float), "hipMalloc");
        hip_check(hipMemcpy(din.ptr, in, n * sizeof(float), hipMemcpyHostToDevice), "hipMemcpy");
        if (launch_eval_math(din.as<float>(), dout.as<float>(), static_cast<int>(n), which, h->stream) != 0) {
            fail(G4GO_ERROR_RUNTIME, "eval kernel launch failed");
        }
        hip_check(hipStreamSynchronize(h->stream), "hipStreamSynchronize");
        hip_check(hipMemcpy(out, dout.ptr, out_n * sizeof(float), hipMemcpyDeviceToHost), "hipMemcpy");
        din.release();
        dout.release();
    });
}

static g4go_status core_eval_fresnel(Core* h, const float* in, float* out, uint64_t n) {
    if (h == nullptr || in == nullptr || out == nullptr || n > (1u << 24)) {
        return G4GO_ERROR_INVALID_ARGUMENT;
    }
    return guarded(h, [&] {
        h->init_hip();
        DevBuf din, dout;
        din.ensure(n * 11 * sizeof(float), "hipMalloc");
        dout.ensure(n * 14 * sizeof(float), "hipMalloc");
        hip_check(hipMemcpy(din.ptr, in, n * 11 * sizeof(float), hipMemcpyHostToDevice), "hipMemcpy");
        if (launch_eval_fresnel(din.as<float>(), dout.as<float>(), static_cast<int>(n), h->stream) != 0) {
            fail(G4GO_ERROR_RUNTIME, "eval kernel launch failed");
        }
        hip_check(hipStreamSynchronize(h->stream), "hipStreamSynchronize");
        hip_check(hipMemcpy(out, dout.ptr, n * 14 * sizeof(float), hipMemcpyDeviceToHost), "hipMemcpy");
        din.release();
        dout.release();
    });
}

// ---------------------------------------------------------------------------
// The handle: a ring of batch slots.  With max_in_flight = 1 (the reference's
// single slot, OptiXTransportHost.cpp:523-525) a second enqueue before the
// completion fails with "transport queue is full"; with up to kMaxSlots the
// next batches are enqueued while earlier ones run, so host staging and
// readback overlap the device work (each slot has its own stream and buffers).
// Batches complete in enqueue order.
// ---------------------------------------------------------------------------
}  // extern "C"

constexpr uint32_t kMaxSlots = 4;

struct g4go_transport {
    g4go_config cfg{};
    std::string error;
    std::vector<std::unique_ptr<Core>> cores;
    uint32_t head = 0;   // slot of the oldest batch in flight
    uint32_t count = 0;  // batches in flight
    int last = -1;       // slot of the last completed batch
    Core* slot(uint32_t k) { return cores[k % cores.size()].get(); }
    g4go_status pass(Core* c, g4go_status s) {
        if (s != G4GO_OK) {
            error = c->error;
        }
        return s;
    }
    g4go_status logic(const char* message) {
        error = message;
        return G4GO_ERROR_LOGIC;
    }
    template <class F>
    g4go_status each(F&& f) {
        for (auto& c : cores) {
            const g4go_status s = f(c.get());
            if (s != G4GO_OK) {
                error = c->error;
                return s;
            }
        }
        return G4GO_OK;
    }
};

extern "C" {

g4go_status g4go_create(const g4go_config* config, g4go_transport** out) {
    if (out == nullptr || config == nullptr) {
        g_create_error = "null argument";
        return G4GO_ERROR_INVALID_ARGUMENT;
    }
    *out = nullptr;
    if (!(config->boundary_tolerance_mm > 0.0f) || !std::isfinite(config->boundary_tolerance_mm)) {
        g_create_error = "boundary tolerance must be positive";
        return G4GO_ERROR_INVALID_ARGUMENT;
    }
    if (config->max_bounces_per_photon == 0) {
        g_create_error = "max bounces must be positive";
        return G4GO_ERROR_INVALID_ARGUMENT;
    }
    if (config->max_in_flight > kMaxSlots) {
        g_create_error = "at most 4 batches can be in flight";
        return G4GO_ERROR_INVALID_ARGUMENT;
    }
    try {
        auto t = std::make_unique<g4go_transport>();
        t->cfg = *config;
        const uint32_t n = std::max<uint32_t>(1, config->max_in_flight);
        for (uint32_t k = 0; k < n; ++k) {
            auto c = std::make_unique<Core>();
            c->cfg = *config;
            c->launch = g4go_launch{1, 512, 256, 4, 2, 0, 0, 0};  // persistent, tuned on crystal batches
            t->cores.push_back(std::move(c));
        }
        *out = t.release();
        return G4GO_OK;
    } catch (...) {
        g_create_error = "host allocation failed";
        return G4GO_ERROR_RUNTIME;
    }
}

void g4go_destroy(g4go_transport* t) {
    if (t == nullptr) {
        return;
    }
    for (auto& c : t->cores) {
        c->release();
    }
    delete t;
}

const char* g4go_last_error(const g4go_transport* t) { return t == nullptr ? g_create_error.c_str() : t->error.c_str(); }

g4go_status g4go_set_launch(g4go_transport* t, const g4go_launch* launch) {
    if (t == nullptr || launch == nullptr) {
        return G4GO_ERROR_INVALID_ARGUMENT;
    }
    if (t->count != 0) {
        return t->logic("launch form cannot change with a pending batch");
    }
    return t->each([&](Core* c) { return core_set_launch(c, launch); });
}

g4go_status g4go_get_launch(const g4go_transport* t, g4go_launch* launch) {
    if (t == nullptr) {
        return G4GO_ERROR_INVALID_ARGUMENT;
    }
    return core_get_launch(t->cores[0].get(), launch);
}

g4go_status g4go_prepare_scene(g4go_transport* t, const g4go_scene* scene) {
    if (t == nullptr || scene == nullptr) {
        return G4GO_ERROR_INVALID_ARGUMENT;
    }
    if (t->count != 0) {
        return t->logic("scene cannot be rebuilt with pending batches");
    }
    t->last = -1;
    return t->each([&](Core* c) { return core_prepare_scene(c, scene); });
}

g4go_status g4go_enqueue_emissions(g4go_transport* t, const g4go_emission* emissions, uint64_t n_emissions,
                                   const uint32_t* event_ids, uint64_t n_events,
                                   const uint32_t* emission_event_indices) {
    if (t == nullptr) {
        return G4GO_ERROR_INVALID_ARGUMENT;
    }
    if (t->count == t->cores.size()) {
        return t->logic("transport queue is full");
    }
    Core* c = t->slot(t->head + t->count);
    const g4go_status s = t->pass(c, core_enqueue_emissions(c, emissions, n_emissions, event_ids, n_events,
                                                            emission_event_indices));
    if (s == G4GO_OK) {
        ++t->count;
    }
    return s;
}

g4go_status g4go_set_stream(g4go_transport* t, void* stream) {
    if (t == nullptr) {
        return G4GO_ERROR_INVALID_ARGUMENT;
    }
    return t->each([&](Core* c) {
        c->user_stream = static_cast<hipStream_t>(stream);
        return G4GO_OK;
    });
}

g4go_status g4go_enqueue_emissions_device(g4go_transport* t, const g4go_emission* emissions_device,
                                          uint64_t n_emissions, const uint32_t* event_ids, uint64_t n_events,
                                          const uint32_t* emission_event_indices_device) {
    if (t == nullptr) {
        return G4GO_ERROR_INVALID_ARGUMENT;
    }
    if (t->count == t->cores.size()) {
        return t->logic("transport queue is full");
    }
    Core* c = t->slot(t->head + t->count);
    std::vector<g4go_emission> em;
    std::vector<uint32_t> idx;
    const g4go_status copied = t->pass(c, guarded(c, [&] {
        if (n_emissions != 0 && (emissions_device == nullptr || emission_event_indices_device == nullptr)) {
            fail(G4GO_ERROR_INVALID_ARGUMENT, "null batch array");
        }
        if (n_emissions > 0xffffffffull) {
            fail(G4GO_ERROR_OVERFLOW, "emission batch exceeds the supported launch size");
        }
        c->init_hip();
        em.resize(n_emissions);
        idx.resize(n_emissions);
        if (n_emissions != 0) {
            // after the caller's work on its stream (the records' producer)
            hipStream_t s = c->user_stream != nullptr ? c->user_stream : c->stream;
            hip_check(hipMemcpyAsync(em.data(), emissions_device, n_emissions * sizeof(g4go_emission),
                                     hipMemcpyDeviceToHost, s),
                      "hipMemcpyAsync emission records");
            hip_check(hipMemcpyAsync(idx.data(), emission_event_indices_device, n_emissions * sizeof(uint32_t),
                                     hipMemcpyDeviceToHost, s),
                      "hipMemcpyAsync emission event indices");
            hip_check(hipStreamSynchronize(s), "hipStreamSynchronize");
        }
    }));
    if (copied != G4GO_OK) {
        return copied;
    }
    return g4go_enqueue_emissions(t, em.data(), n_emissions, event_ids, n_events, idx.data());
}

g4go_status g4go_enqueue_synthetic(g4go_transport* t, const g4go_synthetic_source* source, const uint32_t* event_ids,
                                   const uint32_t* photon_counts, uint64_t n_events) {
    if (t == nullptr) {
        return G4GO_ERROR_INVALID_ARGUMENT;
    }
    if (t->count == t->cores.size()) {
        return t->logic("transport queue is full");
    }
    Core* c = t->slot(t->head + t->count);
    const g4go_status s = t->pass(c, core_enqueue_synthetic(c, source, event_ids, photon_counts, n_events));
    if (s == G4GO_OK) {
        ++t->count;
    }
    return s;
}

static g4go_status complete_oldest(g4go_transport* t, g4go_batch_result* out, const void** hits_device) {
    if (t == nullptr || out == nullptr) {
        return G4GO_ERROR_INVALID_ARGUMENT;
    }
    if (t->count == 0) {
        return t->logic("transport queue is empty");
    }
    const uint32_t k = t->head % static_cast<uint32_t>(t->cores.size());
    Core* c = t->slot(k);
    const g4go_status s = t->pass(c, hits_device != nullptr ? core_complete_device(c, out, hits_device)
                                                            : core_complete(c, out));
    if (!c->busy) {  // the slot is consumed once its completion started
        t->head = (k + 1) % static_cast<uint32_t>(t->cores.size());
        --t->count;
        t->last = static_cast<int>(k);
    }
    return s;
}

g4go_status g4go_complete(g4go_transport* t, g4go_batch_result* out) { return complete_oldest(t, out, nullptr); }

g4go_status g4go_complete_device(g4go_transport* t, g4go_batch_result* out, const void** hits_device) {
    if (hits_device == nullptr) {
        const void* ignored = nullptr;
        return complete_oldest(t, out, &ignored);
    }
    return complete_oldest(t, out, hits_device);
}

g4go_status g4go_transport_synthetic_noalloc(g4go_transport* t, const g4go_synthetic_source* source,
                                             const uint32_t* event_ids_device, const uint32_t* photon_counts_device,
                                             uint32_t n_events, const g4go_device_outputs* outputs, void* stream) {
    if (t == nullptr) {
        return G4GO_ERROR_INVALID_ARGUMENT;
    }
    // any slot whose scene copy is resident serves (the slots hold the same scene)
    Core* c = t->cores[0].get();
    for (auto& k : t->cores) {
        if (k->scene_uploaded) {
            c = k.get();
            break;
        }
    }
    return t->pass(c, core_transport_synthetic_noalloc(c, source, event_ids_device, photon_counts_device, n_events,
                                                       outputs, stream));
}

g4go_status g4go_device_info(g4go_transport* t, char* name, uint64_t name_len, int32_t* cu_count) {
    if (t == nullptr) {
        return G4GO_ERROR_INVALID_ARGUMENT;
    }
    Core* c = t->cores[0].get();
    return t->pass(c, core_device_info(c, name, name_len, cu_count));
}

g4go_status g4go_set_parameter_layout(g4go_transport* t, const g4go_param_table* tables, uint32_t n_tables,
                                      uint32_t* n_params) {
    if (t == nullptr) {
        return G4GO_ERROR_INVALID_ARGUMENT;
    }
    if (t->count != 0) {
        return t->logic("the parameter layout cannot change with a pending batch");
    }
    t->last = -1;
    return t->each([&](Core* c) { return core_set_parameter_layout(c, tables, n_tables, n_params); });
}

g4go_status g4go_get_parameters(g4go_transport* t, double* theta, uint64_t n) {
    if (t == nullptr) {
        return G4GO_ERROR_INVALID_ARGUMENT;
    }
    Core* c = t->cores[0].get();
    return t->pass(c, core_get_parameters(c, theta, n));
}

g4go_status g4go_set_parameters(g4go_transport* t, const double* theta, uint64_t n) {
    if (t == nullptr) {
        return G4GO_ERROR_INVALID_ARGUMENT;
    }
    if (t->count != 0) {
        return t->logic("parameters cannot change with a pending batch");
    }
    t->last = -1;
    return t->each([&](Core* c) { return core_set_parameters(c, theta, n); });
}

g4go_status g4go_adjoint(g4go_transport* t, const double* slot_weight, double* grad, uint64_t n,
                         g4go_adjoint_stats* stats) {
    if (t == nullptr) {
        return G4GO_ERROR_INVALID_ARGUMENT;
    }
    if (t->last < 0) {
        return t->logic("the adjoint replays the last completed batch: complete it first");
    }
    Core* c = t->slot(static_cast<uint32_t>(t->last));
    return t->pass(c, core_adjoint(c, slot_weight, grad, n, stats));
}

g4go_status g4go_trace_synthetic(g4go_transport* t, const g4go_synthetic_source* source, uint32_t event_id,
                                 uint32_t photon_id, uint32_t max_steps, float* out, uint32_t* n_steps) {
    if (t == nullptr) {
        return G4GO_ERROR_INVALID_ARGUMENT;
    }
    Core* c = t->cores[0].get();
    return t->pass(c, core_trace_synthetic(c, source, event_id, photon_id, max_steps, out, n_steps));
}

g4go_status g4go_eval_math(g4go_transport* t, int which, const float* in, float* out, uint64_t n) {
    if (t == nullptr) {
        return G4GO_ERROR_INVALID_ARGUMENT;
    }
    Core* c = t->cores[0].get();
    return t->pass(c, core_eval_math(c, which, in, out, n));
}

g4go_status g4go_eval_fresnel(g4go_transport* t, const float* in, float* out, uint64_t n) {
    if (t == nullptr) {
        return G4GO_ERROR_INVALID_ARGUMENT;
    }
    Core* c = t->cores[0].get();
    return t->pass(c, core_eval_fresnel(c, in, out, n));
}

}  // extern "C"
