// Flat, pointer-free scene layout uploaded once per PrepareScene.
//
// Reference model: include/g4go/optical/Scene.hpp:61-164 and the device copy
// in src/optical/optix/OptiXDeviceData.cuh:68-144.  Differences:
//   * a volume's solid is an axis-aligned box (analytic slab tests), a closed
//     convex triangle mesh held as the planes of its faces (clipping tests),
//     or any other closed triangle mesh held as its triangles under a BVH
//     (watertight ray/triangle tests), in place of the OptiX traversal, with
//     a parent relation for nesting;
//   * every property table lives in one float pool (energies then values),
//     addressed by offset, so the whole scene is a handful of arrays;
//   * the border/skin surface of every static box face, and of every
//     parent->daughter entry, is resolved on the host (Scene.cpp:120-150).
#pragma once

#include <stddef.h>
#include <stdint.h>

namespace g4go {

constexpr uint32_t kFaceDynamic = 0xfffffffeu;  // neighbour found by a point query at run time

// 16 bytes, 16-byte aligned in every record that holds one, so a lookup
// reads the whole descriptor with one load (load_prop).
struct alignas(16) Prop {
    uint32_t offset;  // energies at pool[offset], values at pool[offset + count]
    uint32_t count;   // 0 = property absent
    uint32_t flags;   // kPropConstant | kPropHotGrid
    float value0;     // values[0]: the value of a constant table, read without the pool
};
// all values equal (OptiXTransportHost.cpp:356-369)
constexpr uint32_t kPropConstant = 1u;
// energies identical to the scene's hot grid (SceneView::hot_offset): evaluated
// from the node index and weight each photon computes once for its energy
constexpr uint32_t kPropHotGrid = 2u;

struct Material {
    Prop rindex;
    Prop group_velocity;
    Prop abs_length;
    Prop rayleigh_length;
    Prop spectrum[3];  // normalised CDF over energy (OptiXTransportHost.cpp:371-410)
    float rindex_max;
    uint32_t pad[3];
};

enum SurfaceType : uint32_t { kDielectricDielectric = 0, kDielectricMetal = 1 };
enum SurfaceModel : uint32_t { kGlisur = 0, kUnified = 1 };
enum SurfaceFinish : uint32_t { kPolished = 0, kGround = 1 };

struct Surface {
    uint32_t type, model, finish;
    float model_value;
    Prop reflectivity, efficiency, transmittance, lobe, spike, backscatter, roughness;
};

struct alignas(8) Face {
    uint32_t neighbor;  // volume across the face, kInvalid (leave the world) or kFaceDynamic
    uint32_t surface;   // surface for a static neighbour, kInvalid = none
};

struct alignas(16) Volume {
    float lo[3];
    uint32_t material;
    float hi[3];
    uint32_t parent;
    uint32_t skin, sensor;
    uint32_t daughter_begin, daughter_count;  // indices into the daughter list
    uint32_t entry_surface;                   // surface crossed when entering from the parent
    uint32_t bvh_root;                        // root node of the daughter BVH, kInvalid = linear list
    // convex mesh solid: planes [plane_begin, plane_begin + plane_count) of
    // the scene's plane list; triangle solid (plane_count has kSolidTriangles
    // set): triangles [plane_begin, plane_begin + count) of the triangle
    // list, BVH root in faces[0].neighbor (kInvalid: no BVH); in both, lo/hi
    // is the solid's bounding box.  plane_count 0: the box lo/hi.
    union {
        uint32_t plane_begin;  // mesh solid (plane_count != 0): its first plane or triangle
        // box solid (plane_count == 0), bit f set: a step leaving through face
        // f can take the lean box step of transport_step (no daughters, a
        // static neighbour, a polished dielectric_metal surface; resolved on
        // the host)
        uint32_t box_step_faces;
    };
    uint32_t plane_count;
    // face = 2*axis + (outward sign > 0); the neighbour and its surface sit
    // side by side so a step reads both with one 8-byte load
    Face faces[6];
};

// Records staged into LDS are copied in 16-byte words.
static_assert(sizeof(Material) % 16 == 0 && sizeof(Surface) % 16 == 0 && sizeof(Volume) % 16 == 0,
              "scene records are whole 16-byte words");

// Face plane of a convex mesh solid (world frame): the solid holds the points
// x with dot(n, x) <= h for every plane of its volume.  `shade` is the unit
// normal the boundary physics uses on the face (the caller's triangle normal,
// else n), turned to face the photon at run time.  The first 16 bytes are all
// a clipping test reads.
//
// Across the face: the host lists the siblings whose boxes reach the points
// just outside the face (the face offset by the boundary tolerance, with a
// margin).  None: the parent is across (surface_parent).  One: that sibling
// (daughter-list slot `across`, surface_across) when it holds the point, else
// the parent; the same answer as testing every sibling.  More: kFaceDynamic,
// and the point query tests only the listed siblings, in daughter-list order
// (sib_count of them from entry sib_first of the scene's sibling lists), or
// every sibling when the list would exceed kMaxFaceSiblings (sib_count 0).
struct alignas(16) Plane {
    float n[3];
    float h;
    float shade[3];
    uint32_t across;  // kInvalid, a daughter-list slot, or kFaceDynamic
    uint32_t surface_parent, surface_across;
    uint32_t sib_first, sib_count;
};
constexpr uint32_t kMaxFaceSiblings = 64;
constexpr uint32_t kMaxMeshPlanes = 1024;  // faces per convex mesh solid

constexpr uint32_t kSolidTriangles = 0x80000000u;  // Volume::plane_count flag
// Volume::plane_count flag of a triangle solid whose mesh is open (or cannot
// be wound consistently): it has no inside.  Its triangles are hit from either
// side, it holds no point (a photon inside it leaves only through a triangle,
// or into the parent where no triangle lies ahead).
constexpr uint32_t kSolidOpen = 0x40000000u;
constexpr uint32_t kSolidCountMask = 0x3fffffffu;  // Volume::plane_count without the flags
constexpr uint32_t kFaceTri = 0x80000000u;         // face reference: a triangle (else a plane)
constexpr uint32_t kMaxSolidTriangles = 65536;     // triangles per triangle solid

// Triangle of a non-convex mesh solid, wound so that (v1 - v0) x (v2 - v0)
// points out of the solid.  `face` holds its plane (outward), boundary normal
// and what lies across it, as for a convex solid's face; `key` is the
// triangle's index in the caller's mesh (ties between equal hit distances go
// to the lower key, whatever order the BVH stores the triangles in).
struct alignas(16) Tri {
    Plane face;
    float v0[3];
    uint32_t key;
    float v1[3];
    uint32_t pad1;
    float v2[3];
    uint32_t pad2;
};

struct Binding {
    uint32_t from, to, surface, pad;  // sorted by (from, to)
};

// Bounding-volume-hierarchy node over the daughter boxes of one volume.
// Leaves hold up to four daughters (indices into the daughter list).
struct BvhNode {
    float lo[3];
    uint32_t left;  // inner: left child node; leaf: first daughter-list slot
    float hi[3];
    uint32_t right_or_count;  // inner: right child node; leaf: count | kBvhLeaf
};
constexpr uint32_t kBvhLeaf = 0x80000000u;
constexpr int kBvhMaxDepth = 15;  // root at depth 0; the host rejects deeper trees

// Box of the daughter in one slot of the daughter list (slot order = the
// list's order, so BVH leaves address it directly).  32 bytes: the BVH walk
// reads two aligned 16-byte words per candidate instead of a Volume record.
struct alignas(16) DaughterBox {
    float lo[3];
    uint32_t volume;
    float hi[3];
    // 0: a box solid.  Else a mesh solid within this box: a convex one's planes
    // as plane_count | plane_begin << kMeshBeginShift (no volume record read
    // to find them), or kMeshInVolume (a triangle solid, or a plane range
    // beyond the encoding): read the volume record.
    uint32_t mesh;
};
constexpr uint32_t kMeshBeginShift = 11;  // plane_count <= kMaxMeshPlanes < 2^11
constexpr uint32_t kMeshInVolume = (1u << kMeshBeginShift) - 1u;

struct SceneView {
    const float* pool;
    const Material* materials;
    const Surface* surfaces;
    const Volume* volumes;
    const Plane* planes;  // faces of the convex mesh solids, then the triangles of the others (scene_tris)
    const Binding* bindings;
    const BvhNode* bvh;
    const DaughterBox* dbox;  // one per daughter-list slot (volume index and box)
    uint32_t n_materials, n_surfaces, n_volumes, n_bindings;
    uint32_t n_bvh_nodes, world, n_daughters;
    // The energy grid shared by the most non-constant tables (energies at
    // pool[hot_offset], hot_count nodes; hot_count = 0: none).  A photon's
    // energy never changes, so its node on this grid is located once.
    uint32_t hot_offset, hot_count;
    uint32_t n_pool;    // floats in the pool
    uint32_t n_planes;  // 0 (and n_tris 0): every solid is a box
    uint32_t n_tris;
    uint32_t n_sib_quads;  // 16-byte words of face sibling lists after the triangles
    uint32_t max_daughters;  // most daughters of any volume
    uint32_t open_solids;    // volumes that are open triangle solids
};

// Scenes with mesh solids (convex or triangle).
inline __host__ __device__ bool has_meshes(const SceneView& sc) { return (sc.n_planes | sc.n_tris) != 0u; }

// Bytes of the plane and triangle records and the face sibling lists (one array).
inline __host__ __device__ size_t mesh_record_bytes(const SceneView& sc) {
    return static_cast<size_t>(sc.n_planes) * sizeof(Plane) + static_cast<size_t>(sc.n_tris) * sizeof(Tri) +
           static_cast<size_t>(sc.n_sib_quads) * 16u;
}

// The triangle records follow the plane records in one array (in global
// memory and in the LDS copy alike), and the face sibling lists (daughter-list
// slots, four per 16-byte word) follow the triangles.
inline __host__ __device__ const Tri* scene_tris(const SceneView& sc) {
    return reinterpret_cast<const Tri*>(sc.planes + sc.n_planes);
}
inline __host__ __device__ const uint32_t* scene_face_sibs(const SceneView& sc) {
    return reinterpret_cast<const uint32_t*>(scene_tris(sc) + sc.n_tris);
}

}  // namespace g4go
