// host_scene.hpp -- host-side scene construction of the Barnacle hot path.
//
// The reference builds its acceleration structures on the CPU before the timed
// region (Extensions/Scene/Render.fs:11-14).  This is the C++ mirror of that
// host logic: System.Numerics arithmetic with .NET semantics, the binned-SAH
// BVH of Util/BVH.fs with identical node order, the degenerate alias table of
// Util/AliasTable.fs, Node.Traverse (Base/Scene.fs:43-51) and the TLAS +
// light list.  Everything is plain float arithmetic compiled with
// -ffp-contract=off so that the host and the device see the same bits.
#pragma once

#include <cmath>
#include <cstdint>
#include <cstring>
#include <functional>
#include <limits>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/barnacle_hip.h"

namespace bnh {

constexpr float kPi = 3.14159265358979323846f;  // MathF.PI
constexpr float kInf = std::numeric_limits<float>::infinity();

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

// float -> int as .NET 9 converts (saturating, NaN -> 0)
inline int sat_int(float f) {
    if (f != f) return 0;
    if (f >= 2147483648.0f) return 2147483647;
    if (f <= -2147483648.0f) return -2147483647 - 1;
    return (int)f;
}

struct V3 {
    float x, y, z;
    float operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
};
inline V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 operator*(V3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
inline V3 operator*(float s, V3 a) { return {a.x * s, a.y * s, a.z * s}; }
inline V3 operator/(V3 a, float s) { return {a.x / s, a.y / s, a.z / s}; }
inline V3 operator-(V3 a) { return {-a.x, -a.y, -a.z}; }
inline float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline V3 cross(V3 a, V3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
inline float length(V3 a) { return std::sqrt(dot(a, a)); }
inline V3 normalize(V3 a) { return a / length(a); }
// Vector3.MinNative / MaxNative (x86 minps / maxps)
inline V3 vmin(V3 a, V3 b) { return {a.x < b.x ? a.x : b.x, a.y < b.y ? a.y : b.y, a.z < b.z ? a.z : b.z}; }
inline V3 vmax(V3 a, V3 b) { return {a.x > b.x ? a.x : b.x, a.y > b.y ? a.y : b.y, a.z > b.z ? a.z : b.z}; }

// Matrix4x4 in row-vector convention, m[r*4+c]; translation is M41..M43.
struct M4 {
    float m[16];
    static M4 identity() { M4 r{}; r.m[0] = r.m[5] = r.m[10] = r.m[15] = 1.0f; return r; }
    V3 translation() const { return {m[12], m[13], m[14]}; }
};
M4 operator*(const M4& a, const M4& b);
V3 transform_point(V3 v, const M4& m);          // Vector3.Transform
inline V3 transform_dir(V3 v, const M4& m) { return transform_point(v, m) - m.translation(); }
M4 invert(const M4& m);                          // Matrix4x4.Invert
M4 keyframe_matrix(const bn_transform& t);       // KeyFrameInfo.ToKeyFrame
M4 transform_eval(const bn_transform* keys, int n, float time);   // Transform.Eval

struct Box {
    V3 lo, hi;                                    // 6 contiguous floats
    static Box empty() { return {{kInf, kInf, kInf}, {-kInf, -kInf, -kInf}}; }
    void grow(const Box& b) { lo = vmin(lo, b.lo); hi = vmax(hi, b.hi); }
    void grow(V3 p) { lo = vmin(lo, p); hi = vmax(hi, p); }
    V3 center() const { return 0.5f * (lo + hi); }
    V3 extent() const { return hi - lo; }
    float half_area_x2() const { V3 d = extent(); return 2.0f * (d.x * d.y + d.y * d.z + d.z * d.x); }
    int longest_axis() const {
        V3 d = extent();
        if (d.x >= d.y && d.x >= d.z) return 0;
        return d.y >= d.z ? 1 : 2;
    }
};
Box transform_box(const Box& b, const M4& m);    // AxisAlignedBoundingBox.Transform

// Binned-SAH build (Util/BVH.fs:128-247).  Returns nodes in the reference's
// pre-order; perm[i] is the input index of the item placed at position i.
void build_bvh(const std::vector<Box>& items, std::vector<bn_bvh_node>& nodes, std::vector<int32_t>& perm);

// AliasTable construction (Util/AliasTable.fs:14-51).
struct AliasTable {
    std::vector<int32_t> alias;
    std::vector<float> prob, pdf;
};
AliasTable build_alias(const float* w, int n);

struct HostMesh {
    bool sphere = false;
    float radius = 1.0f;
    std::vector<V3> verts;
    std::vector<int32_t> tri;            // 3 per triangle, BVH order
    std::vector<bn_bvh_node> nodes;
    AliasTable alias;
    Box bounds() const;
};

struct HostInstance {
    int prim, material, light;
    M4 o2w, w2o;
    Box bounds;
};

struct HostScene {
    std::vector<HostMesh> meshes;
    std::vector<HostInstance> insts;     // TLAS order
    std::vector<int32_t> inst_desc;      // TLAS order -> descriptor instance
    std::vector<bn_bvh_node> tlas;
    std::vector<int32_t> lights;         // TLAS-order indices with a light
    std::vector<bn_material> materials;
    std::vector<bn_light> light_defs;
    bn_camera camera{};
    M4 c2w{};
};

// Builds the BVHs of several item sets at once (sets[s]: counts[s] boxes of 6
// floats, all finite) with the nodes and permutations of build_bvh.
using BatchBvhBuild = std::function<void(const std::vector<const float*>& sets, const std::vector<int>& counts,
                                         std::vector<std::vector<bn_bvh_node>>& nodes,
                                         std::vector<std::vector<int32_t>>& perms)>;

// Scene.Load + Traverse + BVHAggregate + UniformLightSampler construction.
// With blas_build, the BLAS of every mesh whose triangle bounds are finite
// are built by it in one call; the others, and the TLAS, by build_bvh.
void build_scene(const bn_scene_desc& d, HostScene& out, const BatchBvhBuild* blas_build = nullptr);

// Film.PostProcess with tone mapping 0 identity, 1 ACES, 2 gamma, clamped to
// [0, 1], and the 8-bit RGBA of Film.Save (film.cpp).  Either output may be
// null.
void postprocess(const float* film, int64_t n, int tone_mapping, float* out_rgb, uint8_t* out_rgba8);

// Triangle area of Mesh.fs:84-87.
inline float tri_area(V3 p0, V3 p1, V3 p2) { return 0.5f * length(cross(p1 - p0, p2 - p0)); }

}  // namespace bnh
