// device.hpp -- gfx950 device code shared by the wavefront kernels.
//
// Arithmetic follows the reference's System.Numerics / MathF semantics so the
// GPU reproduces the CPU renderer bit for bit: everything is compiled with
// -ffp-contract=off (only explicit fmaf fuses), division and sqrt are the
// correctly rounded HIP defaults, fp32 denormals are kept (Single.Epsilon
// compares in Mesh.fs:25-31 depend on them), MathF transcendentals are the
// float rounding of the double-precision result, and min/max reproduce both
// the x86 minps/maxps rule (Vector3.MinNative/MaxNative) and .NET's
// NaN-propagating Math.Min/Max.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace bnd {

constexpr float kPi = 3.14159265358979323846f;
constexpr float kInf = __builtin_inff();
constexpr float kEps = 1.40129846e-45f;   // Single.Epsilon

// ---------------------------------------------------------------- scalars
__device__ __forceinline__ float sin_f(float x) { return (float)::sin((double)x); }
__device__ __forceinline__ float cos_f(float x) { return (float)::cos((double)x); }
__device__ __forceinline__ float atan_f(float x) { return (float)::atan((double)x); }
__device__ __forceinline__ float log_f(float x) { return (float)::log((double)x); }
__device__ __forceinline__ float exp_f(float x) { return (float)::exp((double)x); }
__device__ __forceinline__ float rcp_e(float x) { return 1.0f / x; }   // MathF.ReciprocalEstimate
__device__ __forceinline__ bool isnan_(float a) { return a != a; }
// Math.Max / Math.Min (NaN-propagating; +0 > -0).
__device__ __forceinline__ float max_n(float a, float b) {
    if (isnan_(a)) return a;
    if (isnan_(b)) return b;
    if (a > b) return a;
    if (b > a) return b;
    return __builtin_signbit(b) ? a : b;
}
__device__ __forceinline__ float min_n(float a, float b) {
    if (isnan_(a)) return a;
    if (isnan_(b)) return b;
    if (a < b) return a;
    if (b < a) return b;
    return __builtin_signbit(a) ? a : b;
}
// Only the comparison outcome of the slab test is observable, so a cheaper
// NaN-propagating max/min whose result may differ in the sign of zero is used
// there: max_s(a,b) is NaN if either is NaN, else the larger value.
__device__ __forceinline__ float max_s(float a, float b) { return (a > b || isnan_(a)) ? a : b; }
__device__ __forceinline__ float min_s(float a, float b) { return (a < b || isnan_(a)) ? a : b; }
__device__ __forceinline__ int sat_int(float f) {
    if (isnan_(f)) return 0;
    if (f >= 2147483648.0f) return 2147483647;
    if (f <= -2147483648.0f) return -2147483647 - 1;
    return (int)f;
}

// ---------------------------------------------------------------- vectors
struct V3 { float x, y, z; };
struct V2 { float x, y; };
__device__ __forceinline__ V3 mk(float x, float y, float z) { return {x, y, z}; }
__device__ __forceinline__ V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ V3 operator*(V3 a, V3 b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
__device__ __forceinline__ V3 operator*(V3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
__device__ __forceinline__ V3 operator*(float s, V3 a) { return {a.x * s, a.y * s, a.z * s}; }
__device__ __forceinline__ V3 operator/(V3 a, float s) { return {a.x / s, a.y / s, a.z / s}; }
__device__ __forceinline__ V3 operator-(V3 a) { return {-a.x, -a.y, -a.z}; }
__device__ __forceinline__ float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ V3 cross(V3 a, V3 b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__device__ __forceinline__ float length(V3 a) { return sqrtf(dot(a, a)); }
__device__ __forceinline__ V3 normalize(V3 a) { return a / length(a); }
__device__ __forceinline__ V3 fma3(V3 a, V3 b, V3 c) { return {fmaf(a.x, b.x, c.x), fmaf(a.y, b.y, c.y), fmaf(a.z, b.z, c.z)}; }
__device__ __forceinline__ float comp(V3 a, int i) { return i == 0 ? a.x : (i == 1 ? a.y : a.z); }
__device__ __forceinline__ V3 vmin(V3 a, V3 b) { return {a.x < b.x ? a.x : b.x, a.y < b.y ? a.y : b.y, a.z < b.z ? a.z : b.z}; }
__device__ __forceinline__ V3 vmax(V3 a, V3 b) { return {a.x > b.x ? a.x : b.x, a.y > b.y ? a.y : b.y, a.z > b.z ? a.z : b.z}; }
__device__ __forceinline__ V3 xyz(float4 v) { return {v.x, v.y, v.z}; }
__device__ __forceinline__ V3 sel(bool c, V3 a, V3 b) { return {c ? a.x : b.x, c ? a.y : b.y, c ? a.z : b.z}; }
__device__ __forceinline__ float4 f4(V3 v, float w) { return make_float4(v.x, v.y, v.z, w); }

// Affine map stored as three columns (M1j, M2j, M3j, M4j): Vector3.Transform.
struct Xf { float4 c0, c1, c2; };
__device__ __forceinline__ V3 xf_point(const Xf& m, V3 v) {
    return {v.x * m.c0.x + v.y * m.c0.y + v.z * m.c0.z + m.c0.w,
            v.x * m.c1.x + v.y * m.c1.y + v.z * m.c1.z + m.c1.w,
            v.x * m.c2.x + v.y * m.c2.y + v.z * m.c2.z + m.c2.w};
}
// Transform(v) - Translation (Ray.fs:21, Camera.fs:15-17, Primitive.fs:35-36)
__device__ __forceinline__ V3 xf_dir(const Xf& m, V3 v) { return xf_point(m, v) - mk(m.c0.w, m.c1.w, m.c2.w); }

// ------------------------------------------------------------------ RNG
__device__ __forceinline__ uint32_t xxhash2(uint32_t x, uint32_t y) {     // Util/Hash.fs:6-15
    uint32_t h = y + 374761393u + x * 3266489917u;
    h = 668265263u * ((h << 17) | (h >> 15));
    h = 2246822519u * (h ^ (h >> 15));
    h = 3266489917u * (h ^ (h >> 13));
    return h ^ (h >> 16);
}
__device__ __forceinline__ uint32_t xxhash3(uint32_t x, uint32_t y, uint32_t z) {  // Hash.fs:17-28
    uint32_t h = z + 374761393u + x * 3266489917u;
    h = 668265263u * ((h << 17) | (h >> 15));
    h = h + y * 3266489917u;
    h = 668265263u * ((h << 17) | (h >> 15));
    h = 2246822519u * (h ^ (h >> 15));
    h = 3266489917u * (h ^ (h >> 13));
    return h ^ (h >> 16);
}
__device__ __forceinline__ float lcg(uint32_t& s) {                       // Hash.fs:30-32
    s = 0x00269ec3u + s * 0x000343fdu;
    return __uint_as_float((s >> 9) | 0x3f800000u) - 1.0f;
}

// --------------------------------------------------------------- scene
// BVH node: the reference's 32-byte layout (Util/BVH.fs:52-73) as two float4:
// a = (pmin.xyz, pmax.x), b = (pmax.yz, right/offset bits, flag bytes).
struct DevInst {
    Xf w2o, o2w;
    float4 lo;      // world bounds min, w = sphere radius
    float4 hi;      // world bounds max, w = EvalPDF constant of a mesh light
    int4 info;      // x = BLAS root node (-1: sphere), y = first triangle, z = material, w = light
    int4 aux;       // x = 1: w2o is exactly the identity (entries +0.0f and 1.0f), yzw unused
};
struct DevScene {
    const float4* tlas;
    const float4* blas;
    const float* tris;          // 9 floats per triangle, BVH order
    const DevInst* insts;
    const float4* mats;         // 2 per material: (kind bits, albedo) (ior, alpha, metallic, 0)
    const float4* lightdefs;    // (emission, two_sided)
    const int4* lrec;           // per light: instance, first triangle, triangle count, light definition
    int n_lights;
    const float* alias_prob;    // per triangle
    const int* alias_idx;
    const float* alias_pdf;
    // camera (Pinhole.fs:12-27 / ThinLens.fs:12-35, Camera.fs:12-20)
    Xf c2w;
    float4 ul;                  // viewport corner, w = push-forward
    float4 du;                  // pixel delta U, w = aperture
    float4 dv;                  // pixel delta V, w = focus distance
    int cam_kind;
    int exact_only;             // scene has non-finite geometry: exact slab tests only
};

struct Ray { V3 o, d; };
__device__ __forceinline__ V3 point_at(const Ray& r, float t) { return fma3(mk(t, t, t), r.d, r.o); }

// AxisAlignedBoundingBox.Intersect (Base/Ray.fs:29-39) with 1/d precomputed.
__device__ __forceinline__ bool slab(V3 lo, V3 hi, V3 o, V3 inv, float t) {
    float t0x = (lo.x - o.x) * inv.x, t1x = (hi.x - o.x) * inv.x;
    float t0y = (lo.y - o.y) * inv.y, t1y = (hi.y - o.y) * inv.y;
    float t0z = (lo.z - o.z) * inv.z, t1z = (hi.z - o.z) * inv.z;
    float ax = t0x < t1x ? t0x : t1x, cx = t0x > t1x ? t0x : t1x;
    float ay = t0y < t1y ? t0y : t1y, cy = t0y > t1y ? t0y : t1y;
    float az = t0z < t1z ? t0z : t1z, cz = t0z > t1z ? t0z : t1z;
    float tmin = max_s(1e-3f, max_s(ax, max_s(ay, az)));
    float tmax = min_s(t, min_s(cx, min_s(cy, cz)));
    return tmin <= tmax;
}

struct Tri { V3 p0, p1, p2; };
__device__ __forceinline__ Tri load_tri(const float* tris, int i) {
    const float* p = tris + 9 * (size_t)i;
    return {{p[0], p[1], p[2]}, {p[3], p[4], p[5]}, {p[6], p[7], p[8]}};
}
// Triangle.Intersect (Mesh.fs:24-82): returns t' when eps < t' < t, else -1.
__device__ __forceinline__ bool tri_hit(const Tri& tr, V3 o, V3 d, float t, float& tout) {
    V3 e0 = tr.p1 - tr.p0, e1 = tr.p2 - tr.p0;
    V3 q = cross(d, e1);
    float det = dot(e0, q);
    if (fabsf(det) < kEps) return false;
    float inv = 1.0f / det;
    V3 s = o - tr.p0;
    float u = inv * dot(s, q);
    if (u < 0.0f || u > 1.0f) return false;
    V3 sc = cross(s, e0);
    float v = inv * dot(d, sc);
    if (v < 0.0f || u + v > 1.0f) return false;
    float tt = inv * dot(e1, sc);
    if (tt > kEps && tt < t) { tout = tt; return true; }
    return false;
}
// SpherePrimitive.Intersect (Sphere.fs:13-77): branch 0 = c/q root, 1 = q/a root.
__device__ __forceinline__ bool sphere_hit(float radius, V3 o, V3 d, float t, float& tout, int& branch) {
    const float eps = 1e-3f;
    float a = dot(d, d);
    float b = -dot(o, d);
    float r2 = radius * radius;
    float c = dot(o, o) - r2;
    V3 fb = o + (b / a) * d;
    float disc = r2 - dot(fb, fb);
    if (disc < 0.0f) return false;
    float q = b + copysignf(sqrtf(a * disc), b);
    float t0 = c / q;
    if (t0 > eps && t0 < t) { tout = t0; branch = 0; return true; }
    float t1 = q / a;
    if (t1 > eps && t1 < t) { tout = t1; branch = 1; return true; }
    return false;
}

struct NodeView { V3 lo, hi; int word; int leaf, axis, count; };
__device__ __forceinline__ NodeView load_node(const float4* nodes, int i) {
    float4 a = nodes[2 * i], b = nodes[2 * i + 1];
    NodeView n;
    n.lo = {a.x, a.y, a.z};
    n.hi = {a.w, b.x, b.y};
    n.word = __float_as_int(b.z);
    uint32_t f = __float_as_uint(b.w);
    n.leaf = f & 0xff;
    n.axis = (int)(int8_t)((f >> 8) & 0xff);
    n.count = (int)(int8_t)((f >> 16) & 0xff);
    return n;
}

// Per-lane traversal stack: a window of K entries in LDS over a per-lane
// overflow area in global memory (entry e of lane g at spill[e * stride + g]).
template <int K>
struct LaneStack {
    uint32_t* lds;        // &shared[0][threadIdx.x], row pitch = blockDim.x
    int pitch;
    uint32_t* spill;      // &spill[gid]
    size_t spill_stride;
    int sp;
    __device__ __forceinline__ void push(uint32_t v) {
        int slot = sp & (K - 1);
        if (sp >= K) spill[(size_t)(sp - K) * spill_stride] = lds[slot * pitch];
        lds[slot * pitch] = v;
        ++sp;
    }
    __device__ __forceinline__ uint32_t pop() {
        --sp;
        int slot = sp & (K - 1);
        uint32_t v = lds[slot * pitch];
        if (sp >= K) lds[slot * pitch] = spill[(size_t)(sp - K) * spill_stride];
        return v;
    }
};

struct TraceCount { uint32_t nodes, tris, insts; };

}  // namespace bnd
