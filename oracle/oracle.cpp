// oracle.cpp -- CPU restatement of the Barnacle hot path (test infrastructure).
//
// Compiled with -O2 -ffp-contract=off: the only fused multiply-adds are the
// explicit FusedMultiplyAdd calls of the reference.  Vector arithmetic follows
// System.Numerics component order: Dot = (x*x' + y*y') + z*z', Cross per
// component, Normalize = v / Length (three divisions), Vector3.Transform =
// ((x*M11 + y*M21) + z*M31) + M41.  MathF transcendentals are taken as the
// float rounding of the double-precision libm result (correctly rounded in all
// but vanishingly rare cases); MathF.ReciprocalEstimate is taken as an exact
// reciprocal.  Both choices are documented parity assumptions.
#include "oracle.h"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <functional>
#include <limits>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

namespace {

thread_local std::string g_err;
int fail(const std::string& m) { g_err = m; return -1; }

constexpr float PI_F = 3.14159265358979323846f;  // MathF.PI
constexpr float INF_F = std::numeric_limits<float>::infinity();

// ------------------------------------------------------------- MathF model
inline float sinF(float x) { return (float)std::sin((double)x); }
inline float cosF(float x) { return (float)std::cos((double)x); }
inline float tanF(float x) { return (float)std::tan((double)x); }
inline float atanF(float x) { return (float)std::atan((double)x); }
inline float logF(float x) { return (float)std::log((double)x); }
inline float expF(float x) { return (float)std::exp((double)x); }
inline float rcpE(float x) { return 1.0f / x; }                 // MathF.ReciprocalEstimate
// Math.Max / Math.Min / MathF.Max / MathF.Min: NaN-propagating.
inline float maxN(float a, float b) { if (a != a) return a; if (b != b) return b; return a > b ? a : (b > a ? b : (std::signbit(b) ? a : b)); }
inline float minN(float a, float b) { if (a != a) return a; if (b != b) return b; return a < b ? a : (b < a ? b : (std::signbit(a) ? a : b)); }
// float -> int conversion of .NET 9 (saturating, NaN -> 0).
inline int satInt(float f) {
    if (f != f) return 0;
    if (f >= 2147483648.0f) return 2147483647;
    if (f <= -2147483648.0f) return (-2147483647 - 1);
    return (int)f;
}

// ---------------------------------------------------------------- vectors
struct V2 { float x, y; };
struct V3 { float x, y, z; };
inline V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 operator*(V3 a, V3 b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
inline V3 operator*(V3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
inline V3 operator*(float s, V3 a) { return {a.x * s, a.y * s, a.z * s}; }
inline V3 operator/(V3 a, float s) { return {a.x / s, a.y / s, a.z / s}; }
inline V3 operator-(V3 a) { return {-a.x, -a.y, -a.z}; }
inline float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline V3 cross(V3 a, V3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
inline float length(V3 a) { return std::sqrt(dot(a, a)); }
inline V3 normalize(V3 a) { return a / length(a); }
inline V3 fma3(V3 a, V3 b, V3 c) { return {std::fma(a.x, b.x, c.x), std::fma(a.y, b.y, c.y), std::fma(a.z, b.z, c.z)}; }
inline float comp(V3 a, int i) { return i == 0 ? a.x : (i == 1 ? a.y : a.z); }
// Vector3.MinNative / MaxNative: x86 minps/maxps (second operand on NaN).
inline V3 minNat(V3 a, V3 b) { return {a.x < b.x ? a.x : b.x, a.y < b.y ? a.y : b.y, a.z < b.z ? a.z : b.z}; }
inline V3 maxNat(V3 a, V3 b) { return {a.x > b.x ? a.x : b.x, a.y > b.y ? a.y : b.y, a.z > b.z ? a.z : b.z}; }

// Matrix4x4, row-major, row-vector convention: M41..M43 = translation.
struct M4 { float m[16]; };
M4 identity() { M4 r{}; r.m[0] = r.m[5] = r.m[10] = r.m[15] = 1.0f; return r; }
M4 mul(const M4& a, const M4& b) {
    M4 r;
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j)
            r.m[i * 4 + j] = a.m[i * 4 + 0] * b.m[0 * 4 + j] + a.m[i * 4 + 1] * b.m[1 * 4 + j] +
                             a.m[i * 4 + 2] * b.m[2 * 4 + j] + a.m[i * 4 + 3] * b.m[3 * 4 + j];
    return r;
}
inline V3 xform(V3 v, const M4& M) {  // Vector3.Transform(position, matrix)
    const float* m = M.m;
    return {v.x * m[0] + v.y * m[4] + v.z * m[8] + m[12],
            v.x * m[1] + v.y * m[5] + v.z * m[9] + m[13],
            v.x * m[2] + v.y * m[6] + v.z * m[10] + m[14]};
}
inline V3 trans(const M4& M) { return {M.m[12], M.m[13], M.m[14]}; }
inline V3 xformDir(V3 v, const M4& M) { return xform(v, M) - trans(M); }  // Ray.fs:21 idiom

// Matrix4x4.Invert: cofactor expansion with 2x2 minors.
M4 invert(const M4& M) {
    const float* x = M.m;
    float a = x[0], b = x[1], c = x[2], d = x[3], e = x[4], f = x[5], g = x[6], h = x[7];
    float i = x[8], j = x[9], k = x[10], l = x[11], m = x[12], n = x[13], o = x[14], p = x[15];
    float kp_lo = k * p - l * o, jp_ln = j * p - l * n, jo_kn = j * o - k * n;
    float ip_lm = i * p - l * m, io_km = i * o - k * m, in_jm = i * n - j * m;
    float a11 = +(f * kp_lo - g * jp_ln + h * jo_kn);
    float a12 = -(e * kp_lo - g * ip_lm + h * io_km);
    float a13 = +(e * jp_ln - f * ip_lm + h * in_jm);
    float a14 = -(e * jo_kn - f * io_km + g * in_jm);
    float det = a * a11 + b * a12 + c * a13 + d * a14;
    M4 r;
    if (std::fabs(det) < std::numeric_limits<float>::denorm_min()) {
        for (float& v : r.m) v = std::numeric_limits<float>::quiet_NaN();
        return r;
    }
    float inv = 1.0f / det;
    r.m[0] = a11 * inv; r.m[4] = a12 * inv; r.m[8] = a13 * inv; r.m[12] = a14 * inv;
    r.m[1] = -(b * kp_lo - c * jp_ln + d * jo_kn) * inv;
    r.m[5] = +(a * kp_lo - c * ip_lm + d * io_km) * inv;
    r.m[9] = -(a * jp_ln - b * ip_lm + d * in_jm) * inv;
    r.m[13] = +(a * jo_kn - b * io_km + c * in_jm) * inv;
    float gp_ho = g * p - h * o, fp_hn = f * p - h * n, fo_gn = f * o - g * n;
    float ep_hm = e * p - h * m, eo_gm = e * o - g * m, en_fm = e * n - f * m;
    r.m[2] = +(b * gp_ho - c * fp_hn + d * fo_gn) * inv;
    r.m[6] = -(a * gp_ho - c * ep_hm + d * eo_gm) * inv;
    r.m[10] = +(a * fp_hn - b * ep_hm + d * en_fm) * inv;
    r.m[14] = -(a * fo_gn - b * eo_gm + c * en_fm) * inv;
    float gl_hk = g * l - h * k, fl_hj = f * l - h * j, fk_gj = f * k - g * j;
    float el_hi = e * l - h * i, ek_gi = e * k - g * i, ej_fi = e * j - f * i;
    r.m[3] = -(b * gl_hk - c * fl_hj + d * fk_gj) * inv;
    r.m[7] = +(a * gl_hk - c * el_hi + d * ek_gi) * inv;
    r.m[11] = -(a * fl_hj - b * el_hi + d * ej_fi) * inv;
    r.m[15] = +(a * fk_gj - b * ek_gi + c * ej_fi) * inv;
    return r;
}
M4 scaleM(float x, float y, float z) { M4 r = identity(); r.m[0] = x; r.m[5] = y; r.m[10] = z; return r; }
M4 translateM(float x, float y, float z) { M4 r = identity(); r.m[12] = x; r.m[13] = y; r.m[14] = z; return r; }
M4 rotX(float a) { float c = cosF(a), s = sinF(a); M4 r = identity(); r.m[5] = c; r.m[6] = s; r.m[9] = -s; r.m[10] = c; return r; }
M4 rotY(float a) { float c = cosF(a), s = sinF(a); M4 r = identity(); r.m[0] = c; r.m[2] = -s; r.m[8] = s; r.m[10] = c; return r; }
M4 rotZ(float a) { float c = cosF(a), s = sinF(a); M4 r = identity(); r.m[0] = c; r.m[1] = s; r.m[4] = -s; r.m[5] = c; return r; }

// ---------------------------------------------------------------- Hash.fs
inline uint32_t xx2(uint32_t x, uint32_t y) {            // Util/Hash.fs:6-15
    const uint32_t P2 = 2246822519u, P3 = 3266489917u, P4 = 668265263u, P5 = 374761393u;
    uint32_t h = y + P5 + x * P3;
    h = P4 * ((h << 17) | (h >> 15));
    h = P2 * (h ^ (h >> 15));
    h = P3 * (h ^ (h >> 13));
    return h ^ (h >> 16);
}
inline uint32_t xx3(uint32_t x, uint32_t y, uint32_t z) { // Util/Hash.fs:17-28
    const uint32_t P2 = 2246822519u, P3 = 3266489917u, P4 = 668265263u, P5 = 374761393u;
    uint32_t h = z + P5 + x * P3;
    h = P4 * ((h << 17) | (h >> 15));
    h = h + y * P3;
    h = P4 * ((h << 17) | (h >> 15));
    h = P2 * (h ^ (h >> 15));
    h = P3 * (h ^ (h >> 13));
    return h ^ (h >> 16);
}
struct Sampler {                                          // Base/Sampler.fs:6-16
    uint32_t s;
    float next() {                                        // Hash.LCG (Hash.fs:30-32)
        s = 0x00269ec3u + s * 0x000343fdu;
        uint32_t b = (s >> 9) | 0x3f800000u;
        float f;
        std::memcpy(&f, &b, 4);
        return f - 1.0f;
    }
    V2 next2() { float a = next(); float b = next(); return {a, b}; }
};

// ---------------------------------------------------------- AliasTable.fs
struct Alias { std::vector<int> alias; std::vector<float> prob, pdf; };
Alias alias_build(const float* w, int n) {               // AliasTable.fs:14-51
    Alias t;
    t.alias.resize(n); t.prob.resize(n); t.pdf.resize(n);
    float sum = 0.0f;
    for (int i = 0; i < n; ++i) sum = sum + w[i];         // Array.sum: left to right
    if (sum == 0.0f) {
        for (int i = 0; i < n; ++i) { t.prob[i] = 1.0f; t.alias[i] = i; t.pdf[i] = 1.0f / (float)n; }
        return t;
    }
    for (int i = 0; i < n; ++i) { t.prob[i] = w[i] / sum; t.alias[i] = i; t.pdf[i] = w[i] / sum; }
    std::vector<int> under, over;                         // Stack<int>: LIFO
    for (int i = 0; i < n; ++i) (t.prob[i] > 1.0f ? over : under).push_back(i);
    while (!under.empty() && !over.empty()) {
        int o = over.back(); over.pop_back();
        int u = under.back(); under.pop_back();
        t.prob[o] = t.prob[o] + t.prob[u] - 1.0f;
        t.alias[u] = o;
        (t.prob[o] > 1.0f ? over : under).push_back(o);
    }
    for (int o : over) { t.alias[o] = o; t.prob[o] = 1.0f; }
    for (int u : under) { t.alias[u] = u; t.prob[u] = 1.0f; }
    return t;
}
// AliasTable.Sample (:52-62).  The reference indexes without a bounds check;
// an index rounded up to n is clamped to n-1 here.
inline void alias_sample(const int* alias, const float* prob, const float* pdf, int n, float u,
                         int& idx, float& p) {
    float uu = u * (float)n;
    int i = satInt(uu);
    if (i > n - 1) i = n - 1;
    if (i < 0) i = 0;
    uu = uu - (float)i;
    if (uu < prob[i]) { idx = i; p = pdf[i]; }
    else { idx = alias[i]; p = pdf[idx]; }
}

// ---------------------------------------------------------------- BVH.fs
struct AABB { V3 mn, mx; };
inline AABB aabbDefault() { return {{INF_F, INF_F, INF_F}, {-INF_F, -INF_F, -INF_F}}; }
inline AABB aunion(const AABB& a, const AABB& b) { return {minNat(a.mn, b.mn), maxNat(a.mx, b.mx)}; }
inline AABB aunion(const AABB& a, V3 p) { return {minNat(a.mn, p), maxNat(a.mx, p)}; }
inline V3 centroid(const AABB& a) { return 0.5f * (a.mn + a.mx); }
inline V3 diag(const AABB& a) { return a.mx - a.mn; }
inline float area(const AABB& a) { V3 d = diag(a); return 2.0f * (d.x * d.y + d.y * d.z + d.z * d.x); }
inline int splitAxisOf(const AABB& a) {
    V3 d = diag(a);
    if (d.x >= d.y && d.x >= d.z) return 0;
    return d.y >= d.z ? 1 : 2;
}
AABB aabbTransform(const AABB& a, const M4& m) {          // Util/BVH.fs:29-40
    AABB r = aabbDefault();
    for (int i = 0; i < 8; ++i) {
        V3 p = {(i & 1) == 0 ? a.mn.x : a.mx.x, (i & 2) == 0 ? a.mn.y : a.mx.y, (i & 4) == 0 ? a.mn.z : a.mx.z};
        V3 q = xform(p, m);
        r.mn = minNat(r.mn, q);
        r.mx = maxNat(r.mx, q);
    }
    return r;
}

struct Node {
    AABB b;
    int32_t rc;      // RightChild | InstanceOffset
    bool leaf;
    int axis;
    int count;
};
struct Item { AABB b; int idx; };

// Tree.Build (Util/BVH.fs:128-223) emitting nodes in the Flatten pre-order
// (:224-237): a node's index is taken before its left subtree, and the right
// child index is the node count when the right subtree starts.
void bvhBuild(std::vector<Item>& items, int first, int last, int depth, std::vector<Node>& nodes) {
    int n = last - first;
    std::vector<Item> sub(items.begin() + first, items.begin() + last);
    AABB bounds = aabbDefault();
    for (int i = 0; i < n; ++i) bounds = aunion(bounds, sub[i].b);
    int me = (int)nodes.size();
    nodes.push_back(Node{});
    if (n <= 4 || depth >= 64) {
        if (n > 127) throw std::runtime_error("BVH leaf larger than the int8 count field");
        nodes[me] = Node{bounds, first, true, 0, n};
        return;
    }
    AABB cb = aabbDefault();
    for (int i = 0; i < n; ++i) cb = aunion(cb, centroid(sub[i].b));
    int axis = splitAxisOf(cb);
    float extent = comp(diag(cb), axis);
    if (extent == 0.0f) {
        int mid = first + n / 2;
        bvhBuild(items, first, mid, depth + 1, nodes);
        int rc = (int)nodes.size();
        bvhBuild(items, mid, last, depth + 1, nodes);
        nodes[me] = Node{bounds, rc, false, axis, 0};
        return;
    }
    const int NB = 12;
    AABB bb[NB];
    int bc[NB];
    for (int i = 0; i < NB; ++i) { bb[i] = aabbDefault(); bc[i] = 0; }
    float cmin = comp(cb.mn, axis);
    // Util/BVH.fs:163-174: a negative index (a NaN centroid can leave cmin
    // above other centroids) makes the Span indexer throw.
    auto binOf = [&](const Item& it) {
        int b = std::min(NB - 1, satInt((12.0f * (comp(centroid(it.b), axis) - cmin)) / extent));
        if (b < 0) throw std::runtime_error("BVH build: bin index out of range (non-finite bounds)");
        return b;
    };
    for (int i = first; i < last; ++i) {
        const Item& it = sub[i - first];
        int b = binOf(it);
        bb[b] = aunion(bb[b], it.b);
        bc[b] = bc[b] + 1;
    }
    float costs[NB - 1];
    for (float& c : costs) c = 0.0f;
    AABB lb = aabbDefault(), rb = aabbDefault();
    int lc = 0, rcnt = 0;
    for (int i = 0; i < NB - 1; ++i) {
        lb = aunion(lb, bb[i]);
        lc = lc + bc[i];
        costs[i] = costs[i] + (float)lc * area(lb);
        rb = aunion(rb, bb[NB - 1 - i]);
        rcnt = rcnt + bc[NB - 1 - i];
        costs[NB - 2 - i] = costs[NB - 2 - i] + (float)rcnt * area(rb);
    }
    float minCost = INF_F;
    int best = 0;
    for (int i = 0; i < NB - 1; ++i) {
        float cost = costs[i] + (float)n * area(cb);
        if (cost < minCost) { minCost = cost; best = i; }
    }
    int left = first, right = last - 1;
    for (int i = 0; i < n; ++i) {
        const Item& it = sub[i];
        if (binOf(it) > best) items[right--] = it;
        else items[left++] = it;
    }
    bvhBuild(items, first, left, depth + 1, nodes);
    int rc = (int)nodes.size();
    bvhBuild(items, right + 1, last, depth + 1, nodes);
    nodes[me] = Node{bounds, rc, false, axis, 0};
}
// BVHNode.Build (:239-247): returns nodes and perm[i] = original index at i.
std::vector<Node> bvhBuildAll(const std::vector<AABB>& bounds, std::vector<int>& perm) {
    std::vector<Item> items(bounds.size());
    for (size_t i = 0; i < bounds.size(); ++i) items[i] = {bounds[i], (int)i};
    std::vector<Node> nodes;
    bvhBuild(items, 0, (int)items.size(), 0, nodes);
    perm.resize(items.size());
    for (size_t i = 0; i < items.size(); ++i) perm[i] = items[i].idx;
    return nodes;
}

// ----------------------------------------------------------------- Ray.fs
struct Ray { V3 o, d; };
inline V3 pointAt(const Ray& r, float t) { return fma3({t, t, t}, r.d, r.o); }
inline Ray rayXform(const Ray& r, const M4& m) { return {xform(r.o, m), xformDir(r.d, m)}; }
// AxisAlignedBoundingBox.Intersect (Base/Ray.fs:29-39)
inline bool slab(const AABB& b, const Ray& r, float t) {
    V3 inv = {1.0f / r.d.x, 1.0f / r.d.y, 1.0f / r.d.z};
    V3 t0 = (b.mn - r.o) * inv;
    V3 t1 = (b.mx - r.o) * inv;
    V3 a = minNat(t0, t1), c = maxNat(t0, t1);
    float tMin = maxN(1e-3f, maxN(a.x, maxN(a.y, a.z)));
    float tMax = minN(t, minN(c.x, minN(c.y, c.z)));
    return tMin <= tMax;
}

// ----------------------------------------------------------- Primitive.fs
struct ONB { V3 n, t, b; };
inline ONB onbFromN(V3 n) {                               // Primitive.fs:15-23
    V3 axis = std::fabs(n.x) > 0.1f ? V3{0, 1, 0} : V3{1, 0, 0};
    V3 t = normalize(cross(n, axis));
    return {n, t, cross(n, t)};
}
inline V3 toLocal(const ONB& o, V3 v) { return {dot(v, o.t), dot(v, o.b), dot(v, o.n)}; }
inline V3 toWorld(const ONB& o, V3 v) { return v.x * o.t + v.y * o.b + v.z * o.n; }
inline ONB onbXform(const ONB& o, const M4& m) {          // Primitive.fs:34-38
    V3 t = normalize(xformDir(o.t, m));
    V3 b = normalize(xformDir(o.b, m));
    return {cross(t, b), t, b};
}

struct Tri { V3 p0, p1, p2; };
inline AABB triBounds(const Tri& t) { return {minNat(minNat(t.p0, t.p1), t.p2), maxNat(maxNat(t.p0, t.p1), t.p2)}; }
inline float triArea(const Tri& t) { return 0.5f * length(cross(t.p1 - t.p0, t.p2 - t.p0)); }

struct Mesh {
    std::vector<V3> v;
    std::vector<int> tri;        // 3 per triangle, BVH order
    std::vector<Node> nodes;
    Alias alias;
    bool sphere = false;
    float radius = 1.0f;
    Tri get(int i) const { return {v[tri[3 * i]], v[tri[3 * i + 1]], v[tri[3 * i + 2]]}; }
    AABB bounds() const {
        if (sphere) return {{-radius, -radius, -radius}, {radius, radius, radius}};
        return nodes[0].b;
    }
};

struct Inst {
    int prim, material, light;
    M4 o2w, w2o;
    AABB bounds;
};

struct Material { int kind; V3 albedo; float ior, alpha, metallic, roughness; };
struct LightD { V3 emission; bool twoSided; };

struct Geom { V3 p; ONB onb; };
struct Hit { int inst; int prim; Geom g; };

}  // namespace

struct orc_scene {
    std::vector<Mesh> meshes;     // per primitive
    std::vector<Inst> insts;      // TLAS order
    std::vector<int> instDesc;    // TLAS order -> desc instance index
    std::vector<Node> tlas;
    std::vector<int> lights;      // TLAS-order indices of light instances
    std::vector<Material> mats;
    std::vector<LightD> lightDefs;
    M4 c2w;
    orc_camera cam;
};

namespace {

// ---------------------------------------------------------------- Mesh.fs
// Triangle.Intersect, closest-hit form (Mesh.fs:50-82).
inline bool triClosest(const Tri& tr, const Ray& r, Geom& g, float& t) {
    const float eps = std::numeric_limits<float>::denorm_min();  // Single.Epsilon
    V3 e0 = tr.p1 - tr.p0, e1 = tr.p2 - tr.p0;
    V3 q = cross(r.d, e1);
    float det = dot(e0, q);
    if (std::fabs(det) < eps) return false;
    float inv = 1.0f / det;
    V3 s = r.o - tr.p0;
    float u = inv * dot(s, q);
    if (u < 0.0f || u > 1.0f) return false;
    V3 sc = cross(s, e0);
    float v = inv * dot(r.d, sc);
    if (v < 0.0f || u + v > 1.0f) return false;
    float tt = inv * dot(e1, sc);
    if (tt > eps && tt < t) {
        V3 n = normalize(cross(e0, e1));
        g = {pointAt(r, tt), onbFromN(n)};
        t = tt;
        return true;
    }
    return false;
}
// Triangle.Intersect, any-hit form (Mesh.fs:24-48).
inline bool triAny(const Tri& tr, const Ray& r, float t) {
    const float eps = std::numeric_limits<float>::denorm_min();
    V3 e0 = tr.p1 - tr.p0, e1 = tr.p2 - tr.p0;
    V3 q = cross(r.d, e1);
    float det = dot(e0, q);
    if (std::fabs(det) < eps) return false;
    float inv = 1.0f / det;
    V3 s = r.o - tr.p0;
    float u = inv * dot(s, q);
    if (u < 0.0f || u > 1.0f) return false;
    V3 sc = cross(s, e0);
    float v = inv * dot(r.d, sc);
    if (v < 0.0f || u + v > 1.0f) return false;
    float tt = inv * dot(e1, sc);
    return tt > eps && tt < t;
}

struct Counters { int64_t nodes = 0, tris = 0, insts = 0; };

// MeshPrimitive.Intersect closest (Mesh.fs:217-242).
bool meshClosest(const Mesh& m, const Ray& r, Geom& g, float& t, int& tag, Counters& c) {
    int stack[160];
    int sp = 0;
    stack[sp++] = 0;
    bool hit = false;
    while (sp != 0) {
        int i = stack[--sp];
        const Node& nd = m.nodes[i];
        c.nodes++;
        if (slab(nd.b, r, t)) {
            if (nd.leaf) {
                for (int k = nd.rc; k < nd.rc + nd.count; ++k) {
                    Tri tr = m.get(k);
                    c.tris++;
                    if (slab(triBounds(tr), r, t) && triClosest(tr, r, g, t)) { tag = k; hit = true; }
                }
            } else if (comp(r.d, nd.axis) > 0.0f) {
                stack[sp++] = nd.rc; stack[sp++] = i + 1;
            } else {
                stack[sp++] = i + 1; stack[sp++] = nd.rc;
            }
        }
    }
    return hit;
}
// MeshPrimitive.Intersect any (Mesh.fs:188-215).  The reference starts with
// stackTop = 1 over a zeroed slot 0, so a miss re-traverses from the root; the
// boolean result is the same without that second pass, which is omitted.
bool meshAny(const Mesh& m, const Ray& r, float t, Counters& c) {
    int stack[160];
    int sp = 0;
    stack[sp++] = 0;
    bool hit = false;
    while (!hit && sp != 0) {
        int i = stack[--sp];
        const Node& nd = m.nodes[i];
        c.nodes++;
        if (slab(nd.b, r, t)) {
            if (nd.leaf) {
                int k = nd.rc;
                while (!hit && k < nd.rc + nd.count) {
                    Tri tr = m.get(k);
                    c.tris++;
                    if (slab(triBounds(tr), r, t)) hit = triAny(tr, r, t);
                    ++k;
                }
            } else if (comp(r.d, nd.axis) > 0.0f) {
                stack[sp++] = nd.rc; stack[sp++] = i + 1;
            } else {
                stack[sp++] = i + 1; stack[sp++] = nd.rc;
            }
        }
    }
    return hit;
}

// -------------------------------------------------------------- Sphere.fs
bool sphereClosest(float radius, const Ray& r, Geom& g, float& t, int& branch) {  // Sphere.fs:35-77
    const float eps = 1e-3f;
    V3 f = r.o;
    float a = dot(r.d, r.d);
    float b = -dot(f, r.d);
    float r2 = radius * radius;
    float c = dot(f, f) - r2;
    V3 fb = f + (b / a) * r.d;
    float d = r2 - dot(fb, fb);
    if (d < 0.0f) return false;
    float q = b + std::copysign(std::sqrt(a * d), b);
    float t0 = c / q;
    if (t0 > eps && t0 < t) {
        t = t0;
        V3 p = pointAt(r, t);
        V3 n = normalize(p);
        if (dot(n, r.d) > 0.0f) n = -n;
        g = {p, onbFromN(n)};
        branch = 0;
        return true;
    }
    float t1 = q / a;
    if (t1 > eps && t1 < t) {
        t = t1;
        V3 p = pointAt(r, t);
        g = {p, onbFromN(normalize(p))};
        branch = 1;
        return true;
    }
    return false;
}
bool sphereAny(float radius, const Ray& r, float t) {    // Sphere.fs:13-33
    const float eps = 1e-3f;
    V3 f = r.o;
    float a = dot(r.d, r.d);
    float b = -dot(f, r.d);
    float r2 = radius * radius;
    float c = dot(f, f) - r2;
    V3 fb = f + (b / a) * r.d;
    float d = r2 - dot(fb, fb);
    if (d < 0.0f) return false;
    float q = b + std::copysign(std::sqrt(a * d), b);
    float t0 = c / q;
    if (t0 > eps && t0 < t) return true;
    float t1 = q / a;
    return t1 > eps && t1 < t;
}

// ------------------------------------------------ aggregate (Aggregate/BVH.fs)
bool instClosest(const orc_scene& s, int ii, const Ray& r, Hit& h, float& t, Counters& c) {
    const Inst& in = s.insts[ii];                          // Primitive.fs:118-129
    if (!slab(in.bounds, r, t)) return false;
    c.insts++;
    Ray o = rayXform(r, in.w2o);
    const Mesh& m = s.meshes[in.prim];
    Geom g;
    int tag = 0;
    bool hit = m.sphere ? sphereClosest(m.radius, o, g, t, tag) : meshClosest(m, o, g, t, tag, c);
    if (!hit) return false;
    // LocalGeometry.Transform (Primitive.fs:57-58): the tag is not carried.
    h.g = {xform(g.p, in.o2w), onbXform(g.onb, in.o2w)};
    h.inst = ii;
    h.prim = tag;
    return true;
}
bool instAny(const orc_scene& s, int ii, const Ray& r, float t, Counters& c) {
    const Inst& in = s.insts[ii];                          // Primitive.fs:111-116
    if (!slab(in.bounds, r, t)) return false;
    c.insts++;
    Ray o = rayXform(r, in.w2o);
    const Mesh& m = s.meshes[in.prim];
    return m.sphere ? sphereAny(m.radius, o, t) : meshAny(m, o, t, c);
}
bool closest(const orc_scene& s, const Ray& r, Hit& h, float& t, Counters& c) {  // BVH.fs:37-58
    int stack[160];
    int sp = 0;
    stack[sp++] = 0;
    bool hit = false;
    while (sp != 0) {
        int i = stack[--sp];
        const Node& nd = s.tlas[i];
        c.nodes++;
        if (slab(nd.b, r, t)) {
            if (nd.leaf) {
                for (int k = nd.rc; k < nd.rc + nd.count; ++k) hit = instClosest(s, k, r, h, t, c) || hit;
            } else if (comp(r.d, nd.axis) > 0.0f) {
                stack[sp++] = nd.rc; stack[sp++] = i + 1;
            } else {
                stack[sp++] = i + 1; stack[sp++] = nd.rc;
            }
        }
    }
    return hit;
}
bool anyHit(const orc_scene& s, const Ray& r, float t, Counters& c) {  // BVH.fs:11-35
    int stack[160];
    int sp = 0;
    stack[sp++] = 0;
    bool hit = false;
    while (!hit && sp != 0) {
        int i = stack[--sp];
        const Node& nd = s.tlas[i];
        c.nodes++;
        if (slab(nd.b, r, t)) {
            if (nd.leaf) {
                int k = nd.rc;
                while (!hit && k < nd.rc + nd.count) { hit = instAny(s, k, r, t, c); ++k; }
            } else if (comp(r.d, nd.axis) > 0.0f) {
                stack[sp++] = nd.rc; stack[sp++] = i + 1;
            } else {
                stack[sp++] = i + 1; stack[sp++] = nd.rc;
            }
        }
    }
    return hit;
}

// ------------------------------------------------------------ materials
struct BSDFEval { V3 bsdf; float pdf; };
struct BSDFSample { BSDFEval e; V3 wi; };

float pbrLambda(const Material& m, V3 w) {                // PBR.fs:13-19
    float s2 = std::fma(w.x, w.x, w.y * w.y);
    if (s2 == 0.0f) return 0.0f;
    float t2 = s2 / (w.z * w.z);
    float a2t2 = m.alpha * m.alpha * t2;
    return (-1.0f + std::sqrt(1.0f + a2t2)) / 2.0f;
}
float pbrD(const Material& m, V3 wh) {                    // PBR.fs:20-23
    float c = std::fabs(wh.z);
    float x = 1.0f + std::fma(m.alpha, m.alpha, -1.0f) * c * c;
    return m.alpha * m.alpha / (PI_F * (x * x));
}
BSDFEval pbrEval(const Material& m, V3 wo, V3 wi) {       // PBR.fs:37-49
    V3 wh = normalize(wo + wi);
    float d = pbrD(m, wh);
    float g = 1.0f / (1.0f + pbrLambda(m, wo) + pbrLambda(m, wi));
    float spec = d * g / (4.0f * std::fabs(wo.z));
    float cosT = dot(wo, wh);
    float c = 1.0f - cosT;
    float c2 = c * c;
    V3 fc = m.albedo + (V3{1, 1, 1} - m.albedo) * (c2 * c2 * c);
    V3 metal = spec * fc;
    const float f0 = 0.04f;
    float f = f0 + (1.0f - f0) * c2 * c2 * c;
    V3 diffuse = m.albedo * std::fabs(wi.z) / PI_F;
    auto mix3 = [](V3 a, V3 b, float tt) { V3 bt = b * tt; float s = 1.0f - tt; return V3{std::fma(a.x, s, bt.x), std::fma(a.y, s, bt.y), std::fma(a.z, s, bt.z)}; };
    V3 bsdf = mix3(mix3(diffuse, V3{spec, spec, spec}, f), metal, m.metallic);
    float pa = d * std::fabs(wh.z) / (4.0f * dot(wo, wh));
    float pb = std::fabs(wi.z) / PI_F;
    float tt = 0.5f * (1.0f - m.metallic);
    float pdf = std::fma(pa, 1.0f - tt, pb * tt);
    return {bsdf, pdf};
}
BSDFEval matEval(const Material& m, V3 wo, V3 wi) {
    switch (m.kind) {
        case 0: {                                           // Lambertian.fs:11-16
            if (wi.z * wo.z < 0.0f || minN(std::fabs(wi.z), std::fabs(wo.z)) < 1e-6f) return {{0, 0, 0}, 0.0f};
            float pdf = std::fabs(wi.z) / PI_F;
            return {m.albedo * pdf, pdf};
        }
        case 3: return pbrEval(m, wo, wi);
        default: return {{0, 0, 0}, 0.0f};                  // Mirror.fs:11-12, Dielectric.fs:13-14
    }
}
BSDFSample matSample(const Material& m, V3 wo, float uLobe, V2 u) {
    switch (m.kind) {
        case 0: {                                           // Lambertian.fs:18-26
            float cosT = std::sqrt(u.x);
            float sinT = std::sqrt(1.0f - u.x);
            float ph = 2.0f * PI_F * u.y;
            float sp = sinF(ph), cp = cosF(ph);
            V3 wi = {sinT * cp, sinT * sp, cosT};
            float pdf = wi.z / PI_F;
            return {{m.albedo * pdf, pdf}, wo.z > 0.0f ? wi : -wi};
        }
        case 1:                                             // Mirror.fs:13-14
            return {{m.albedo, 1.0f}, {-wo.x, -wo.y, wo.z}};
        case 2: {                                           // Dielectric.fs:15-31
            float ior = wo.z > 0.0f ? 1.0f / m.ior : m.ior;
            float c2t = 1.0f - ior * ior * std::fma(-wo.z, wo.z, 1.0f);
            if (c2t <= 0.0f) return {{m.albedo, 1.0f}, {-wo.x, -wo.y, wo.z}};
            float a = ior - 1.0f, b = ior + 1.0f;
            float r0 = a * a / (b * b);
            float c = 1.0f - std::fabs(wo.z);
            float c2 = c * c;
            float r = r0 + (1.0f - r0) * c2 * c2 * c;
            if (uLobe < r) return {{r * m.albedo, r}, {-wo.x, -wo.y, wo.z}};
            float cosT = std::sqrt(c2t);
            return {{(1.0f - r) * m.albedo, 1.0f - r}, {-wo.x * ior, -wo.y * ior, -std::copysign(cosT, wo.z)}};
        }
        default: {                                          // PBR.fs:50-64
            V3 wi;
            if (uLobe < 1.0f - 0.5f * (1.0f - m.metallic)) {
                float th = atanF(m.alpha * std::sqrt(u.x / (1.0f - u.x)));
                float ph = 2.0f * PI_F * u.y;
                float st = sinF(th), ct = cosF(th), sp = sinF(ph), cp = cosF(ph);
                V3 wh = {st * cp, st * sp, ct};
                wi = 2.0f * dot(wo, wh) * wh - wo;
            } else {
                float cosT = std::sqrt(u.x);
                float sinT = std::sqrt(1.0f - u.x);
                float ph = 2.0f * PI_F * u.y;
                float sp = sinF(ph), cp = cosF(ph);
                wi = {sinT * cp, sinT * sp, cosT};
            }
            return {pbrEval(m, wo, wi), wi};
        }
    }
}
Material toMaterial(const orc_material& d) {
    Material m;
    m.kind = d.kind;
    m.albedo = {d.albedo[0], d.albedo[1], d.albedo[2]};
    m.ior = d.ior;
    m.alpha = maxN(d.roughness * d.roughness, 1e-3f);      // PBR.fs:11
    m.roughness = d.roughness;
    float mt = d.metallic;                                  // Single.Clamp (PBR.fs:12)
    m.metallic = mt < 0.0f ? 0.0f : (mt > 1.0f ? 1.0f : mt);
    return m;
}

// ---------------------------------------------------------------- lights
inline V3 lightEmit(const LightD& l, float woz) {         // Light.fs:49-52
    if (std::fabs(woz) > 1e-6f && (woz > 0.0f || l.twoSided)) return l.emission;
    return {0, 0, 0};
}
struct LightSample { V3 p, L; float pdf; V3 wi; int ldef; float facing; };
// MeshInstance.Sample (Mesh.fs:289-298) / SphereInstance.Sample (Sphere.fs:90-113)
float instSample(const orc_scene& s, const Inst& in, float uSel, V2 u, V3& p, V3& n) {
    const Mesh& m = s.meshes[in.prim];
    if (m.sphere) {
        float a = 2.0f * PI_F * u.x;
        float st = sinF(a), ct = cosF(a);
        float cp = std::fma(-2.0f, u.y, 1.0f);
        float sp = std::sqrt(std::fma(-cp, cp, 1.0f));
        V3 nn = {ct * sp, st * sp, cp};
        ONB o = onbFromN(nn);
        V3 pl = nn * m.radius;
        V3 pw = xform(pl, in.o2w);
        V3 tw = xformDir(o.t, in.o2w), bw = xformDir(o.b, in.o2w);
        V3 nw = cross(tw, bw);
        float invJ = 1.0f / length(nw);
        p = pw;
        n = invJ * nw;
        return invJ / (4.0f * PI_F * m.radius * m.radius);
    }
    int i;
    float pdfTri;
    alias_sample(m.alias.alias.data(), m.alias.prob.data(), m.alias.pdf.data(), (int)m.alias.pdf.size(), uSel, i, pdfTri);
    Tri tr = m.get(i);
    tr = {xform(tr.p0, in.o2w), xform(tr.p1, in.o2w), xform(tr.p2, in.o2w)};
    V2 uv = u.x < u.y ? V2{0.5f * u.x, std::fma(-0.5f, u.x, u.y)} : V2{std::fma(-0.5f, u.y, u.x), 0.5f * u.y};  // Mesh.fs:89-99
    p = uv.x * tr.p1 + uv.y * tr.p2 + (1.0f - uv.x - uv.y) * tr.p0;
    V3 nc = cross(tr.p1 - tr.p0, tr.p2 - tr.p0);
    float pdfArea = 2.0f / length(nc);
    n = (0.5f * pdfArea) * nc;
    return pdfTri * pdfArea;
}
// UniformLightSampler.Sample (Uniform.fs:13-29)
LightSample lightSample(const orc_scene& s, V3 p, float uSel, V2 u) {
    int N = (int)s.lights.size();
    float us = uSel * (float)N;
    int id = std::min(satInt(us), N - 1);
    us = us - (float)id;
    const Inst& in = s.insts[s.lights[id]];
    V3 lp, ln;
    float pdfS = instSample(s, in, us, u, lp, ln);
    V3 wo = normalize(p - lp);
    float cosWo = dot(ln, wo);
    V3 dv = p - lp;
    float dist2 = dot(dv, dv);
    LightSample r;
    r.p = lp;
    const LightD& ld = s.lightDefs[in.light];
    const float woz = dot(wo, ln);
    r.L = lightEmit(ld, woz);
    r.pdf = dist2 * pdfS / (maxN(std::fabs(cosWo), 1e-6f) * (float)N);
    r.wi = -wo;
    r.ldef = in.light;
    r.facing = std::fabs(woz) > 1e-6f && (woz > 0.0f || ld.twoSided) ? 1.0f : 0.0f;
    return r;
}
// PrimitiveInstance.EvalPDF (Mesh.fs:300-304 with the tag lost; Sphere.fs:115-126)
float instEvalPdf(const orc_scene& s, const Inst& in, const Hit& h) {
    const Mesh& m = s.meshes[in.prim];
    if (m.sphere) {
        V3 tw = xformDir(h.g.onb.t, in.w2o), bw = xformDir(h.g.onb.b, in.w2o);
        float invJ = length(cross(tw, bw));
        return invJ / (4.0f * PI_F * m.radius * m.radius);
    }
    Tri tr = m.get(0);
    tr = {xform(tr.p0, in.o2w), xform(tr.p1, in.o2w), xform(tr.p2, in.o2w)};
    return m.alias.pdf[0] / triArea(tr);
}
struct LightEval { V3 L; float pdf; };
LightEval lightEval(const orc_scene& s, V3 p, const Hit& h) {  // Uniform.fs:40-49
    const Inst& in = s.insts[h.inst];
    V3 wo = normalize(p - h.g.p);
    float cosWo = dot(h.g.onb.n, wo);
    float pdfS = instEvalPdf(s, in, h);
    V3 dv = p - h.g.p;
    float dist2 = dot(dv, dv);
    return {lightEmit(s.lightDefs[in.light], toLocal(h.g.onb, wo).z),
            dist2 * pdfS / (maxN(std::fabs(cosWo), 1e-6f) * (float)s.lights.size())};
}

// ---------------------------------------------------------------- camera
Ray generatePrimary(const orc_scene& s, int W, int H, int x, int y, V2 uPix, V2 uLens) {
    const orc_camera& c = s.cam;                           // Pinhole.fs:12-27
    float vh = 2.0f * tanF(c.fov * PI_F / 360.0f);
    float vw = vh * c.aspect;
    V3 U = vw * V3{1, 0, 0}, V = vh * V3{0, 1, 0};
    V3 dU = U / (float)W, dV = V / (float)H;
    V3 ul = -V3{0, 0, 1} - 0.5f * (U + V);
    V3 pl = ul + ((float)x + uPix.x) * dU + ((float)y + uPix.y) * dV;
    Ray r = {{0, 0, 0}, normalize(pl)};
    if (c.kind == 1 && c.aperture > 0.0f) {                // ThinLens.fs:12-35
        V2 u = {uLens.x * 2.0f - 1.0f, uLens.y * 2.0f - 1.0f};
        V2 d = {0, 0};
        if (u.x == 0.0f || u.y == 0.0f) d = {0, 0};
        else if (std::fabs(u.x) > std::fabs(u.y)) {
            float rr = u.x, th = PI_F / 4.0f * (u.y / u.x);
            d = {rr * cosF(th), rr * sinF(th)};
        } else {
            float rr = u.y, th = PI_F / 2.0f - PI_F / 4.0f * (u.x / u.y);
            d = {rr * cosF(th), rr * sinF(th)};
        }
        V2 pLens = {c.aperture * d.x, c.aperture * d.y};
        V3 o = r.o + V3{pLens.x, pLens.y, 0.0f};
        V3 dir = normalize(pointAt(r, c.focus_distance) - o);
        r = {o, dir};
    }
    // CameraBase.GeneratePrimaryRay (Base/Camera.fs:12-20)
    Ray w;
    w.o = xform(r.o, s.c2w);
    w.d = normalize(xformDir(r.d, s.c2w));
    w.o = pointAt(w, c.push_forward);
    return w;
}

// ------------------------------------------------------------ integrators
inline void evalBSDF(const orc_scene& s, const Hit& h, V3 wo, V3 wi, BSDFEval& e) {
    e = matEval(s.mats[s.insts[h.inst].material], toLocal(h.g.onb, wo), toLocal(h.g.onb, wi));
}
inline BSDFSample sampleBSDF(const orc_scene& s, const Hit& h, V3 wo, float ul, V2 u) {
    BSDFSample b = matSample(s.mats[s.insts[h.inst].material], toLocal(h.g.onb, wo), ul, u);
    b.wi = toWorld(h.g.onb, b.wi);
    return b;
}

struct Stats {
    Counters cc, cs;
    int64_t raysC = 0, raysS = 0, verts = 0, paths = 0;
    void add(const Stats& o) {
        cc.nodes += o.cc.nodes; cc.tris += o.cc.tris; cc.insts += o.cc.insts;
        cs.nodes += o.cs.nodes; cs.tris += o.cs.tris; cs.insts += o.cs.insts;
        raysC += o.raysC; raysS += o.raysS; verts += o.verts; paths += o.paths;
    }
};

// PathTracingIntegrator.Li (PathTracing.fs:14-81)
V3 liPT(const orc_scene& s, Ray ray, Sampler& smp, int maxDepth, int rrDepth, Stats& st) {
    V3 L = {0, 0, 0}, beta = {1, 1, 1};
    int depth = 0;
    float prevPdf = 0.0f;                                   // bsdfSample.eval.pdf
    while (depth < maxDepth) {
        float t = INF_F;
        Hit h;
        st.raysC++;
        if (!closest(s, ray, h, t, st.cc)) { depth = maxDepth; continue; }
        st.verts++;
        const Inst& in = s.insts[h.inst];
        if (in.light >= 0) {
            LightEval le = lightEval(s, ray.o, h);
            float w = depth == 0 ? 1.0f : prevPdf * rcpE(le.pdf + prevPdf);
            L = fma3(beta, le.L * w, L);
        }
        if (in.material >= 0) {
            float uSel = smp.next();
            V2 uL = smp.next2();
            LightSample ls = lightSample(s, h.g.p, uSel, uL);
            Ray sh = {h.g.p, ls.wi};
            float dist = length(ls.p - h.g.p);
            if (ls.pdf != 0.0f) {
                st.raysS++;
                if (!anyHit(s, sh, dist - 1e-3f, st.cs)) {
                    BSDFEval e;
                    evalBSDF(s, h, -ray.d, ls.wi, e);
                    L = fma3(beta * e.bsdf, ls.L * rcpE(e.pdf + ls.pdf), L);
                }
            }
            float ul = smp.next();
            V2 ub = smp.next2();
            BSDFSample bs = sampleBSDF(s, h, -ray.d, ul, ub);
            prevPdf = bs.e.pdf;
            if (bs.e.pdf == 0.0f) { depth = maxDepth; continue; }
            ray = {h.g.p, bs.wi};
            beta = beta * bs.e.bsdf * rcpE(bs.e.pdf);
            if (depth >= rrDepth) {
                float q = minN(1.0f, maxN(beta.x, maxN(beta.y, beta.z)));
                if (smp.next() < q) beta = beta * rcpE(q);
                else depth = maxDepth;
            }
            depth = depth + 1;
        } else {
            depth = maxDepth;
        }
    }
    return L;
}
// NormalIntegrator.Li (Normal.fs:10-17)
V3 liNormal(const orc_scene& s, const Ray& ray, Stats& st) {
    float t = INF_F;
    Hit h;
    st.raysC++;
    if (closest(s, ray, h, t, st.cc)) return 0.5f * (h.g.onb.n + V3{1, 1, 1});
    return {0, 0, 0};
}
// DirectIntegrator.Li (Direct.fs:10-40)
V3 liDirect(const orc_scene& s, const Ray& ray, Sampler& smp, Stats& st) {
    float t = INF_F;
    Hit h;
    V3 L = {0, 0, 0};
    st.raysC++;
    if (closest(s, ray, h, t, st.cc)) {
        const Inst& in = s.insts[h.inst];
        if (in.light >= 0) L = L + lightEmit(s.lightDefs[in.light], toLocal(h.g.onb, -ray.d).z);
        if (in.material >= 0) {
            float uSel = smp.next();
            V2 uL = smp.next2();
            LightSample ls = lightSample(s, h.g.p, uSel, uL);
            Ray sh = {h.g.p, ls.wi};
            float dist = length(ls.p - h.g.p);
            st.raysS++;
            if (!anyHit(s, sh, dist - 1e-3f, st.cs)) {
                BSDFEval e;
                evalBSDF(s, h, -ray.d, ls.wi, e);
                L = fma3(e.bsdf, ls.L * (1.0f / ls.pdf), L);
            }
            float ul = smp.next();
            V2 ub = smp.next2();
            BSDFSample bs = sampleBSDF(s, h, -ray.d, ul, ub);
            Ray r2 = {h.g.p, bs.wi};
            Hit h2;
            float t2 = INF_F;
            st.raysC++;
            if (closest(s, r2, h2, t2, st.cc) && s.insts[h2.inst].light >= 0) {
                LightEval le = lightEval(s, r2.o, h2);
                L = fma3(bs.e.bsdf, le.L * (1.0f / le.pdf), L);
            }
        }
    }
    return L;
}

// ProgressiveIntegrator.RenderTile (Base/Integrator.fs:22-44) for one pixel.
V3 renderPixel(const orc_scene& s, int integ, const orc_pt_params& p, int x, int y, Stats& st) {
    float invSpp = rcpE((float)p.spp);
    V3 acc = {0, 0, 0};
    for (int k = 0; k < p.spp; ++k) {
        Sampler smp{xx3((uint32_t)x, (uint32_t)y, (uint32_t)(p.frame_id * p.spp + k))};
        V2 uPix = smp.next2();
        V2 uLens = smp.next2();
        Ray r = generatePrimary(s, p.width, p.height, x, y, uPix, uLens);
        st.paths++;
        V3 Li = integ == 0 ? liNormal(s, r, st) : integ == 1 ? liDirect(s, r, smp, st)
                                                             : liPT(s, r, smp, p.max_depth, p.rr_depth, st);
        V3 rad = Li * rcpE(1.0f);                            // camera pdf is 1
        acc = fma3({invSpp, invSpp, invSpp}, rad, acc);
    }
    return acc;
}

// ---------------------------------------------------------------- PSSMLT
struct PrimarySample { float v, vb; int lm, mb; };         // PSSMLT.fs:20-33
inline float erfInv(float x) {                            // PSSMLT.fs:69-94 (Giles 2010)
    x = minN(maxN(x, -0.99999f), 0.99999f);
    float w = -logF(std::fma(x, -x, 1.0f));
    float p;
    if (w < 5.0f) {
        w = w - 2.5f;
        p = 2.81022636e-08f;
        p = std::fma(p, w, 3.43273939e-07f);
        p = std::fma(p, w, -3.5233877e-06f);
        p = std::fma(p, w, -4.39150654e-06f);
        p = std::fma(p, w, 0.00021858087f);
        p = std::fma(p, w, -0.00125372503f);
        p = std::fma(p, w, -0.00417768164f);
        p = std::fma(p, w, 0.246640727f);
        return std::fma(p, w, 1.50140941f) * x;
    }
    w = std::sqrt(w) - 3.0f;
    p = -0.000200214257f;
    p = std::fma(p, w, 0.000100950558f);
    p = std::fma(p, w, 0.00134934322f);
    p = std::fma(p, w, -0.00367342844f);
    p = std::fma(p, w, 0.00573950773f);
    p = std::fma(p, w, -0.0076224613f);
    p = std::fma(p, w, 0.00943887047f);
    p = std::fma(p, w, 1.00167406f);
    return std::fma(p, w, 2.83297682f) * x;
}
struct MLTSampler {                                       // PSSMLT.fs:35-150
    Sampler inner;
    float pL;
    int strategy;
    float sigma, epsMin, epsMax;
    PrimarySample* xs;
    bool large = false;
    int lastLarge = 0, cur = 0, idx = 0, init = 0;
    void start() {
        large = cur == 0 || inner.next() < pL;
        cur = cur + 1;
        idx = 0;
    }
    float ensure(int i) {
        PrimarySample& x = xs[i];
        if (init <= i) { x = {0.0f, 0.0f, 0, 0}; init = i + 1; }
        if (x.lm < lastLarge) { x.v = inner.next(); x.lm = lastLarge; }
        x.vb = x.v; x.mb = x.lm;
        float val;
        if (large) {
            val = inner.next();
        } else {
            if (strategy == 0) {
                float ns = std::sqrt(2.0f) * erfInv(std::fma(2.0f, inner.next(), -1.0f));
                float es = sigma * std::sqrt((float)(cur - x.lm));
                val = std::fma(ns, es, x.v);
            } else {
                float value = x.v;
                float a = logF(epsMax / epsMin);
                for (int k = x.lm; k <= cur - 1; ++k) {
                    float u1 = inner.next() - 0.5f;
                    float u2 = u1 < 0.0f ? 1.0f + 2.0f * u1 : 2.0f * u1;
                    value = value + std::copysign(epsMax * expF(-a * u2), u1);
                }
                val = value;
            }
            val = val - std::floor(val);
        }
        x.v = val;
        x.lm = cur;
        return x.v;
    }
    float next() { return ensure(idx++); }
    V2 next2() { float a = next(); float b = next(); return {a, b}; }
    void reject() {
        for (int i = 0; i < init; ++i) { xs[i].v = xs[i].vb; xs[i].lm = xs[i].mb; }
        cur = cur - 1;
    }
    void accept() { if (large) lastLarge = cur; }
};
// PSSMLTIntegrator.Li (PSSMLT.fs:172-245): seven primary samples per hit.
V3 liMLT(const orc_scene& s, Ray ray, MLTSampler& smp, int maxDepth, int rrDepth, Stats& st) {
    V3 L = {0, 0, 0}, beta = {1, 1, 1};
    int depth = 0;
    float prevPdf = 0.0f;
    while (depth < maxDepth) {
        float t = INF_F;
        Hit h;
        st.raysC++;
        if (!closest(s, ray, h, t, st.cc)) { depth = maxDepth; continue; }
        st.verts++;
        const Inst& in = s.insts[h.inst];
        if (in.light >= 0) {
            LightEval le = lightEval(s, ray.o, h);
            float w = depth == 0 ? 1.0f : prevPdf * rcpE(le.pdf + prevPdf);
            L = fma3(beta, le.L * w, L);
        }
        float uLight = smp.next();
        V2 uEmit = smp.next2();
        float uLobe = smp.next();
        V2 uBSDF = smp.next2();
        float uRR = smp.next();
        if (in.material >= 0) {
            LightSample ls = lightSample(s, h.g.p, uLight, uEmit);
            Ray sh = {h.g.p, ls.wi};
            float dist = length(ls.p - h.g.p);
            if (ls.pdf != 0.0f) {
                st.raysS++;
                if (!anyHit(s, sh, dist - 1e-3f, st.cs)) {
                    BSDFEval e;
                    evalBSDF(s, h, -ray.d, ls.wi, e);
                    L = fma3(beta * e.bsdf, ls.L * rcpE(e.pdf + ls.pdf), L);
                }
            }
            BSDFSample bs = sampleBSDF(s, h, -ray.d, uLobe, uBSDF);
            prevPdf = bs.e.pdf;
            if (bs.e.pdf == 0.0f) { depth = maxDepth; continue; }
            ray = {h.g.p, bs.wi};
            beta = beta * bs.e.bsdf * rcpE(bs.e.pdf);
            if (depth >= rrDepth) {
                float q = minN(1.0f, maxN(beta.x, maxN(beta.y, beta.z)));
                if (uRR < q) beta = beta * rcpE(q);
                else depth = maxDepth;
            }
            depth = depth + 1;
        } else {
            depth = maxDepth;
        }
    }
    return L;
}
inline float luminance(V3 L) { return dot(L, V3{0.2126f, 0.7152f, 0.0722f}); }

struct MLTPath { V3 L; float y; int px, py; };
MLTPath mltPath(const orc_scene& s, const orc_mlt_params& p, MLTSampler& m, Stats& st) {
    // PSSMLT.fs:254-273 / 292-310 / 333-351
    m.start();
    V2 u = m.next2();
    V2 up = {u.x * (float)p.width, u.y * (float)p.height};
    int ix = std::min(p.width - 1, satInt(up.x));
    int iy = std::min(p.height - 1, satInt(up.y));
    V2 frac = {up.x - (float)ix, up.y - (float)iy};
    V2 uLens = m.next2();
    Ray r = generatePrimary(s, p.width, p.height, ix, iy, frac, uLens);
    st.paths++;
    V3 L = liMLT(s, r, m, p.max_depth, p.rr_depth, st) * rcpE(1.0f);
    return {L, luminance(L), ix, iy};
}

// ------------------------------------------------------------ key frames
// KeyFrameInfo.ToKeyFrame (Loader.fs:20-41): S * Rx * Ry * Rz * T.
M4 keyFrameMatrix(const orc_transform& t) {
    M4 S = (t.flags & 1) ? scaleM(t.scale[0], t.scale[1], t.scale[2]) : identity();
    M4 R = (t.flags & 2) ? mul(mul(rotX(t.rotation[0]), rotY(t.rotation[1])), rotZ(t.rotation[2])) : identity();
    M4 T = (t.flags & 4) ? translateM(t.translation[0], t.translation[1], t.translation[2]) : identity();
    return mul(mul(S, R), T);
}
struct Q4 { float x, y, z, w; };
// Quaternion.CreateFromRotationMatrix (System.Numerics): the largest of the
// trace and the diagonal picks the component computed from a square root.
Q4 rotToQuat(const float* m) {
    const float tr = m[0] + m[5] + m[10];
    if (tr > 0.0f) {
        float s = std::sqrt(tr + 1.0f);
        float w = s * 0.5f;
        s = 0.5f / s;
        return {(m[6] - m[9]) * s, (m[8] - m[2]) * s, (m[1] - m[4]) * s, w};
    }
    if (m[0] >= m[5] && m[0] >= m[10]) {
        float s = std::sqrt(1.0f + m[0] - m[5] - m[10]);
        float k = 0.5f / s;
        return {0.5f * s, (m[1] + m[4]) * k, (m[2] + m[8]) * k, (m[6] - m[9]) * k};
    }
    if (m[5] > m[10]) {
        float s = std::sqrt(1.0f + m[5] - m[0] - m[10]);
        float k = 0.5f / s;
        return {(m[4] + m[1]) * k, 0.5f * s, (m[9] + m[6]) * k, (m[8] - m[2]) * k};
    }
    float s = std::sqrt(1.0f + m[10] - m[0] - m[5]);
    float k = 0.5f / s;
    return {(m[8] + m[2]) * k, (m[9] + m[6]) * k, 0.5f * s, (m[1] - m[4]) * k};
}
// Matrix4x4.GetDeterminant of a matrix whose fourth row and column are those
// of the identity (the only kind Decompose builds): its cofactor expansion
// reduces to the 3 x 3 one below (the dropped terms add zeros).
float det3of4(const float* q) {
    return q[0] * (q[5] * q[10] - q[6] * q[9]) - q[1] * (q[4] * q[10] - q[6] * q[8]) +
           q[2] * (q[4] * q[9] - q[5] * q[8]);
}
// Matrix4x4.Decompose (System.Numerics).
void decomposeM(const M4& M, V3& sc, Q4& q, V3& tr) {
    const float EPS = 1e-4f;
    tr = {M.m[12], M.m[13], M.m[14]};
    V3 r[3];
    for (int i = 0; i < 3; ++i) r[i] = {M.m[4 * i], M.m[4 * i + 1], M.m[4 * i + 2]};
    float s[3] = {length(r[0]), length(r[1]), length(r[2])};
    int a, b, c;                                               // ranking, largest first
    if (s[0] < s[1]) {
        if (s[1] < s[2]) { a = 2; b = 1; c = 0; } else if (s[0] < s[2]) { a = 1; b = 2; c = 0; } else { a = 1; b = 0; c = 2; }
    } else {
        if (s[0] < s[2]) { a = 2; b = 0; c = 1; } else if (s[1] < s[2]) { a = 0; b = 2; c = 1; } else { a = 0; b = 1; c = 2; }
    }
    const V3 axes[3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
    if (s[a] < EPS) r[a] = axes[a];
    r[a] = normalize(r[a]);
    if (s[b] < EPS) {
        float fx = std::fabs(r[a].x), fy = std::fabs(r[a].y), fz = std::fabs(r[a].z);
        int cc;
        if (fx < fy) { if (fy < fz) cc = 0; else if (fx < fz) cc = 0; else cc = 2; }
        else { if (fx < fz) cc = 1; else if (fy < fz) cc = 1; else cc = 2; }
        r[b] = cross(r[a], axes[cc]);
    }
    r[b] = normalize(r[b]);
    if (s[c] < EPS) r[c] = cross(r[a], r[b]);
    r[c] = normalize(r[c]);
    float basis[16] = {r[0].x, r[0].y, r[0].z, 0, r[1].x, r[1].y, r[1].z, 0, r[2].x, r[2].y, r[2].z, 0, 0, 0, 0, 1};
    float det = det3of4(basis);
    if (det < 0.0f) {
        s[a] = -s[a];
        r[a] = -1.0f * r[a];
        basis[4 * a] = r[a].x; basis[4 * a + 1] = r[a].y; basis[4 * a + 2] = r[a].z;
        det = -det;
    }
    det = det - 1.0f;
    det = det * det;
    q = (EPS < det) ? Q4{0, 0, 0, 1} : rotToQuat(basis);
    sc = {s[0], s[1], s[2]};
}
// Quaternion.Slerp (System.Numerics), epsilon 1e-6.
Q4 slerpQ(Q4 p, Q4 q, float t) {
    float co = p.x * q.x + p.y * q.y + p.z * q.z + p.w * q.w;
    bool neg = co < 0.0f;
    if (neg) co = -co;
    float a, b;
    if (co > 1.0f - 1e-6f) {
        a = 1.0f - t;
        b = neg ? -t : t;
    } else {
        float om = (float)std::acos((double)co);
        float inv = 1.0f / sinF(om);
        a = sinF((1.0f - t) * om) * inv;
        b = sinF(t * om) * inv;
        if (neg) b = -b;
    }
    return {p.x * a + q.x * b, p.y * a + q.y * b, p.z * a + q.z * b, p.w * a + q.w * b};
}
// Matrix4x4.CreateFromQuaternion (System.Numerics).
M4 quatToM(Q4 q) {
    M4 r = identity();
    float xx = q.x * q.x, yy = q.y * q.y, zz = q.z * q.z;
    float xy = q.x * q.y, wz = q.z * q.w, xz = q.z * q.x, wy = q.y * q.w, yz = q.y * q.z, wx = q.x * q.w;
    r.m[0] = 1.0f - 2.0f * (yy + zz); r.m[1] = 2.0f * (xy + wz); r.m[2] = 2.0f * (xz - wy);
    r.m[4] = 2.0f * (xy - wz); r.m[5] = 1.0f - 2.0f * (zz + xx); r.m[6] = 2.0f * (yz + wx);
    r.m[8] = 2.0f * (xz + wy); r.m[9] = 2.0f * (yz - wx); r.m[10] = 1.0f - 2.0f * (yy + xx);
    return r;
}
// Transform.Eval (Base/Scene.fs:27-36) with KeyFrame.Interpolate (:13-19):
// decompose both key frames, Lerp (Vector3.Lerp: a (1 - t) + b t) and Slerp,
// then Transform.Compose (Util/Transform.fs:6-9) as written:
// CreateTranslation(scale) * CreateFromQuaternion(rotation) * CreateScale(translation).
M4 evalTransform(const orc_transform* k, int n, float t) {
    int prev = -1;
    for (int i = 0; i < n; ++i)
        if (k[i].time <= t) prev = i;
    if (prev < 0) return keyFrameMatrix(k[0]);
    if (prev == n - 1) return keyFrameMatrix(k[prev]);
    float ratio = (t - k[prev].time) / (k[prev + 1].time - k[prev].time);
    V3 s1, t1, s2, t2;
    Q4 q1, q2;
    decomposeM(keyFrameMatrix(k[prev]), s1, q1, t1);
    decomposeM(keyFrameMatrix(k[prev + 1]), s2, q2, t2);
    auto lerp3 = [](V3 a, V3 b, float u) { return a * (1.0f - u) + b * u; };
    V3 tr = lerp3(t1, t2, ratio), sc = lerp3(s1, s2, ratio);
    Q4 q = slerpQ(q1, q2, ratio);
    return mul(mul(translateM(sc.x, sc.y, sc.z), quatToM(q)), scaleM(tr.x, tr.y, tr.z));
}

int desc_build(const orc_scene_desc* d, orc_scene* s);

}  // namespace

// ====================================================================== C API
extern "C" {

const char* orc_last_error(void) { return g_err.c_str(); }

int orc_struct_sizes(int32_t* out, int32_t n) {
    int32_t v[] = {(int32_t)sizeof(orc_primitive), (int32_t)sizeof(orc_material), (int32_t)sizeof(orc_light),
                   (int32_t)sizeof(orc_instance), (int32_t)sizeof(orc_transform), (int32_t)sizeof(orc_node),
                   (int32_t)sizeof(orc_camera), (int32_t)sizeof(orc_scene_desc), (int32_t)sizeof(orc_pt_params),
                   (int32_t)sizeof(orc_mlt_params), (int32_t)sizeof(orc_mlt_stats), (int32_t)sizeof(orc_stats),
                   (int32_t)sizeof(orc_bvh_node)};
    int m = (int)(sizeof(v) / sizeof(v[0]));
    for (int i = 0; i < n && i < m; ++i) out[i] = v[i];
    return m;
}

uint32_t orc_xxhash32_two(uint32_t x, uint32_t y) { return xx2(x, y); }
float orc_erfinv(float x) { return erfInv(x); }
uint32_t orc_xxhash32_three(uint32_t x, uint32_t y, uint32_t z) { return xx3(x, y, z); }
float orc_lcg(uint32_t* state) { Sampler s{*state}; float f = s.next(); *state = s.s; return f; }

int orc_alias_build(const float* w, int32_t n, int32_t* alias, float* prob, float* pdf) {
    if (n <= 0) return fail("empty alias table");
    Alias t = alias_build(w, n);
    for (int i = 0; i < n; ++i) { alias[i] = t.alias[i]; prob[i] = t.prob[i]; pdf[i] = t.pdf[i]; }
    return 0;
}
int orc_alias_sample(const int32_t* alias, const float* prob, const float* pdf, int32_t n, float u,
                     int32_t* idx, float* out_pdf) {
    int i; float p;
    alias_sample(alias, prob, pdf, n, u, i, p);
    *idx = i; *out_pdf = p;
    return 0;
}

static void exportNodes(const std::vector<Node>& nodes, orc_bvh_node* out) {
    for (size_t i = 0; i < nodes.size(); ++i) {
        const Node& n = nodes[i];
        orc_bvh_node o{};
        o.pmin[0] = n.b.mn.x; o.pmin[1] = n.b.mn.y; o.pmin[2] = n.b.mn.z;
        o.pmax[0] = n.b.mx.x; o.pmax[1] = n.b.mx.y; o.pmax[2] = n.b.mx.z;
        o.right_or_offset = n.rc;
        o.is_leaf = n.leaf ? 1 : 0;
        o.split_axis = (int8_t)(n.leaf ? 0 : n.axis);
        o.count = (int8_t)(n.leaf ? n.count : 0);
        o.mask = 0;
        out[i] = o;
    }
}

int orc_bvh_build(const float* b, int32_t n, orc_bvh_node* nodes, int32_t* n_nodes, int32_t* perm) {
    if (n <= 0) return fail("empty BVH input");
    std::vector<AABB> bb(n);
    for (int i = 0; i < n; ++i) bb[i] = {{b[6 * i], b[6 * i + 1], b[6 * i + 2]}, {b[6 * i + 3], b[6 * i + 4], b[6 * i + 5]}};
    std::vector<int> pm;
    std::vector<Node> nd;
    try { nd = bvhBuildAll(bb, pm); } catch (const std::exception& e) { return fail(e.what()); }
    if ((int)nd.size() > *n_nodes) return fail("node buffer too small");
    exportNodes(nd, nodes);
    *n_nodes = (int)nd.size();
    for (int i = 0; i < n; ++i) perm[i] = pm[i];
    return 0;
}

int orc_material_sample(const orc_material* md, const float* wo, float u_lobe, const float* u2, float* out7) {
    Material m = toMaterial(*md);
    BSDFSample s = matSample(m, {wo[0], wo[1], wo[2]}, u_lobe, {u2[0], u2[1]});
    float v[7] = {s.e.bsdf.x, s.e.bsdf.y, s.e.bsdf.z, s.e.pdf, s.wi.x, s.wi.y, s.wi.z};
    std::memcpy(out7, v, sizeof(v));
    return 0;
}
int orc_material_eval(const orc_material* md, const float* wo, const float* wi, float* out4) {
    Material m = toMaterial(*md);
    BSDFEval e = matEval(m, {wo[0], wo[1], wo[2]}, {wi[0], wi[1], wi[2]});
    float v[4] = {e.bsdf.x, e.bsdf.y, e.bsdf.z, e.pdf};
    std::memcpy(out4, v, sizeof(v));
    return 0;
}

int orc_material_eval_n(const orc_material* md, int32_t n, const float* wo, const float* wi, float* out4) {
    Material m = toMaterial(*md);
    for (int32_t k = 0; k < n; ++k) {
        BSDFEval e = matEval(m, {wo[3 * k], wo[3 * k + 1], wo[3 * k + 2]}, {wi[3 * k], wi[3 * k + 1], wi[3 * k + 2]});
        out4[4 * k] = e.bsdf.x; out4[4 * k + 1] = e.bsdf.y; out4[4 * k + 2] = e.bsdf.z; out4[4 * k + 3] = e.pdf;
    }
    return 0;
}

int orc_scene_create(const orc_scene_desc* d, orc_scene** out) {
    auto* s = new orc_scene();
    int rc;
    try { rc = desc_build(d, s); } catch (const std::exception& e) { rc = fail(e.what()); }
    if (rc != 0) { delete s; return rc; }
    *out = s;
    return 0;
}
void orc_scene_destroy(orc_scene* s) { delete s; }

int orc_scene_export(orc_scene* s, orc_bvh_node* tlas, int32_t* n_tlas, int32_t* inst_order, int32_t* n_inst,
                     float* o2w, float* w2o, int32_t* light_insts, int32_t* n_lights, float* c2w) {
    if ((int)s->tlas.size() > *n_tlas || (int)s->insts.size() > *n_inst || (int)s->lights.size() > *n_lights)
        return fail("export buffer too small");
    exportNodes(s->tlas, tlas);
    *n_tlas = (int)s->tlas.size();
    for (size_t i = 0; i < s->insts.size(); ++i) {
        inst_order[i] = s->instDesc[i];
        std::memcpy(o2w + 16 * i, s->insts[i].o2w.m, 64);
        std::memcpy(w2o + 16 * i, s->insts[i].w2o.m, 64);
    }
    *n_inst = (int)s->insts.size();
    for (size_t i = 0; i < s->lights.size(); ++i) light_insts[i] = s->lights[i];
    *n_lights = (int)s->lights.size();
    std::memcpy(c2w, s->c2w.m, 64);
    return 0;
}
int orc_scene_export_blas(orc_scene* s, int32_t prim, orc_bvh_node* nodes, int32_t* n_nodes, int32_t* tri,
                          int32_t* n_tris, float* pdf) {
    if (prim < 0 || prim >= (int)s->meshes.size() || s->meshes[prim].sphere) return fail("not a mesh primitive");
    const Mesh& m = s->meshes[prim];
    int nt = (int)m.tri.size() / 3;
    if ((int)m.nodes.size() > *n_nodes || nt > *n_tris) return fail("export buffer too small");
    exportNodes(m.nodes, nodes);
    *n_nodes = (int)m.nodes.size();
    std::memcpy(tri, m.tri.data(), sizeof(int) * m.tri.size());
    for (int i = 0; i < nt; ++i) pdf[i] = m.alias.pdf[i];
    *n_tris = nt;
    return 0;
}

static void fillStats(const Stats& a, orc_stats* o) {
    if (!o) return;
    o->rays_closest = a.raysC; o->rays_shadow = a.raysS;
    o->nodes_closest = a.cc.nodes; o->tris_closest = a.cc.tris; o->insts_closest = a.cc.insts;
    o->nodes_shadow = a.cs.nodes; o->tris_shadow = a.cs.tris; o->insts_shadow = a.cs.insts;
    o->vertices = a.verts; o->paths = a.paths;
}

// ------------------------------------------------------------ adjoint
// FP64 restatement of the derivative the HIP path-replay adjoint computes:
// every path is traced exactly as liPT does (float32, same decisions), its
// vertices are kept, and the derivative of each radiance contribution with
// respect to material albedo / roughness and light emission is accumulated in
// double by direct differentiation (product rule over the BSDF factors of the
// earlier vertices).  Sampled directions, light samples, Russian-roulette
// outcomes and the importance factors 1/pdf are held fixed; the MIS weights
// are differentiated (they depend on roughness through the BSDF pdf).

// PBRMaterial.Eval value (PBR.fs:37-49) in double.
void pbrFD(const double alb[3], double alpha, double met, const double wo[3], const double wi[3], double f[3],
           double* pdf = nullptr) {
    double h[3] = {wo[0] + wi[0], wo[1] + wi[1], wo[2] + wi[2]};
    double hl = std::sqrt(h[0] * h[0] + h[1] * h[1] + h[2] * h[2]);
    for (double& v : h) v /= hl;
    double c = std::fabs(h[2]);
    double x = 1.0 + (alpha * alpha - 1.0) * c * c;
    double d = alpha * alpha / (M_PI * x * x);
    auto lam = [&](const double w[3]) {
        double s2 = w[0] * w[0] + w[1] * w[1];
        if (s2 == 0.0) return 0.0;
        double t2 = s2 / (w[2] * w[2]);
        return (-1.0 + std::sqrt(1.0 + alpha * alpha * t2)) / 2.0;
    };
    double g = 1.0 / (1.0 + lam(wo) + lam(wi));
    double spec = d * g / (4.0 * std::fabs(wo[2]));
    double k = 1.0 - (wo[0] * h[0] + wo[1] * h[1] + wo[2] * h[2]);
    double k5 = k * k * k * k * k;
    double F = 0.04 + 0.96 * k5;
    for (int ch = 0; ch < 3; ++ch) {
        double diff = alb[ch] * std::fabs(wi[2]) / M_PI;
        double fc = alb[ch] + (1.0 - alb[ch]) * k5;
        double inner = diff * (1.0 - F) + spec * F;
        f[ch] = inner * (1.0 - met) + spec * fc * met;
    }
    if (pdf) {
        double cw = wo[0] * h[0] + wo[1] * h[1] + wo[2] * h[2];
        double tt = 0.5 * (1.0 - met);
        *pdf = d * c / (4.0 * cw) * (1.0 - tt) + std::fabs(wi[2]) / M_PI * tt;
    }
}
// Its derivatives: d f_c / d albedo_c and d f_c / d alpha.
void pbrGradD(const double alb[3], double alpha, double met, const double wo[3], const double wi[3], double da[3],
              double dal[3], double* dpdf = nullptr) {
    double h[3] = {wo[0] + wi[0], wo[1] + wi[1], wo[2] + wi[2]};
    double hl = std::sqrt(h[0] * h[0] + h[1] * h[1] + h[2] * h[2]);
    for (double& v : h) v /= hl;
    double c = std::fabs(h[2]);
    double a2 = alpha * alpha;
    double x = 1.0 + (a2 - 1.0) * c * c;
    double d = a2 / (M_PI * x * x);
    double dd = 2.0 * alpha / (M_PI * x * x) - 4.0 * alpha * a2 * c * c / (M_PI * x * x * x);
    auto lam = [&](const double w[3], double& dl) {
        double s2 = w[0] * w[0] + w[1] * w[1];
        dl = 0.0;
        if (s2 == 0.0) return 0.0;
        double t2 = s2 / (w[2] * w[2]);
        double r = std::sqrt(1.0 + a2 * t2);
        dl = alpha * t2 / (2.0 * r);
        return (-1.0 + r) / 2.0;
    };
    double dlo, dli;
    double lo = lam(wo, dlo), li = lam(wi, dli);
    double g = 1.0 / (1.0 + lo + li);
    double dg = -g * g * (dlo + dli);
    double spec = d * g / (4.0 * std::fabs(wo[2]));
    double dspec = (dd * g + d * dg) / (4.0 * std::fabs(wo[2]));
    double k = 1.0 - (wo[0] * h[0] + wo[1] * h[1] + wo[2] * h[2]);
    double k5 = k * k * k * k * k;
    double F = 0.04 + 0.96 * k5;
    for (int ch = 0; ch < 3; ++ch) {
        double fc = alb[ch] + (1.0 - alb[ch]) * k5;
        da[ch] = (1.0 - met) * (1.0 - F) * std::fabs(wi[2]) / M_PI + met * spec * (1.0 - k5);
        dal[ch] = dspec * ((1.0 - met) * F + met * fc);
    }
    if (dpdf) {
        double cw = wo[0] * h[0] + wo[1] * h[1] + wo[2] * h[2];
        *dpdf = dd * c / (4.0 * cw) * (1.0 - 0.5 * (1.0 - met));
    }
}
// d f / d albedo and d f / d roughness at fixed directions for every kind;
// f_pdf is the pdf the forward returned with the value (Lambertian albedo*pdf,
// mirror albedo*1, dielectric albedo*{r, 1-r, 1}).
// dpdf: d pdf / d roughness (the BSDF pdf enters the MIS weights).
void matGradD(const Material& m, V3 wo, V3 wi, float f_pdf, double da[3], double dr[3], double* dpdf = nullptr) {
    for (int ch = 0; ch < 3; ++ch) { da[ch] = 0.0; dr[ch] = 0.0; }
    if (dpdf) *dpdf = 0.0;
    if (m.kind == 3) {
        double alb[3] = {m.albedo.x, m.albedo.y, m.albedo.z};
        double o[3] = {wo.x, wo.y, wo.z}, i[3] = {wi.x, wi.y, wi.z};
        double dal[3], dp;
        pbrGradD(alb, m.alpha, m.metallic, o, i, da, dal, &dp);
        float r = m.roughness;
        double dadr = r * r > 1e-3f ? 2.0 * r : 0.0;
        for (int ch = 0; ch < 3; ++ch) dr[ch] = dal[ch] * dadr;
        if (dpdf) *dpdf = dp * dadr;
        return;
    }
    for (int ch = 0; ch < 3; ++ch) da[ch] = f_pdf;
}

struct GradFactor { int mat; double la[3], lr[3]; };

void liPTGrad(const orc_scene& s, Ray ray, Sampler& smp, int maxDepth, int rrDepth, const double delta[3],
              double* gm, double* gl) {
    if (delta[0] == 0.0 && delta[1] == 0.0 && delta[2] == 0.0) return;   // no adjoint: no gradient
    V3 beta = {1, 1, 1};
    int depth = 0;
    float prevPdf = 0.0f;
    int prevMat = -1;                                       // material whose pdf is prevPdf
    double dPrevPdf = 0.0;                                  // d prevPdf / d roughness of prevMat
    std::vector<GradFactor> fac;
    Stats st;
    // one contribution c (rgb) reaching the film through beta: the earlier
    // BSDF factors' terms
    // A term group (one parameter entry's rgb + roughness, or one light's
    // rgb) that is not finite is dropped, as the GPU adjoint drops it.
    auto add4 = [](double* g, const double v[4]) {
        for (int k = 0; k < 4; ++k)
            if (!std::isfinite(v[k])) return;
        for (int k = 0; k < 4; ++k) g[k] += v[k];
    };
    auto add3 = [](double* g, const double v[3]) {
        for (int k = 0; k < 3; ++k)
            if (!std::isfinite(v[k])) return;
        for (int k = 0; k < 3; ++k) g[k] += v[k];
    };
    auto through = [&](const double c[3]) {
        for (const GradFactor& f : fac) {
            double v[4] = {0, 0, 0, 0};
            for (int ch = 0; ch < 3; ++ch) {
                v[ch] = delta[ch] * c[ch] * f.la[ch];
                v[3] += delta[ch] * c[ch] * f.lr[ch];
            }
            add4(gm + 4 * (size_t)f.mat, v);
        }
    };
    while (depth < maxDepth) {
        float t = INF_F;
        Hit h;
        if (!closest(s, ray, h, t, st.cc)) break;
        const Inst& in = s.insts[h.inst];
        if (in.light >= 0) {
            LightEval le = lightEval(s, ray.o, h);
            float w = depth == 0 ? 1.0f : prevPdf * rcpE(le.pdf + prevPdf);
            V3 lw = le.L * w;
            double c[3] = {(double)beta.x * lw.x, (double)beta.y * lw.y, (double)beta.z * lw.z};
            through(c);
            float woz = toLocal(h.g.onb, normalize(ray.o - h.g.p)).z;
            const LightD& ld = s.lightDefs[in.light];
            double facing = std::fabs(woz) > 1e-6f && (woz > 0.0f || ld.twoSided) ? 1.0 : 0.0;
            double bw[3] = {beta.x, beta.y, beta.z};
            double ve[3];
            for (int ch = 0; ch < 3; ++ch) ve[ch] = delta[ch] * bw[ch] * w * facing;
            add3(gl + 3 * (size_t)in.light, ve);
            if (depth > 0 && prevMat >= 0) {
                // w = prev / (lpdf + prev): d w / d prev = lpdf / (lpdf + prev)^2
                double sden = (double)le.pdf + prevPdf;
                double dw = (double)le.pdf / (sden * sden) * dPrevPdf;
                double lv[3] = {le.L.x, le.L.y, le.L.z};
                double vp[4] = {0, 0, 0, 0};
                for (int ch = 0; ch < 3; ++ch) vp[3] += delta[ch] * bw[ch] * lv[ch] * dw;
                add4(gm + 4 * (size_t)prevMat, vp);
            }
        }
        if (in.material < 0) break;
        const Material& m = s.mats[in.material];
        float uSel = smp.next();
        V2 uL = smp.next2();
        LightSample ls = lightSample(s, h.g.p, uSel, uL);
        Ray sh = {h.g.p, ls.wi};
        float dist = length(ls.p - h.g.p);
        if (ls.pdf != 0.0f && !anyHit(s, sh, dist - 1e-3f, st.cs)) {
            BSDFEval e;
            evalBSDF(s, h, -ray.d, ls.wi, e);
            float rc = rcpE(e.pdf + ls.pdf);
            V3 lb = ls.L * rc;
            V3 a = beta * e.bsdf;
            double c[3] = {(double)a.x * lb.x, (double)a.y * lb.y, (double)a.z * lb.z};
            through(c);
            double da[3], dr[3], dpb;
            matGradD(m, toLocal(h.g.onb, -ray.d), toLocal(h.g.onb, ls.wi), e.pdf, da, dr, &dpb);
            double bl[3] = {(double)beta.x * lb.x, (double)beta.y * lb.y, (double)beta.z * lb.z};
            // d / d roughness of f / (pdf_b + pdf_l): f' / S - f * pdf_b' / S^2
            double dmis = dpb * (double)rc;
            double vm[4] = {0, 0, 0, 0};
            for (int ch = 0; ch < 3; ++ch) {
                vm[ch] = delta[ch] * bl[ch] * da[ch];
                vm[3] += delta[ch] * bl[ch] * dr[ch] - delta[ch] * c[ch] * dmis;
            }
            add4(gm + 4 * (size_t)in.material, vm);
            // the sampled light's emission (Light.fs:49-52 facing test)
            const double av[3] = {a.x, a.y, a.z};
            double vl[3];
            for (int ch = 0; ch < 3; ++ch) vl[ch] = delta[ch] * av[ch] * rc * ls.facing;
            add3(gl + 3 * (size_t)ls.ldef, vl);
        }
        float ul = smp.next();
        V2 ub = smp.next2();
        BSDFSample bs = sampleBSDF(s, h, -ray.d, ul, ub);
        prevPdf = bs.e.pdf;
        if (bs.e.pdf == 0.0f) break;
        {
            GradFactor f;
            f.mat = in.material;
            double da[3], dr[3];
            matGradD(m, toLocal(h.g.onb, -ray.d), toLocal(h.g.onb, bs.wi), bs.e.pdf, da, dr, &dPrevPdf);
            prevMat = in.material;
            const float fv[3] = {bs.e.bsdf.x, bs.e.bsdf.y, bs.e.bsdf.z};
            for (int ch = 0; ch < 3; ++ch) {
                f.la[ch] = fv[ch] != 0.0f ? da[ch] / fv[ch] : 0.0;
                f.lr[ch] = fv[ch] != 0.0f ? dr[ch] / fv[ch] : 0.0;
            }
            fac.push_back(f);
        }
        ray = {h.g.p, bs.wi};
        beta = beta * bs.e.bsdf * rcpE(bs.e.pdf);
        if (depth >= rrDepth) {
            float q = minN(1.0f, maxN(beta.x, maxN(beta.y, beta.z)));
            if (smp.next() < q) beta = beta * rcpE(q);
            else break;
        }
        depth = depth + 1;
    }
}

int orc_render(orc_scene* s, int32_t integ, const orc_pt_params* p, int32_t n_threads, float* film,
               orc_stats* stats, int32_t x0, int32_t y0, int32_t x1, int32_t y1) {
    // Scene.Render builds the UniformLightSampler before any integrator runs
    // (Extensions/Scene/Render.fs:13-14), so the normal view fails as well.
    if (s->lights.empty()) return fail("LightSamplerBase: No light primitives found.");
    if (x1 <= x0 || y1 <= y0) { x0 = 0; y0 = 0; x1 = p->width; y1 = p->height; }
    const int TS = 16;                                      // Integrator.fs:16
    int tx = (p->width + TS - 1) / TS, ty = (p->height + TS - 1) / TS;
    int nt = std::max(1, (int)n_threads);
    std::vector<Stats> st(nt);
    std::atomic<int> next(0);
    auto work = [&](int w) {
        for (;;) {
            int tile = next.fetch_add(1);
            if (tile >= tx * ty) break;
            int bx = tile % tx, by = tile / tx;
            for (int y = std::max(by * TS, y0); y < std::min(std::min((by + 1) * TS, p->height), y1); ++y)
                for (int x = std::max(bx * TS, x0); x < std::min(std::min((bx + 1) * TS, p->width), x1); ++x) {
                    V3 c = renderPixel(*s, integ, *p, x, y, st[w]);
                    size_t idx = (size_t)(p->height - y - 1) * p->width + x;   // Film.fs:43
                    film[3 * idx] = c.x; film[3 * idx + 1] = c.y; film[3 * idx + 2] = c.z;
                }
        }
    };
    std::vector<std::thread> th;
    for (int w = 1; w < nt; ++w) th.emplace_back(work, w);
    work(0);
    for (auto& t : th) t.join();
    Stats tot;
    for (auto& a : st) tot.add(a);
    fillStats(tot, stats);
    return 0;
}

int orc_trace_sample(orc_scene* s, const orc_pt_params* p, int32_t x, int32_t y, int32_t k, float* rgb) {
    if (s->lights.empty()) return fail("LightSamplerBase: No light primitives found.");
    Stats st;
    Sampler smp{xx3((uint32_t)x, (uint32_t)y, (uint32_t)(p->frame_id * p->spp + k))};
    V2 uPix = smp.next2();
    V2 uLens = smp.next2();
    Ray r = generatePrimary(*s, p->width, p->height, x, y, uPix, uLens);
    V3 L = liPT(*s, r, smp, p->max_depth, p->rr_depth, st);
    rgb[0] = L.x; rgb[1] = L.y; rgb[2] = L.z;
    return 0;
}

int orc_set_params(orc_scene* s, const float* mp, int32_t n_mat, const float* le, int32_t n_light) {
    if (mp) {
        if (n_mat != (int)s->mats.size()) return fail("material parameter count mismatch");
        for (int i = 0; i < n_mat; ++i) {
            Material& m = s->mats[i];
            m.albedo = {mp[4 * i], mp[4 * i + 1], mp[4 * i + 2]};
            m.roughness = mp[4 * i + 3];
            m.alpha = maxN(m.roughness * m.roughness, 1e-3f);
        }
    }
    if (le) {
        if (n_light != (int)s->lightDefs.size()) return fail("light parameter count mismatch");
        for (int i = 0; i < n_light; ++i) s->lightDefs[i].emission = {le[3 * i], le[3 * i + 1], le[3 * i + 2]};
    }
    return 0;
}

int orc_adjoint_pt(orc_scene* s, const orc_pt_params* p, const float* d_film, int32_t n_threads, int32_t x0,
                   int32_t y0, int32_t x1, int32_t y1, double* d_mat, double* d_light) {
    if (s->lights.empty()) return fail("LightSamplerBase: No light primitives found.");
    if (x1 <= x0 || y1 <= y0) { x0 = 0; y0 = 0; x1 = p->width; y1 = p->height; }
    const size_t nm = 4 * s->mats.size(), nl = 3 * s->lightDefs.size();
    int nt = std::max(1, (int)n_threads);
    std::vector<std::vector<double>> gm(nt, std::vector<double>(nm, 0.0)), gl(nt, std::vector<double>(nl, 0.0));
    std::atomic<int> next(0);
    const float inv_spp = 1.0f / (float)p->spp;
    auto work = [&](int w) {
        for (;;) {
            int y = y0 + next.fetch_add(1);
            if (y >= y1) break;
            for (int x = x0; x < x1; ++x) {
                size_t idx = (size_t)(p->height - y - 1) * p->width + x;
                const double delta[3] = {(double)(d_film[3 * idx] * inv_spp), (double)(d_film[3 * idx + 1] * inv_spp),
                                         (double)(d_film[3 * idx + 2] * inv_spp)};
                for (int k = 0; k < p->spp; ++k) {
                    Sampler smp{xx3((uint32_t)x, (uint32_t)y, (uint32_t)(p->frame_id * p->spp + k))};
                    V2 uPix = smp.next2();
                    V2 uLens = smp.next2();
                    Ray r = generatePrimary(*s, p->width, p->height, x, y, uPix, uLens);
                    liPTGrad(*s, r, smp, p->max_depth, p->rr_depth, delta, gm[w].data(), gl[w].data());
                }
            }
        }
    };
    std::vector<std::thread> th;
    for (int w = 1; w < nt; ++w) th.emplace_back(work, w);
    work(0);
    for (auto& t : th) t.join();
    for (size_t i = 0; i < nm; ++i) { d_mat[i] = 0.0; for (int w = 0; w < nt; ++w) d_mat[i] += gm[w][i]; }
    for (size_t i = 0; i < nl; ++i) { d_light[i] = 0.0; for (int w = 0; w < nt; ++w) d_light[i] += gl[w][i]; }
    return 0;
}

int orc_pbr_grad_d(const double* alb, double alpha, double metallic, const double* wo, const double* wi,
                   double* f, double* d_albedo, double* d_alpha) {
    pbrFD(alb, alpha, metallic, wo, wi, f, f + 3);
    pbrGradD(alb, alpha, metallic, wo, wi, d_albedo, d_alpha, d_alpha + 3);
    return 0;
}

int orc_camera_ray(orc_scene* s, const orc_pt_params* p, int32_t x, int32_t y, int32_t k, float* out) {
    Sampler smp{xx3((uint32_t)x, (uint32_t)y, (uint32_t)(p->frame_id * p->spp + k))};
    V2 uPix = smp.next2();
    V2 uLens = smp.next2();
    Ray r = generatePrimary(*s, p->width, p->height, x, y, uPix, uLens);
    out[0] = r.o.x; out[1] = r.o.y; out[2] = r.o.z;
    out[3] = r.d.x; out[4] = r.d.y; out[5] = r.d.z;
    return 0;
}

// Camera ray of pixel (x, y) for given uPixel / uLens draws (Camera.fs:12-20).
int orc_camera_ray_u(orc_scene* s, int32_t W, int32_t H, int32_t x, int32_t y, const float* uPix, const float* uLens,
                     float* out) {
    Ray r = generatePrimary(*s, W, H, x, y, {uPix[0], uPix[1]}, {uLens[0], uLens[1]});
    out[0] = r.o.x; out[1] = r.o.y; out[2] = r.o.z;
    out[3] = r.d.x; out[4] = r.d.y; out[5] = r.d.z;
    return 0;
}

int orc_first_hit(orc_scene* s, const orc_pt_params* p, int32_t k, int32_t* inst, int32_t* prim, float* tout) {
    for (int y = 0; y < p->height; ++y)
        for (int x = 0; x < p->width; ++x) {
            Sampler smp{xx3((uint32_t)x, (uint32_t)y, (uint32_t)(p->frame_id * p->spp + k))};
            V2 uPix = smp.next2();
            V2 uLens = smp.next2();
            Ray r = generatePrimary(*s, p->width, p->height, x, y, uPix, uLens);
            Hit h;
            float t = INF_F;
            Counters c;
            size_t i = (size_t)y * p->width + x;
            if (closest(*s, r, h, t, c)) { inst[i] = h.inst; prim[i] = h.prim; tout[i] = t; }
            else { inst[i] = -1; prim[i] = -1; tout[i] = INF_F; }
        }
    return 0;
}

int orc_render_pssmlt(orc_scene* s, const orc_mlt_params* p, int32_t n_threads, float* film,
                      orc_mlt_stats* stats, int32_t c0, int32_t c1, float* bootOut, int32_t* chainAcc) {
    if (s->lights.empty()) return fail("LightSamplerBase: No light primitives found.");
    const int W = p->width, H = p->height;
    std::memset(film, 0, sizeof(float) * 3 * (size_t)W * H);   // Film.Clear (Render.fs:11)
    int nt = std::max(1, (int)n_threads);
    int nx = 4 + 7 * p->max_depth;
    std::vector<float> boot(p->n_bootstrap);
    // PSSMLT.fs:382-392: bootstrap paths.
    {
        std::atomic<int> next(0);
        auto work = [&]() {
            std::vector<PrimarySample> xs(nx);
            Stats st;
            for (;;) {
                int b0 = next.fetch_add(256);
                if (b0 >= p->n_bootstrap) break;
                for (int b = b0; b < std::min(b0 + 256, p->n_bootstrap); ++b) {
                    MLTSampler m{Sampler{xx2((uint32_t)p->frame_id, (uint32_t)b)}, p->large_step_prob, p->strategy,
                                 p->sigma, p->eps_min, p->eps_max, xs.data()};
                    boot[b] = mltPath(*s, *p, m, st).y;
                }
            }
        };
        std::vector<std::thread> th;
        for (int w = 1; w < nt; ++w) th.emplace_back(work);
        work();
        for (auto& t : th) t.join();
    }
    if (bootOut) std::memcpy(bootOut, boot.data(), sizeof(float) * p->n_bootstrap);
    float sum = 0.0f;                                        // Array.average (PSSMLT.fs:394)
    for (int i = 0; i < p->n_bootstrap; ++i) sum = sum + boot[i];
    float B = sum / (float)p->n_bootstrap;
    int mpc = (int)(((uint64_t)p->spp * (uint64_t)W * (uint64_t)H + (uint64_t)p->n_chains - 1ull) / (uint64_t)p->n_chains);
    if (stats) { stats->b = B; stats->mutations_per_chain = mpc; stats->accepted = 0; stats->proposed = 0; }
    if (chainAcc) std::memset(chainAcc, 0, sizeof(int32_t) * p->n_chains);
    if (B == 0.0f) return 0;                                 // PSSMLT.fs:396-397
    Alias table = alias_build(boot.data(), p->n_bootstrap);  // PSSMLT.fs:399
    if (c1 <= c0) { c0 = 0; c1 = p->n_chains; }
    float invEff = rcpE((float)mpc * (float)p->n_chains / (float)(W * H));
    float invB = rcpE(B);
    std::vector<std::vector<float>> films(nt, std::vector<float>(3 * (size_t)W * H, 0.0f));
    std::vector<int64_t> acc(nt, 0), prop(nt, 0);
    auto work = [&](int w) {
        std::vector<PrimarySample> xs(nx);
        Stats st;
        float* f = films[w].data();
        auto accum = [&](int x, int y, V3 c) {               // Film.Accumulate (Film.fs:48-53)
            size_t i = (size_t)(H - y - 1) * W + x;
            f[3 * i] = f[3 * i] + c.x; f[3 * i + 1] = f[3 * i + 1] + c.y; f[3 * i + 2] = f[3 * i + 2] + c.z;
        };
        for (int ch = c0 + w; ch < c1; ch += nt) {            // PSSMLT.fs:275-377
            Sampler outer{xx2((uint32_t)p->frame_id, (uint32_t)ch)};
            int bid; float bp;
            alias_sample(table.alias.data(), table.prob.data(), table.pdf.data(), p->n_bootstrap, outer.next(), bid, bp);
            MLTSampler m{Sampler{xx2((uint32_t)p->frame_id, (uint32_t)bid)}, p->large_step_prob, p->strategy,
                         p->sigma, p->eps_min, p->eps_max, xs.data()};
            MLTPath cur = mltPath(*s, *p, m, st);
            m.accept();
            m.inner = Sampler{xx3((uint32_t)ch, (uint32_t)bid, (uint32_t)p->frame_id)};
            V3 rad = {0, 0, 0};
            int accepted = 0;
            for (int it = 0; it < mpc; ++it) {
                MLTPath nw = mltPath(*s, *p, m, st);
                float a = minN(1.0f, nw.y / cur.y);
                float wOld = (1.0f - a) / std::fma(cur.y, invB, p->large_step_prob);
                rad = rad + wOld * cur.L;
                float wNew = (a + (m.large ? 1.0f : 0.0f)) / std::fma(nw.y, invB, p->large_step_prob);
                if (outer.next() < a) {
                    accepted++;
                    accum(cur.px, cur.py, rad * invEff);
                    rad = wNew * nw.L;
                    cur = nw;
                    m.accept();
                } else {
                    if (a > 0.0f) accum(nw.px, nw.py, (wNew * invEff) * nw.L);
                    m.reject();
                }
            }
            accum(cur.px, cur.py, rad * invEff);
            acc[w] += accepted;
            prop[w] += mpc;
            if (chainAcc) chainAcc[ch] = accepted;
        }
    };
    std::vector<std::thread> th;
    for (int w = 1; w < nt; ++w) th.emplace_back(work, w);
    work(0);
    for (auto& t : th) t.join();
    for (int w = 0; w < nt; ++w)
        for (size_t i = 0; i < films[w].size(); ++i) film[i] = film[i] + films[w][i];
    if (stats) for (int w = 0; w < nt; ++w) { stats->accepted += acc[w]; stats->proposed += prop[w]; }
    return 0;
}

}  // extern "C"

namespace {
// Scene.Load + Scene.Traverse + BVHAggregate + UniformLightSampler
// (Loader.fs:253-272, Base/Scene.fs:43-51, Render.fs:11-14).
int desc_build(const orc_scene_desc* d, orc_scene* s) {
    s->cam = d->camera;
    for (int i = 0; i < d->n_materials; ++i) s->mats.push_back(toMaterial(d->materials[i]));
    for (int i = 0; i < d->n_lights; ++i) {
        const orc_light& l = d->lights[i];
        s->lightDefs.push_back({{l.emission[0], l.emission[1], l.emission[2]}, l.two_sided != 0});
    }
    // MeshPrimitive construction: BLAS build permutes the triangle indices
    // (Mesh.fs:157-169); the area alias table is built after it (:181-186).
    for (int i = 0; i < d->n_primitives; ++i) {
        const orc_primitive& p = d->primitives[i];
        Mesh m;
        if (p.kind == 1) { m.sphere = true; m.radius = p.radius; s->meshes.push_back(std::move(m)); continue; }
        if (p.n_indices < 3 || p.n_indices % 3 != 0) return fail("mesh without triangles");
        for (int k = 0; k < p.n_verts; ++k) {
            const float* v = d->vertices + 3 * ((int64_t)p.vert_offset + k);
            m.v.push_back({v[0], v[1], v[2]});
        }
        int nt = p.n_indices / 3;
        std::vector<int> ix(d->indices + p.index_offset, d->indices + p.index_offset + p.n_indices);
        for (int v : ix) if (v < 0 || v >= p.n_verts) return fail("mesh index out of range");
        std::vector<AABB> bb(nt);
        for (int k = 0; k < nt; ++k) bb[k] = triBounds({m.v[ix[3 * k]], m.v[ix[3 * k + 1]], m.v[ix[3 * k + 2]]});
        std::vector<int> perm;
        m.nodes = bvhBuildAll(bb, perm);
        m.tri.resize(3 * nt);
        for (int k = 0; k < nt; ++k)
            for (int j = 0; j < 3; ++j) m.tri[3 * k + j] = ix[3 * perm[k] + j];
        std::vector<float> w(nt);
        for (int k = 0; k < nt; ++k) w[k] = triArea(m.get(k));
        m.alias = alias_build(w.data(), nt);
        s->meshes.push_back(std::move(m));
    }
    std::vector<M4> locals(d->n_transforms);
    for (int i = 0; i < d->n_transforms; ++i) {               // Transform.Eval (Base/Scene.fs:27-36)
        const int nk = d->transforms[i].n_keys > 0 ? d->transforms[i].n_keys : 1;
        if (i + nk > d->n_transforms) return fail("key frame run past the transform array");
        locals[i] = evalTransform(d->transforms + i, nk, d->time);
    }
    // Node.Traverse (Base/Scene.fs:43-51).  Instances are shared objects: the
    // last UpdateTransform of an instance wins for every list entry of it.
    std::vector<int> order;
    std::vector<M4> instO2W(d->n_instances, identity());
    M4 c2w{};                                                // Unchecked.defaultof<Matrix4x4>
    int visits = 0;
    std::function<void(int, const M4&, int)> trav = [&](int ni, const M4& parent, int depth) {
        if (depth > 1000 || ++visits > 10000000) throw std::runtime_error("node graph has a cycle");
        const orc_node& n = d->nodes[ni];
        M4 local = n.transform >= 0 ? locals[n.transform] : identity();
        M4 o2w = mul(local, parent);
        if (n.has_camera) c2w = o2w;
        for (int k = 0; k < n.n_inst; ++k) {
            int ii = d->node_refs[n.inst_offset + k];
            instO2W[ii] = o2w;
            order.push_back(ii);
        }
        for (int k = 0; k < n.n_child; ++k) trav(d->node_refs[n.child_offset + k], o2w, depth + 1);
    };
    trav(d->root, identity(), 0);
    s->c2w = c2w;
    if (order.empty()) return fail("scene has no primitive instances");
    std::vector<Inst> list;
    std::vector<AABB> bb;
    for (int ii : order) {
        const orc_instance& di = d->instances[ii];
        Inst in;
        in.prim = di.primitive; in.material = di.material; in.light = di.light;
        in.o2w = instO2W[ii];
        in.w2o = invert(in.o2w);                              // Primitive.fs:131-138
        in.bounds = aabbTransform(s->meshes[in.prim].bounds(), in.o2w);
        list.push_back(in);
        bb.push_back(in.bounds);
    }
    std::vector<int> perm;
    s->tlas = bvhBuildAll(bb, perm);                          // Extensions/Aggregate/BVH.fs:9
    for (int p : perm) { s->insts.push_back(list[p]); s->instDesc.push_back(order[p]); }
    for (int i = 0; i < (int)s->insts.size(); ++i)            // Base/LightSampler.fs:7
        if (s->insts[i].light >= 0) s->lights.push_back(i);
    return 0;
}
}  // namespace
