// host_scene.cpp -- see host_scene.hpp.
#include "host_scene.hpp"

#include <algorithm>

namespace bnh {

M4 operator*(const M4& a, const M4& b) {
    // Matrix4x4 product, each element summed left to right.
    M4 r;
    for (int i = 0; i < 4; ++i) {
        const float* ar = a.m + 4 * i;
        for (int j = 0; j < 4; ++j)
            r.m[4 * i + j] = ar[0] * b.m[j] + ar[1] * b.m[4 + j] + ar[2] * b.m[8 + j] + ar[3] * b.m[12 + j];
    }
    return r;
}

V3 transform_point(V3 v, const M4& M) {
    const float* m = M.m;
    return {v.x * m[0] + v.y * m[4] + v.z * m[8] + m[12],
            v.x * m[1] + v.y * m[5] + v.z * m[9] + m[13],
            v.x * m[2] + v.y * m[6] + v.z * m[10] + m[14]};
}

M4 invert(const M4& M) {
    // Cofactor expansion through the 2x2 minors of the lower two rows (for
    // the first two result columns) and of rows 2/4 and 2/3 (for the others).
    const float* q = M.m;
    const float a = q[0], b = q[1], c = q[2], d = q[3];
    const float e = q[4], f = q[5], g = q[6], h = q[7];
    const float i = q[8], j = q[9], k = q[10], l = q[11];
    const float m = q[12], n = q[13], o = q[14], p = q[15];
    const float m0 = k * p - l * o, m1 = j * p - l * n, m2 = j * o - k * n;
    const float m3 = i * p - l * m, m4 = i * o - k * m, m5 = i * n - j * m;
    const float c11 = +(f * m0 - g * m1 + h * m2);
    const float c12 = -(e * m0 - g * m3 + h * m4);
    const float c13 = +(e * m1 - f * m3 + h * m5);
    const float c14 = -(e * m2 - f * m4 + g * m5);
    const float det = a * c11 + b * c12 + c * c13 + d * c14;
    M4 r;
    if (std::fabs(det) < std::numeric_limits<float>::denorm_min()) {
        for (float& v : r.m) v = std::numeric_limits<float>::quiet_NaN();
        return r;
    }
    const float s = 1.0f / det;
    r.m[0] = c11 * s; r.m[4] = c12 * s; r.m[8] = c13 * s; r.m[12] = c14 * s;
    r.m[1] = -(b * m0 - c * m1 + d * m2) * s;
    r.m[5] = +(a * m0 - c * m3 + d * m4) * s;
    r.m[9] = -(a * m1 - b * m3 + d * m5) * s;
    r.m[13] = +(a * m2 - b * m4 + c * m5) * s;
    const float n0 = g * p - h * o, n1 = f * p - h * n, n2 = f * o - g * n;
    const float n3 = e * p - h * m, n4 = e * o - g * m, n5 = e * n - f * m;
    r.m[2] = +(b * n0 - c * n1 + d * n2) * s;
    r.m[6] = -(a * n0 - c * n3 + d * n4) * s;
    r.m[10] = +(a * n1 - b * n3 + d * n5) * s;
    r.m[14] = -(a * n2 - b * n4 + c * n5) * s;
    const float k0 = g * l - h * k, k1 = f * l - h * j, k2 = f * k - g * j;
    const float k3 = e * l - h * i, k4 = e * k - g * i, k5 = e * j - f * i;
    r.m[3] = -(b * k0 - c * k1 + d * k2) * s;
    r.m[7] = +(a * k0 - c * k3 + d * k4) * s;
    r.m[11] = -(a * k1 - b * k3 + d * k5) * s;
    r.m[15] = +(a * k2 - b * k4 + c * k5) * s;
    return r;
}

static float cos_f(float x) { return (float)std::cos((double)x); }
static float sin_f(float x) { return (float)std::sin((double)x); }

M4 keyframe_matrix(const bn_transform& t) {
    M4 S = M4::identity(), R = M4::identity(), T = M4::identity();
    if (t.flags & BN_XF_SCALE) { S.m[0] = t.scale[0]; S.m[5] = t.scale[1]; S.m[10] = t.scale[2]; }
    if (t.flags & BN_XF_ROTATION) {
        // CreateRotationX * CreateRotationY * CreateRotationZ, radians.
        M4 X = M4::identity(), Y = M4::identity(), Z = M4::identity();
        float cx = cos_f(t.rotation[0]), sx = sin_f(t.rotation[0]);
        X.m[5] = cx; X.m[6] = sx; X.m[9] = -sx; X.m[10] = cx;
        float cy = cos_f(t.rotation[1]), sy = sin_f(t.rotation[1]);
        Y.m[0] = cy; Y.m[2] = -sy; Y.m[8] = sy; Y.m[10] = cy;
        float cz = cos_f(t.rotation[2]), sz = sin_f(t.rotation[2]);
        Z.m[0] = cz; Z.m[1] = sz; Z.m[4] = -sz; Z.m[5] = cz;
        R = X * Y * Z;
    }
    if (t.flags & BN_XF_TRANSLATION) { T.m[12] = t.translation[0]; T.m[13] = t.translation[1]; T.m[14] = t.translation[2]; }
    return S * R * T;
}

// ---------------------------------------------------- key-frame animation
// Transform.Eval (Base/Scene.fs:27-36) and KeyFrame.Interpolate (:13-19) over
// the System.Numerics routines they call (.NET 9 BCL, which the reference
// does not vendor): Matrix4x4.Decompose, Quaternion.CreateFromRotationMatrix,
// Quaternion.Slerp, Vector3.Lerp, and Transform.Compose with its argument
// order as written (Util/Transform.fs:6-9): CreateTranslation(scale) *
// CreateFromQuaternion(rotation) * CreateScale(translation).
namespace {

struct Quat { float x, y, z, w; };

float det4(const M4& M) {                                 // Matrix4x4.GetDeterminant
    const float* q = M.m;
    const float a = q[0], b = q[1], c = q[2], d = q[3], e = q[4], f = q[5], g = q[6], h = q[7];
    const float i = q[8], j = q[9], k = q[10], l = q[11], m = q[12], n = q[13], o = q[14], p = q[15];
    const float kp_lo = k * p - l * o, jp_ln = j * p - l * n, jo_kn = j * o - k * n;
    const float ip_lm = i * p - l * m, io_km = i * o - k * m, in_jm = i * n - j * m;
    return a * (f * kp_lo - g * jp_ln + h * jo_kn) - b * (e * kp_lo - g * ip_lm + h * io_km) +
           c * (e * jp_ln - f * ip_lm + h * in_jm) - d * (e * jo_kn - f * io_km + g * in_jm);
}

Quat quat_from_rotation(const M4& M) {                    // Quaternion.CreateFromRotationMatrix
    const float* m = M.m;
    const float m11 = m[0], m12 = m[1], m13 = m[2], m21 = m[4], m22 = m[5], m23 = m[6];
    const float m31 = m[8], m32 = m[9], m33 = m[10];
    const float trace = m11 + m22 + m33;
    Quat q;
    if (trace > 0.0f) {
        float s = std::sqrt(trace + 1.0f);
        q.w = s * 0.5f;
        s = 0.5f / s;
        q.x = (m23 - m32) * s; q.y = (m31 - m13) * s; q.z = (m12 - m21) * s;
    } else if (m11 >= m22 && m11 >= m33) {
        const float s = std::sqrt(1.0f + m11 - m22 - m33), inv = 0.5f / s;
        q.x = 0.5f * s; q.y = (m12 + m21) * inv; q.z = (m13 + m31) * inv; q.w = (m23 - m32) * inv;
    } else if (m22 > m33) {
        const float s = std::sqrt(1.0f + m22 - m11 - m33), inv = 0.5f / s;
        q.x = (m21 + m12) * inv; q.y = 0.5f * s; q.z = (m32 + m23) * inv; q.w = (m31 - m13) * inv;
    } else {
        const float s = std::sqrt(1.0f + m33 - m11 - m22), inv = 0.5f / s;
        q.x = (m31 + m13) * inv; q.y = (m32 + m23) * inv; q.z = 0.5f * s; q.w = (m12 - m21) * inv;
    }
    return q;
}

// Matrix4x4.Decompose: scales are the lengths of the basis rows; the rows are
// normalised largest scale first, a degenerate row (scale < 1e-4) replaced by
// a canonical axis or a cross product; a left-handed basis negates the
// largest scale; a basis whose determinant is not 1 (within 1e-4 after
// squaring the difference) gives the identity rotation.
void decompose(const M4& M, V3& scale, Quat& rot, V3& trans) {
    constexpr float eps = 1e-4f;
    const V3 canon[3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
    trans = {M.m[12], M.m[13], M.m[14]};
    V3 row[3] = {{M.m[0], M.m[1], M.m[2]}, {M.m[4], M.m[5], M.m[6]}, {M.m[8], M.m[9], M.m[10]}};
    float sc[3] = {length(row[0]), length(row[1]), length(row[2])};
    int a, b, c;
    const float x = sc[0], y = sc[1], z = sc[2];
    if (x < y) {
        if (y < z) { a = 2; b = 1; c = 0; }
        else { a = 1; if (x < z) { b = 2; c = 0; } else { b = 0; c = 2; } }
    } else {
        if (x < z) { a = 2; b = 0; c = 1; }
        else { a = 0; if (y < z) { b = 2; c = 1; } else { b = 1; c = 2; } }
    }
    if (sc[a] < eps) row[a] = canon[a];
    row[a] = normalize(row[a]);
    if (sc[b] < eps) {
        const float ax = std::fabs(row[a].x), ay = std::fabs(row[a].y), az = std::fabs(row[a].z);
        int cc;
        if (ax < ay) cc = (ay < az) ? 0 : (ax < az ? 0 : 2);
        else cc = (ax < az) ? 1 : (ay < az ? 1 : 2);
        row[b] = cross(row[a], canon[cc]);
    }
    row[b] = normalize(row[b]);
    if (sc[c] < eps) row[c] = cross(row[a], row[b]);
    row[c] = normalize(row[c]);
    M4 basis = M4::identity();
    for (int r = 0; r < 3; ++r) { basis.m[4 * r] = row[r].x; basis.m[4 * r + 1] = row[r].y; basis.m[4 * r + 2] = row[r].z; }
    float det = det4(basis);
    if (det < 0.0f) {
        sc[a] = -sc[a];
        row[a] = -row[a];
        basis.m[4 * a] = row[a].x; basis.m[4 * a + 1] = row[a].y; basis.m[4 * a + 2] = row[a].z;
        det = -det;
    }
    det -= 1.0f;
    det *= det;
    rot = eps < det ? Quat{0, 0, 0, 1} : quat_from_rotation(basis);
    scale = {sc[0], sc[1], sc[2]};
}

Quat slerp(Quat q1, Quat q2, float t) {                   // Quaternion.Slerp
    constexpr float eps = 1e-6f;
    float cosw = q1.x * q2.x + q1.y * q2.y + q1.z * q2.z + q1.w * q2.w;
    bool flip = false;
    if (cosw < 0.0f) { flip = true; cosw = -cosw; }
    float s1, s2;
    if (cosw > 1.0f - eps) {
        s1 = 1.0f - t;
        s2 = flip ? -t : t;
    } else {
        const float w = (float)std::acos((double)cosw);
        const float inv = 1.0f / (float)std::sin((double)w);
        s1 = (float)std::sin((double)((1.0f - t) * w)) * inv;
        const float b = (float)std::sin((double)(t * w)) * inv;
        s2 = flip ? -b : b;
    }
    return {q1.x * s1 + q2.x * s2, q1.y * s1 + q2.y * s2, q1.z * s1 + q2.z * s2, q1.w * s1 + q2.w * s2};
}

V3 lerp(V3 a, V3 b, float t) { return a * (1.0f - t) + b * t; }   // Vector3.Lerp

M4 from_quat(Quat q) {                                    // Matrix4x4.CreateFromQuaternion
    const float xx = q.x * q.x, yy = q.y * q.y, zz = q.z * q.z;
    const float xy = q.x * q.y, wz = q.z * q.w, xz = q.z * q.x, wy = q.y * q.w, yz = q.y * q.z, wx = q.x * q.w;
    M4 r = M4::identity();
    r.m[0] = 1.0f - 2.0f * (yy + zz); r.m[1] = 2.0f * (xy + wz); r.m[2] = 2.0f * (xz - wy);
    r.m[4] = 2.0f * (xy - wz); r.m[5] = 1.0f - 2.0f * (zz + xx); r.m[6] = 2.0f * (yz + wx);
    r.m[8] = 2.0f * (xz + wy); r.m[9] = 2.0f * (yz - wx); r.m[10] = 1.0f - 2.0f * (yy + xx);
    return r;
}

}  // namespace

M4 transform_eval(const bn_transform* keys, int n, float time) {
    if (n <= 0) throw Error(BN_E_INVALID, "KeyFrames cannot be empty");
    int prev = -1;                                        // Array.tryFindIndexBack (k.Time <= t)
    for (int i = n - 1; i >= 0; --i)
        if (keys[i].time <= time) { prev = i; break; }
    if (prev < 0) return keyframe_matrix(keys[0]);
    if (prev == n - 1) return keyframe_matrix(keys[prev]);
    const bn_transform& A = keys[prev];
    const bn_transform& B = keys[prev + 1];
    const float ratio = (time - A.time) / (B.time - A.time);
    V3 s1, t1, s2, t2;
    Quat r1, r2;
    decompose(keyframe_matrix(A), s1, r1, t1);
    decompose(keyframe_matrix(B), s2, r2, t2);
    const V3 tr = lerp(t1, t2, ratio);
    const Quat rq = slerp(r1, r2, ratio);
    const V3 sc = lerp(s1, s2, ratio);
    M4 T = M4::identity(), S = M4::identity();            // Compose(scale, rotation, translation)
    T.m[12] = sc.x; T.m[13] = sc.y; T.m[14] = sc.z;       // CreateTranslation(scale)
    S.m[0] = tr.x; S.m[5] = tr.y; S.m[10] = tr.z;         // CreateScale(translation)
    return T * from_quat(rq) * S;
}

Box transform_box(const Box& b, const M4& m) {
    Box r = Box::empty();
    for (int c = 0; c < 8; ++c) {
        V3 p = {(c & 1) ? b.hi.x : b.lo.x, (c & 2) ? b.hi.y : b.lo.y, (c & 4) ? b.hi.z : b.lo.z};
        r.grow(transform_point(p, m));
    }
    return r;
}

Box HostMesh::bounds() const {
    if (sphere) return {{-radius, -radius, -radius}, {radius, radius, radius}};
    const bn_bvh_node& n = nodes[0];
    return {{n.pmin[0], n.pmin[1], n.pmin[2]}, {n.pmax[0], n.pmax[1], n.pmax[2]}};
}

// ------------------------------------------------------------------ BVH
namespace {
struct BuildItem { Box b; int32_t id; };
struct Task { int first, last, depth; int parent; };   // parent >= 0: fill its right child

bn_bvh_node make_node(const Box& b, int32_t word, bool leaf, int axis, int count) {
    bn_bvh_node n{};
    n.pmin[0] = b.lo.x; n.pmin[1] = b.lo.y; n.pmin[2] = b.lo.z;
    n.pmax[0] = b.hi.x; n.pmax[1] = b.hi.y; n.pmax[2] = b.hi.z;
    n.right_or_offset = word;
    n.is_leaf = leaf ? 1 : 0;
    n.split_axis = (int8_t)axis;
    n.count = (int8_t)count;
    n.mask = 0;
    return n;
}
}  // namespace

void build_bvh(const std::vector<Box>& in, std::vector<bn_bvh_node>& nodes, std::vector<int32_t>& perm) {
    constexpr int kBins = 12, kLeaf = 4, kMaxDepth = 64;       // BVHBuildConfig (Util/BVH.fs:99-107)
    const int n = (int)in.size();
    std::vector<BuildItem> items(n), scratch(n);
    for (int i = 0; i < n; ++i) items[i] = {in[i], i};
    nodes.clear();
    nodes.reserve(2 * (size_t)n);
    // Depth-first with an explicit stack: the left task is popped first, so a
    // node's index is assigned in pre-order exactly like Tree.Flatten.
    std::vector<Task> stack;
    stack.push_back({0, n, 0, -1});
    while (!stack.empty()) {
        Task t = stack.back();
        stack.pop_back();
        const int me = (int)nodes.size();
        if (t.parent >= 0) nodes[t.parent].right_or_offset = me;
        const int cnt = t.last - t.first;
        Box bounds = Box::empty();
        for (int i = t.first; i < t.last; ++i) bounds.grow(items[i].b);
        if (cnt <= kLeaf || t.depth >= kMaxDepth) {
            if (cnt > 127) throw Error(BN_E_CAPACITY, "BVH leaf larger than the int8 count field");
            nodes.push_back(make_node(bounds, t.first, true, 0, cnt));
            continue;
        }
        Box cb = Box::empty();
        for (int i = t.first; i < t.last; ++i) cb.grow(items[i].b.center());
        const int axis = cb.longest_axis();
        const float extent = cb.extent()[axis];
        int split_first;  // first index of the right range
        int right_first;
        if (extent == 0.0f) {
            split_first = t.first + cnt / 2;
            right_first = split_first;
        } else {
            const float lo = cb.lo[axis];
            // A NaN among the centroids can leave `lo` above other centroids,
            // so the index can be negative; the reference's bounds-checked
            // Span indexer throws there (Util/BVH.fs:163-174), and so does this.
            auto bin_of = [&](const Box& b) {
                const int k = std::min(kBins - 1, sat_int((12.0f * (b.center()[axis] - lo)) / extent));
                if (k < 0) throw Error(BN_E_INVALID, "BVH build: bin index out of range (non-finite bounds)");
                return k;
            };
            Box bb[kBins];
            int bc[kBins] = {};
            for (int k = 0; k < kBins; ++k) bb[k] = Box::empty();
            for (int i = t.first; i < t.last; ++i) {
                int k = bin_of(items[i].b);
                bb[k].grow(items[i].b);
                bc[k] += 1;
            }
            // Prefix (left) and suffix (right) sweeps; cost[s] = L(0..s) + R(s+1..11).
            float lcost[kBins - 1], rcost[kBins - 1];
            Box acc = Box::empty();
            int c = 0;
            for (int s = 0; s < kBins - 1; ++s) {
                acc.grow(bb[s]);
                c += bc[s];
                lcost[s] = (float)c * acc.half_area_x2();
            }
            acc = Box::empty();
            c = 0;
            for (int s = kBins - 2; s >= 0; --s) {
                acc.grow(bb[s + 1]);
                c += bc[s + 1];
                rcost[s] = (float)c * acc.half_area_x2();
            }
            const float base = (float)cnt * cb.half_area_x2();
            float best_cost = kInf;
            int best = 0;
            for (int s = 0; s < kBins - 1; ++s) {
                float cost = (lcost[s] + rcost[s]) + base;
                if (cost < best_cost) { best_cost = cost; best = s; }
            }
            // Stable partition to the left, reversed fill from the end on the right.
            std::copy(items.begin() + t.first, items.begin() + t.last, scratch.begin() + t.first);
            int l = t.first, r = t.last - 1;
            for (int i = t.first; i < t.last; ++i) {
                if (bin_of(scratch[i].b) > best) items[r--] = scratch[i];
                else items[l++] = scratch[i];
            }
            split_first = l;
            right_first = r + 1;
        }
        nodes.push_back(make_node(bounds, 0, false, axis, 0));
        stack.push_back({right_first, t.last, t.depth + 1, me});
        stack.push_back({t.first, split_first, t.depth + 1, -1});
    }
    perm.resize(n);
    for (int i = 0; i < n; ++i) perm[i] = items[i].id;
}

AliasTable build_alias(const float* w, int n) {
    AliasTable t;
    t.alias.resize(n);
    t.prob.resize(n);
    t.pdf.resize(n);
    float total = 0.0f;
    for (int i = 0; i < n; ++i) total = total + w[i];
    if (total == 0.0f) {
        for (int i = 0; i < n; ++i) { t.alias[i] = i; t.prob[i] = 1.0f; t.pdf[i] = 1.0f / (float)n; }
        return t;
    }
    std::vector<int32_t> small, large;
    for (int i = 0; i < n; ++i) {
        t.alias[i] = i;
        t.prob[i] = w[i] / total;
        t.pdf[i] = w[i] / total;
        if (t.prob[i] > 1.0f) large.push_back(i); else small.push_back(i);
    }
    while (!small.empty() && !large.empty()) {
        int32_t o = large.back(); large.pop_back();
        int32_t u = small.back(); small.pop_back();
        t.prob[o] = t.prob[o] + t.prob[u] - 1.0f;
        t.alias[u] = o;
        if (t.prob[o] > 1.0f) large.push_back(o); else small.push_back(o);
    }
    for (int32_t o : large) { t.alias[o] = o; t.prob[o] = 1.0f; }
    for (int32_t u : small) { t.alias[u] = u; t.prob[u] = 1.0f; }
    return t;
}

// ---------------------------------------------------------------- scene
void build_scene(const bn_scene_desc& d, HostScene& S, const BatchBvhBuild* blas_build) {
    static_assert(sizeof(Box) == 6 * sizeof(float), "Box is six packed floats");
    S = HostScene{};
    S.camera = d.camera;
    for (int i = 0; i < d.n_materials; ++i) S.materials.push_back(d.materials[i]);
    for (int i = 0; i < d.n_lights; ++i) S.light_defs.push_back(d.lights[i]);
    S.meshes.resize(d.n_primitives);
    std::vector<std::vector<Box>> tri_boxes(d.n_primitives);
    std::vector<std::vector<int32_t>> perms(d.n_primitives);
    std::vector<int> deferred;                                        // meshes for blas_build
    for (int pi = 0; pi < d.n_primitives; ++pi) {
        const bn_primitive& p = d.primitives[pi];
        HostMesh& m = S.meshes[pi];
        if (p.kind == BN_PRIM_SPHERE) { m.sphere = true; m.radius = p.radius; continue; }
        if (p.kind != BN_PRIM_MESH) throw Error(BN_E_INVALID, "unknown primitive kind");
        if (p.n_indices < 3 || p.n_indices % 3) throw Error(BN_E_INVALID, "mesh without triangles");
        if ((int64_t)p.vert_offset + p.n_verts > d.n_vertex_floats / 3 ||
            (int64_t)p.index_offset + p.n_indices > d.n_indices)
            throw Error(BN_E_INVALID, "mesh range outside the vertex/index pools");
        m.verts.resize(p.n_verts);
        for (int k = 0; k < p.n_verts; ++k) {
            const float* v = d.vertices + 3 * ((int64_t)p.vert_offset + k);
            m.verts[k] = {v[0], v[1], v[2]};
        }
        const int32_t* ix = d.indices + p.index_offset;
        const int nt = p.n_indices / 3;
        std::vector<Box>& tb = tri_boxes[pi];
        tb.resize(nt);
        bool finite = true;
        for (int k = 0; k < nt; ++k) {
            for (int j = 0; j < 3; ++j)
                if (ix[3 * k + j] < 0 || ix[3 * k + j] >= p.n_verts) throw Error(BN_E_INVALID, "mesh index out of range");
            V3 a = m.verts[ix[3 * k]], b = m.verts[ix[3 * k + 1]], c = m.verts[ix[3 * k + 2]];
            tb[k] = {vmin(vmin(a, b), c), vmax(vmax(a, b), c)};        // Mesh.fs:162-168
            finite = finite && std::isfinite(tb[k].lo.x) && std::isfinite(tb[k].lo.y) && std::isfinite(tb[k].lo.z) &&
                     std::isfinite(tb[k].hi.x) && std::isfinite(tb[k].hi.y) && std::isfinite(tb[k].hi.z);
        }
        if (blas_build && finite) deferred.push_back(pi);
    }
    if (!deferred.empty()) {
        std::vector<const float*> sets;
        std::vector<int> counts;
        for (int pi : deferred) {
            sets.push_back(reinterpret_cast<const float*>(tri_boxes[pi].data()));
            counts.push_back((int)tri_boxes[pi].size());
        }
        std::vector<std::vector<bn_bvh_node>> nodes;
        std::vector<std::vector<int32_t>> built;
        (*blas_build)(sets, counts, nodes, built);
        for (size_t k = 0; k < deferred.size(); ++k) {
            S.meshes[deferred[k]].nodes = std::move(nodes[k]);
            perms[deferred[k]] = std::move(built[k]);
        }
    }
    for (int pi = 0; pi < d.n_primitives; ++pi) {
        const bn_primitive& p = d.primitives[pi];
        HostMesh& m = S.meshes[pi];
        if (m.sphere) continue;
        const int32_t* ix = d.indices + p.index_offset;
        const int nt = p.n_indices / 3;
        if (perms[pi].empty()) build_bvh(tri_boxes[pi], m.nodes, perms[pi]);
        std::vector<Box>().swap(tri_boxes[pi]);
        const std::vector<int32_t>& perm = perms[pi];
        m.tri.resize(3 * (size_t)nt);
        for (int k = 0; k < nt; ++k)
            for (int j = 0; j < 3; ++j) m.tri[3 * k + j] = ix[3 * perm[k] + j];
        std::vector<float> area(nt);
        for (int k = 0; k < nt; ++k)
            area[k] = tri_area(m.verts[m.tri[3 * k]], m.verts[m.tri[3 * k + 1]], m.verts[m.tri[3 * k + 2]]);
        m.alias = build_alias(area.data(), nt);
    }
    std::vector<M4> local(d.n_transforms);
    // A transform is a run of key frame records sorted by time (Loader.fs:47);
    // nodes reference its first record, whose n_keys gives the run length.
    for (int i = 0; i < d.n_transforms; ++i) {
        const int nk = d.transforms[i].n_keys > 0 ? d.transforms[i].n_keys : 1;
        if (i + nk > d.n_transforms) throw Error(BN_E_INVALID, "key frame run past the transform array");
        local[i] = transform_eval(d.transforms + i, nk, d.time);
    }

    // Node.Traverse: DFS, world = local * parent.  An instance object listed
    // by several nodes keeps the transform of its last visit for every entry.
    std::vector<int32_t> order;
    std::vector<M4> inst_xf(d.n_instances, M4::identity());
    M4 c2w{};
    struct Frame { int node; M4 parent; int depth; };
    std::vector<Frame> work;
    work.push_back({d.root, M4::identity(), 0});
    int64_t visits = 0;
    while (!work.empty()) {
        Frame f = work.back();
        work.pop_back();
        if (f.node < 0 || f.node >= d.n_nodes) throw Error(BN_E_INVALID, "node index out of range");
        if (f.depth > 1000 || ++visits > 10000000) throw Error(BN_E_INVALID, "node graph has a cycle");
        const bn_node& n = d.nodes[f.node];
        const M4 w = (n.transform >= 0 ? local[n.transform] : M4::identity()) * f.parent;
        if (n.has_camera) c2w = w;
        for (int k = 0; k < n.n_inst; ++k) {
            int32_t ii = d.node_refs[n.inst_offset + k];
            if (ii < 0 || ii >= d.n_instances) throw Error(BN_E_INVALID, "instance index out of range");
            inst_xf[ii] = w;
            order.push_back(ii);
        }
        for (int k = n.n_child - 1; k >= 0; --k) work.push_back({d.node_refs[n.child_offset + k], w, f.depth + 1});
    }
    S.c2w = c2w;
    if (order.empty()) throw Error(BN_E_INVALID, "scene has no primitive instances");
    std::vector<HostInstance> list;
    std::vector<Box> boxes;
    for (int32_t ii : order) {
        const bn_instance& di = d.instances[ii];
        if (di.primitive < 0 || di.primitive >= d.n_primitives || di.material >= d.n_materials ||
            di.light >= d.n_lights)
            throw Error(BN_E_INVALID, "instance references an undefined primitive/material/light");
        HostInstance h;
        h.prim = di.primitive; h.material = di.material; h.light = di.light;
        h.o2w = inst_xf[ii];
        h.w2o = invert(h.o2w);                                      // Primitive.fs:131-138
        h.bounds = transform_box(S.meshes[h.prim].bounds(), h.o2w);
        list.push_back(h);
        boxes.push_back(h.bounds);
    }
    std::vector<int32_t> perm;
    build_bvh(boxes, S.tlas, perm);                                 // Aggregate/BVH.fs:9
    for (int32_t p : perm) { S.insts.push_back(list[p]); S.inst_desc.push_back(order[p]); }
    for (int i = 0; i < (int)S.insts.size(); ++i)
        if (S.insts[i].light >= 0) S.lights.push_back(i);           // LightSampler.fs:7
}

}  // namespace bnh
