// traverse.hpp -- persistent, lane-refilling BVH traversal for gfx950.
//
// Semantics are those of BVHAggregate.Intersect (Extensions/Aggregate/BVH.fs:
// 11-58) over PrimitiveInstance.Intersect (Base/Primitive.fs:111-129) and
// MeshPrimitive.Intersect (Extensions/Primitive/Mesh.fs:188-242): the same
// items are visited in the same order with the same current t, so hit
// (instance, triangle, t) and the traversal counters are identical to the
// reference's.  The schedule differs: every loop iteration processes exactly
// one item per lane, and an item is one of
//
//    TLAS node | BLAS node | triangle of a BLAS leaf | instance of a TLAS leaf
//
// all of which start with the same slab test (node bounds, triangle bounds,
// instance world bounds), so a wave executes one shared slab test per
// iteration instead of serialising whole leaves.  Lanes whose ray finished are
// refilled from the work queue in batches (one atomic per wave).
#pragma once
#include "device.hpp"

namespace bnd {

// Slab test with hardware min/max.  Equal in outcome to the exact test when
// the ray origin and 1/d are finite and the box is finite: then no t0/t1 is
// NaN, and min/max results can differ only in the sign of a zero, which the
// final comparison does not see.
__device__ __forceinline__ bool slab_fast(V3 lo, V3 hi, V3 o, V3 inv, float t) {
    const float t0x = (lo.x - o.x) * inv.x, t1x = (hi.x - o.x) * inv.x;
    const float t0y = (lo.y - o.y) * inv.y, t1y = (hi.y - o.y) * inv.y;
    const float t0z = (lo.z - o.z) * inv.z, t1z = (hi.z - o.z) * inv.z;
    const float tmin = fmaxf(fmaxf(1e-3f, fminf(t0x, t1x)), fmaxf(fminf(t0y, t1y), fminf(t0z, t1z)));
    const float tmax = fminf(fminf(t, fmaxf(t0x, t1x)), fminf(fmaxf(t0y, t1y), fmaxf(t0z, t1z)));
    return tmin <= tmax;
}
__device__ __forceinline__ bool finite3(V3 v) {
    return __builtin_isfinite(v.x) && __builtin_isfinite(v.y) && __builtin_isfinite(v.z);
}

enum : int { kItemNone = 0, kItemTNode = 1, kItemBNode = 2, kItemTri = 3, kItemInst = 4 };

// The geometry a lane traverses (per lane, so a batched launch can mix scenes).
struct TraceScene {
    const float4* tlas;
    const float4* blas;
    const float* tris;
    const DevInst* insts;
    __device__ __forceinline__ void set(const DevScene& S) { tlas = S.tlas; blas = S.blas; tris = S.tris; insts = S.insts; }
};

template <int K>
struct Traversal {
    TraceScene G;
    // world ray and the ray of the current space (world, or object inside a BLAS)
    V3 wo, wd, winv;
    V3 co, cd, cinv;
    float t;
    int hit_inst, hit_prim;
    bool exact;          // take the NaN-exact slab path
    // traversal position
    LaneStack<K> st;
    int blas_base;       // stack depth at BLAS entry, -1 in the TLAS
    int tl_cur, tl_end;  // instances left in the current TLAS leaf
    int bl_cur, bl_end;  // triangles left in the current BLAS leaf (global indices)
    int blas_root, tri_base, cur_inst;

    __device__ __forceinline__ void start(const DevScene& S, const Ray& r, float tmax) {
        G.set(S);
        const bool exact_scene = S.exact_only != 0;
        wo = r.o; wd = r.d;
        winv = {1.0f / r.d.x, 1.0f / r.d.y, 1.0f / r.d.z};
        co = wo; cd = wd; cinv = winv;
        t = tmax;
        hit_inst = -1; hit_prim = -1;
        exact = exact_scene || !finite3(wo) || !finite3(winv);
        st.sp = 0;
        st.push(0);
        blas_base = -1;
        tl_cur = tl_end = 0;
        bl_cur = bl_end = 0;
        blas_root = tri_base = cur_inst = 0;
    }

    __device__ __forceinline__ bool box(V3 lo, V3 hi) const {
        return exact ? slab(lo, hi, co, cinv, t) : slab_fast(lo, hi, co, cinv, t);
    }

    // One item.  Returns false when the traversal has finished (no items left,
    // or an any-hit query found its hit); `found` reports an any-hit.
    template <bool ANY, bool COUNT>
    __device__ __forceinline__ bool step(TraceCount& cnt, bool& found) {
        // leave a finished BLAS
        if (blas_base >= 0 && bl_cur >= bl_end && st.sp == blas_base) {
            blas_base = -1;
            co = wo; cd = wd; cinv = winv;
        }
        int kind;
        int idx = 0;
        if (blas_base >= 0) {
            if (bl_cur < bl_end) { kind = kItemTri; idx = bl_cur++; }
            else { kind = kItemBNode; idx = (int)st.pop(); }
        } else if (tl_cur < tl_end) {
            kind = kItemInst; idx = tl_cur++;
        } else if (st.sp > 0) {
            kind = kItemTNode; idx = (int)st.pop();
        } else {
            return false;
        }
        V3 lo, hi;
        Tri tr;
        NodeView n;
        if (kind == kItemTri) {
            tr = load_tri(G.tris, idx);
            lo = vmin(vmin(tr.p0, tr.p1), tr.p2);
            hi = vmax(vmax(tr.p0, tr.p1), tr.p2);
            if (COUNT) cnt.tris++;
        } else if (kind == kItemInst) {
            const float4 a = G.insts[idx].lo, b = G.insts[idx].hi;
            lo = xyz(a); hi = xyz(b);
        } else {
            n = load_node(kind == kItemTNode ? G.tlas : G.blas + 2 * (size_t)blas_root, idx);
            lo = n.lo; hi = n.hi;
            if (COUNT) cnt.nodes++;
        }
        if (!box(lo, hi)) return true;
        if (kind == kItemTri) {
            float tt;
            if (tri_hit(tr, co, cd, t, tt)) {
                if (ANY) { found = true; return false; }
                t = tt;
                hit_inst = cur_inst;
                hit_prim = idx - tri_base;
            }
        } else if (kind == kItemInst) {
            if (COUNT) cnt.insts++;
            const DevInst& I = G.insts[idx];
            const Xf w2o = I.w2o;
            const V3 oo = xf_point(w2o, wo), od = xf_dir(w2o, wd);
            const int4 info = I.info;
            if (info.x < 0) {
                float tt;
                int br;
                if (sphere_hit(I.lo.w, oo, od, t, tt, br)) {
                    if (ANY) { found = true; return false; }
                    t = tt;
                    hit_inst = idx;
                    hit_prim = br;
                }
            } else {
                co = oo; cd = od;
                cinv = {1.0f / od.x, 1.0f / od.y, 1.0f / od.z};
                exact = exact || !finite3(co) || !finite3(cinv);
                cur_inst = idx;
                blas_root = info.x;
                tri_base = info.y;
                blas_base = st.sp;
                st.push(0);
            }
        } else if (n.leaf) {
            if (kind == kItemTNode) { tl_cur = n.word; tl_end = n.word + n.count; }
            else { bl_cur = tri_base + n.word; bl_end = bl_cur + n.count; }
        } else if (comp(cd, n.axis) > 0.0f) {
            st.push((uint32_t)n.word); st.push((uint32_t)(idx + 1));
        } else {
            st.push((uint32_t)(idx + 1)); st.push((uint32_t)n.word);
        }
        return true;
    }
};

}  // namespace bnd
