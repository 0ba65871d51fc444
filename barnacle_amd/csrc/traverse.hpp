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
// Three-way min / max in one instruction.  Used only on the finite path, where
// they agree with the compare-and-select Vector3.Min/Max up to the sign of a
// zero, which no slab comparison can see.
__device__ __forceinline__ float min3f(float a, float b, float c) {
    float r;
    asm volatile("v_min3_f32 %0, %1, %2, %3" : "=v"(r) : "v"(a), "v"(b), "v"(c));
    return r;
}
__device__ __forceinline__ float max3f(float a, float b, float c) {
    float r;
    asm volatile("v_max3_f32 %0, %1, %2, %3" : "=v"(r) : "v"(a), "v"(b), "v"(c));
    return r;
}
__device__ __forceinline__ bool finite3(V3 v) {
    return __builtin_isfinite(v.x) && __builtin_isfinite(v.y) && __builtin_isfinite(v.z);
}

enum : int { kItemNone = 0, kItemTNode = 1, kItemBNode = 2, kItemTri = 3, kItemInst = 4 };
constexpr uint32_t kNoNode = 0xffffffffu;
constexpr uint32_t kNeedTest = 0x80000000u;   // node whose box is still to be tested
constexpr uint32_t kNodeMask = 0x7fffffffu;

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
    uint32_t nxt;        // node to take before the stack (kNoNode: none); kNeedTest: box not tested yet
    int blas_base;       // stack depth at BLAS entry, -1 in the TLAS
    int tl_cur, tl_end;  // instances left in the current TLAS leaf
    int bl_cur, bl_end;  // triangles left in the current BLAS leaf (global indices)
    int blas_root, tri_base, cur_inst;

    // A new ray over the geometry already in G (set once per launch for a
    // single scene, per ray in a batch).
    __device__ __forceinline__ void start(const Ray& r, float tmax, bool exact_scene) {
        wo = r.o; wd = r.d;
        winv = {1.0f / r.d.x, 1.0f / r.d.y, 1.0f / r.d.z};
        co = wo; cd = wd; cinv = winv;
        t = tmax;
        hit_inst = -1; hit_prim = -1;
        exact = exact_scene || !finite3(wo) || !finite3(winv);
        st.sp = 0;
        nxt = kNeedTest;                 // the TLAS root
        blas_base = -1;
        tl_cur = tl_end = 0;
        bl_cur = bl_end = 0;
        blas_root = tri_base = cur_inst = 0;
    }

    // No traversal at all (placeholder rays, NaN t max).
    __device__ __forceinline__ void empty() {
        st.sp = 0;
        nxt = kNoNode;
        tl_cur = tl_end = 0;
    }
    // The fast traversal tests a root when it is reached: the TLAS root when
    // the ray starts, a BLAS root when its instance is entered.  A failing
    // root leaves nothing to do, a leaf root hands over its items at once,
    // and an interior root is expanded next without a second test.  (The
    // counting traversal leaves the TLAS root to step(), which tests and
    // counts it.)
    __device__ __forceinline__ void take_tlas_root() { take_root(G.tlas, false); }
    __device__ __forceinline__ void take_root(const float4* base, bool in_blas) {
        const float4 a = base[0], b = base[1];
        nxt = kNoNode;
        if (!box(mk(a.x, a.y, a.z), mk(a.w, b.x, b.y))) return;
        const uint32_t f = __float_as_uint(b.w);
        if (f & 0xff) {
            const int cnt = (int)(int8_t)((f >> 16) & 0xff);
            const int w = __float_as_int(b.z);
            if (in_blas) { bl_cur = tri_base + w; bl_end = bl_cur + cnt; }
            else { tl_cur = w; tl_end = w + cnt; }
        } else {
            nxt = 0u;
        }
    }
    // An idle lane: no item left anywhere, so next_kind() reports kItemNone
    // and a step does nothing.
    __device__ __forceinline__ void clear() {
        st.sp = 0;
        nxt = kNoNode;
        blas_base = -1;
        tl_cur = tl_end = 0;
        bl_cur = bl_end = 0;
    }

    // The fast test is evaluated first, in the block that computes its
    // products, so its min/max need no operand canonicalisation; the exact
    // test replaces it on the rare non-finite rays.
    __device__ __forceinline__ bool box(V3 lo, V3 hi) const {
        bool r = slab_fast(lo, hi, co, cinv, t);
        if (exact) r = slab(lo, hi, co, cinv, t);
        return r;
    }
    // A triangle's bounding-box pre-test (Mesh.fs:204).
    __device__ __forceinline__ bool box_tri(const Tri& tr) const {
        if (!exact)
            return slab_fast(mk(min3f(tr.p0.x, tr.p1.x, tr.p2.x), min3f(tr.p0.y, tr.p1.y, tr.p2.y),
                                min3f(tr.p0.z, tr.p1.z, tr.p2.z)),
                             mk(max3f(tr.p0.x, tr.p1.x, tr.p2.x), max3f(tr.p0.y, tr.p1.y, tr.p2.y),
                                max3f(tr.p0.z, tr.p1.z, tr.p2.z)),
                             co, cinv, t);
        return slab(vmin(vmin(tr.p0, tr.p1), tr.p2), vmax(vmax(tr.p0, tr.p1), tr.p2), co, cinv, t);
    }

    // One item.  Returns false when the traversal has finished (no items left,
    // or an any-hit query found its hit); `found` reports an any-hit.
    // The reference's visit sequence item by item (every popped node is
    // tested); the counting kernels use it, so the counters are the
    // reference's.
    template <bool ANY, bool COUNT>
    __device__ __forceinline__ bool step(TraceCount& cnt, bool& found) {
        // leave a finished BLAS
        if (blas_base >= 0 && bl_cur >= bl_end && nxt == kNoNode && st.sp == blas_base) {
            blas_base = -1;
            co = wo; cd = wd; cinv = winv;
        }
        int kind;
        int idx = 0;
        if (blas_base >= 0) {
            if (bl_cur < bl_end) { kind = kItemTri; idx = bl_cur++; }
            else if (nxt != kNoNode) { kind = kItemBNode; idx = (int)(nxt & kNodeMask); nxt = kNoNode; }
            else { kind = kItemBNode; idx = (int)st.pop(); }
        } else if (tl_cur < tl_end) {
            kind = kItemInst; idx = tl_cur++;
        } else if (nxt != kNoNode) {
            kind = kItemTNode; idx = (int)(nxt & kNodeMask); nxt = kNoNode;
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
                nxt = kNeedTest;         // the BLAS root
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

    // The fast traversal (do_tri / do_inst / do_node): the same hits with
    // fewer iterations.  An interior node tests both children when it is
    // expanded, with the t of that moment, and only the children that pass
    // are kept.  A kept child is not re-tested when it is taken later with a
    // smaller t.  That changes no hit:
    // if its box would fail then, the box lies beyond the current t, and so do
    // the boxes of every triangle and instance below it, which are still
    // tested (Mesh.fs:204,231 pre-test every triangle; instances have their
    // world bounds); a box that passes with the current t passed with the
    // larger one too.  Every triangle the reference tests is tested, in the
    // same order, so ties resolve alike.
    // What this lane does next (leaving a finished BLAS first): kItemTri,
    // kItemInst, kItemTNode (any node), or kItemNone when it has finished.
    __device__ __forceinline__ int next_kind() {
        // Selects rather than a branch: the ray registers are then updated in
        // place instead of through copies at the join.
        const bool leave = blas_base >= 0 && bl_cur >= bl_end && nxt == kNoNode && st.sp == blas_base;
        blas_base = leave ? -1 : blas_base;
        co = sel(leave, wo, co);
        cd = sel(leave, wd, cd);
        cinv = sel(leave, winv, cinv);
        if (blas_base >= 0 && bl_cur < bl_end) return kItemTri;
        if (blas_base < 0 && tl_cur < tl_end) return kItemInst;
        if (nxt != kNoNode || st.sp > (blas_base >= 0 ? blas_base : 0)) return kItemTNode;
        return kItemNone;
    }

    // One triangle of the current BLAS leaf.  Returns false on an any-hit.
    template <bool ANY>
    __device__ __forceinline__ bool do_tri(bool& found) {
        const int idx = bl_cur++;
        const Tri tr = load_tri(G.tris, idx);
        if (!box_tri(tr)) return true;
        float tt;
        if (tri_hit(tr, co, cd, t, tt)) {
            if (ANY) { found = true; return false; }
            t = tt;
            hit_inst = cur_inst;
            hit_prim = idx - tri_base;
        }
        return true;
    }

    // One instance of the current TLAS leaf: a sphere is intersected, a mesh
    // BLAS is entered.  Returns false on an any-hit.
    template <bool ANY>
    __device__ __forceinline__ bool do_inst(bool& found) {
        const int idx = tl_cur++;
        const DevInst& I = G.insts[idx];
        if (!box(xyz(I.lo), xyz(I.hi))) return true;
        const int4 info = I.info;
        if (info.x >= 0 && I.aux.x != 0 && !exact) {
            // Identity world-to-object map on a finite ray: Transform gives
            // o + 0 per coordinate and d unchanged (no component of d is 0
            // when 1/d is finite), so the object ray's 1/d is the world one.
            co = mk(wo.x + 0.0f, wo.y + 0.0f, wo.z + 0.0f);
            cd = wd;
            cinv = winv;
            enter_blas(idx, info);
            return true;
        }
        const Xf w2o = I.w2o;
        const V3 oo = xf_point(w2o, wo), od = xf_dir(w2o, wd);
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
            enter_blas(idx, info);
        }
        return true;
    }

    // Entry into the BLAS of instance idx with the object ray set; its root
    // is tested at once, as the reference pops and tests it first
    // (Mesh.fs:217-242).  A lane whose root fails leaves the BLAS in its next
    // step.
    __device__ __forceinline__ void enter_blas(int idx, int4 info) {
        cur_inst = idx;
        blas_root = info.x;
        tri_base = info.y;
        blas_base = st.sp;
        take_root(G.blas + 2 * (size_t)info.x, true);
    }

    // One item of the fast traversal (all lanes of a wave advance together,
    // whatever kind of item each has: measured faster than advancing only the
    // lanes of the most common kind).  Returns false when the traversal has
    // finished or an any-hit query found its hit.
    template <bool ANY>
    __device__ __forceinline__ bool step_fast(bool& found) {
        const int kind = next_kind();
        if (kind == kItemTri) return do_tri<ANY>(found);
        if (kind == kItemInst) return do_inst<ANY>(found);
        if (kind == kItemTNode) { do_node(); return true; }
        return false;
    }

    // Expand the next node: a leaf sets its item range; an interior node
    // tests both children with the current t and keeps those that pass (the
    // near one in `nxt`, the far one on the stack, near first as in
    // Extensions/Aggregate/BVH.fs:47-56 and Mesh.fs:233-240).
    __device__ __forceinline__ void do_node() {
        uint32_t e;
        if (nxt != kNoNode) { e = nxt; nxt = kNoNode; }
        else e = st.pop();
        const float4* base = blas_base >= 0 ? G.blas + 2 * (size_t)blas_root : G.tlas;
        const int node = (int)e;
        const float4 b = base[2 * node + 1];
        const int word = __float_as_int(b.z);
        const uint32_t f = __float_as_uint(b.w);
        if (f & 0xff) {
            const int count = (int)(int8_t)((f >> 16) & 0xff);
            if (blas_base < 0) { tl_cur = word; tl_end = word + count; }
            else { bl_cur = tri_base + word; bl_end = bl_cur + count; }
            return;
        }
        const int axis = (int)(int8_t)((f >> 8) & 0xff);
        const int l = node + 1, r = word;
        const float4 la = base[2 * l], lb = base[2 * l + 1], ra = base[2 * r], rb = base[2 * r + 1];
        const bool hl = box(mk(la.x, la.y, la.z), mk(la.w, lb.x, lb.y));
        const bool hr = box(mk(ra.x, ra.y, ra.z), mk(ra.w, rb.x, rb.y));
        const bool left_near = comp(cd, axis) > 0.0f;
        const int near = left_near ? l : r, far = left_near ? r : l;
        const bool hn = left_near ? hl : hr, hf = left_near ? hr : hl;
        if (hn && hf) st.push((uint32_t)far);
        if (hn || hf) {
            // The child taken next: a leaf hands over its items at once, which
            // saves the step that would only pop it (same items, same order).
            const bool take_left = hn ? left_near : !left_near;
            const float4 tb = take_left ? lb : rb;
            const uint32_t tf = __float_as_uint(tb.w);
            if (tf & 0xff) {
                const int cnt = (int)(int8_t)((tf >> 16) & 0xff);
                const int w = __float_as_int(tb.z);
                if (blas_base < 0) { tl_cur = w; tl_end = w + cnt; }
                else { bl_cur = tri_base + w; bl_end = bl_cur + cnt; }
            } else {
                nxt = (uint32_t)(take_left ? l : r);
            }
        }
    }
};

}  // namespace bnd
