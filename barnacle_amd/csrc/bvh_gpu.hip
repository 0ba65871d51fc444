// bvh_gpu.hip -- binned-SAH BVH construction on the GPU, node for node the
// build of BVHNode.Build (Util/BVH.fs:128-247): the same nodes in the same
// pre-order and the same item permutation as the recursive host build.
//
// What makes a parallel build exact:
//  * Every bound is a fold of MinNative/MaxNative over items in their order
//    (AABB.Union(acc, x): x wins ties, so of -0 and +0 the later one stays).
//    "Later wins ties" is associative, so the folds are evaluated as trees
//    that only ever join a range with the range right after it; with finite
//    inputs the result is bit-identical to the sequential fold.  Non-finite
//    bounds break associativity and are refused (the host build takes them).
//  * Bin counts, partition counts and positions are integers; the partition
//    is the reference's: left items in order from `first`, right items in
//    order from `last - 1` downwards (Util/BVH.fs:205-217).
//  * The SAH sweep of a node runs in one thread in the reference's order.
//
// Schedule: a task is a node's item range.  Tasks of more than `small` items
// run level by level over chunks of kChunk items (one block per chunk):
// bounds, bins and the scatter are chunk kernels, the per-node decisions one
// thread per task.  A task of at most `small` items (or at depth 64) is built
// whole by one wave, recursively with an LDS stack, into a node region of its
// own.  The host numbers the upper nodes and regions in pre-order, and a
// last pass writes the flattened nodes.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "abi_internal.hpp"
#include "bvh_gpu.hpp"

namespace bvhg {

constexpr int kBins = 12, kLeaf = 4, kMaxDepth = 64;        // BVHBuildConfig (Util/BVH.fs:99-107)
constexpr int kThreads = 256, kIpt = 8;
constexpr int kChunk = kThreads * kIpt;                     // items per chunk block
constexpr int kWavesPerBlock = kThreads / 64;
constexpr float kInf = __builtin_huge_valf();

// 32-byte item: bounds and the item's index in its set.
struct alignas(16) Item { float lx, ly, lz; int32_t id; float hx, hy, hz; int32_t pad; };
struct DBox { float l[3], h[3]; };
struct BigTask { int first, last, depth, unode, set, chunk0, nchunks, pad; };
struct BigInfo { int mode, axis, best, nl; float lo_axis, extent, pad0, pad1; DBox cb; };   // mode 0 bin, 1 middle
struct SmallTask { int first, last, depth, buf, set, pad0, pad1, pad2; };
struct UNode { DBox b; int axis, left, right, set; };      // child < 0: small task ~child
struct Ctl { int n_next_big, n_small, n_unodes, n_chunks, err, pad[3]; };

__device__ __forceinline__ float omin(float a, float b) { return a < b ? a : b; }
__device__ __forceinline__ float omax(float a, float b) { return a > b ? a : b; }
__device__ __forceinline__ DBox empty_box() { return {{kInf, kInf, kInf}, {-kInf, -kInf, -kInf}}; }
// a covers items before b
__device__ __forceinline__ DBox join(const DBox& a, const DBox& b) {
    DBox r;
#pragma unroll
    for (int k = 0; k < 3; ++k) { r.l[k] = omin(a.l[k], b.l[k]); r.h[k] = omax(a.h[k], b.h[k]); }
    return r;
}
__device__ __forceinline__ Item load_item(const Item* p) {
    const float4 a = reinterpret_cast<const float4*>(p)[0], b = reinterpret_cast<const float4*>(p)[1];
    Item it;
    it.lx = a.x; it.ly = a.y; it.lz = a.z; it.id = __float_as_int(a.w);
    it.hx = b.x; it.hy = b.y; it.hz = b.z; it.pad = 0;
    return it;
}
__device__ __forceinline__ void store_item(Item* p, const Item& it) {
    reinterpret_cast<float4*>(p)[0] = make_float4(it.lx, it.ly, it.lz, __int_as_float(it.id));
    reinterpret_cast<float4*>(p)[1] = make_float4(it.hx, it.hy, it.hz, 0.0f);
}
__device__ __forceinline__ DBox item_box(const Item& it) { return {{it.lx, it.ly, it.lz}, {it.hx, it.hy, it.hz}}; }
// AABB.Centroid = 0.5f * (pMin + pMax), as a point box for Union(acc, point)
__device__ __forceinline__ DBox centroid_box(const Item& it) {
    const float cx = 0.5f * (it.lx + it.hx), cy = 0.5f * (it.ly + it.hy), cz = 0.5f * (it.lz + it.hz);
    return {{cx, cy, cz}, {cx, cy, cz}};
}
__device__ __forceinline__ float centroid_on(const Item& it, int axis) {
    return axis == 0 ? 0.5f * (it.lx + it.hx) : axis == 1 ? 0.5f * (it.ly + it.hy) : 0.5f * (it.lz + it.hz);
}
// findBinIndex (Util/BVH.fs:163-168): int of (12 * (c - lo)) / extent, at most 11
__device__ __forceinline__ int bin_of(const Item& it, int axis, float lo, float extent) {
    const float v = (12.0f * (centroid_on(it, axis) - lo)) / extent;
    return min(kBins - 1, (int)v);
}
// AABB.SurfaceArea: 2 * (dx dy + dy dz + dz dx)
__device__ __forceinline__ float area(const DBox& b) {
    const float dx = b.h[0] - b.l[0], dy = b.h[1] - b.l[1], dz = b.h[2] - b.l[2];
    return 2.0f * (dx * dy + dy * dz + dz * dx);
}
// AABB.SplitAxis of the centroid bounds
__device__ __forceinline__ int split_axis(const DBox& b) {
    const float dx = b.h[0] - b.l[0], dy = b.h[1] - b.l[1], dz = b.h[2] - b.l[2];
    if (dx >= dy && dx >= dz) return 0;
    return dy >= dz ? 1 : 2;
}

// Ordered union over the lanes of a wave (lane i's items precede lane i+1's);
// the result is broadcast to every lane.
__device__ DBox wave_union(DBox v) {
#pragma unroll
    for (int s = 1; s < 64; s <<= 1) {
        DBox o;
#pragma unroll
        for (int k = 0; k < 3; ++k) { o.l[k] = __shfl_down(v.l[k], s); o.h[k] = __shfl_down(v.h[k], s); }
        v = join(v, o);                       // meaningful in lanes that are multiples of 2s
    }
#pragma unroll
    for (int k = 0; k < 3; ++k) { v.l[k] = __shfl(v.l[k], 0); v.h[k] = __shfl(v.h[k], 0); }
    return v;
}
__device__ __forceinline__ int wave_sum(int x) {
#pragma unroll
    for (int s = 32; s >= 1; s >>= 1) x += __shfl_xor(x, s);
    return x;
}
__device__ __forceinline__ int wave_excl_scan(int x) {
    const int lane = threadIdx.x & 63;
    int v = x;
#pragma unroll
    for (int s = 1; s < 64; s <<= 1) {
        const int y = __shfl_up(v, s);
        if (lane >= s) v += y;
    }
    return v - x;
}
// Ordered union over a block (waves in order); every thread gets the result.
__device__ DBox block_union(DBox v, DBox* sh) {
    v = wave_union(v);
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) sh[w] = v;
    __syncthreads();
    DBox r = sh[0];
    for (int k = 1; k < kWavesPerBlock; ++k) r = join(r, sh[k]);
    __syncthreads();
    return r;
}
__device__ int block_sum(int x, int* sh) {
    x = wave_sum(x);
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) sh[w] = x;
    __syncthreads();
    int r = 0;
    for (int k = 0; k < kWavesPerBlock; ++k) r += sh[k];
    __syncthreads();
    return r;
}
__device__ int block_excl_scan(int x, int* sh) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    int v = x;
#pragma unroll
    for (int s = 1; s < 64; s <<= 1) {
        const int y = __shfl_up(v, s);
        if (lane >= s) v += y;
    }
    if (lane == 63) sh[w] = v;
    __syncthreads();
    int off = 0;
    for (int k = 0; k < w; ++k) off += sh[k];
    __syncthreads();
    return off + v - x;
}

struct Bufs {
    Item* items[2];
    const BigTask* big;
    BigInfo* info;
    BigTask* next_big;
    SmallTask* small;
    const int* chunk_task;
    const int* chunk_start;
    int* chunk_task_next;
    int* chunk_start_next;
    DBox* part_b;       // per chunk: item bounds
    DBox* part_cb;      // per chunk: centroid bounds
    DBox* bin_b;        // per chunk x bin
    int* bin_n;
    int* chunk_lbase;
    int* chunk_rbase;
    UNode* unodes;
    bn_bvh_node* region; // 2 * first: the region of the small task starting at `first`
    int* small_size;     // nodes of each small task's region
    int32_t* perm;       // per position: item id
    const int* set_off;
    Ctl* ctl;
    int big_cap, small_cap, unode_cap, chunk_cap;
};

// The thread's items of a chunk: [s + t * kIpt, ...) within [s, e).
__device__ __forceinline__ void thread_range(int s, int e, int& b, int& en) {
    b = min(s + (int)threadIdx.x * kIpt, e);
    en = min(b + kIpt, e);
}

// K1: item bounds and centroid bounds of every chunk.
__global__ void __launch_bounds__(kThreads) k_chunk_bounds(Bufs B, int src) {
    __shared__ DBox sh[kWavesPerBlock];
    const int c = blockIdx.x;
    const BigTask t = B.big[B.chunk_task[c]];
    const int s = B.chunk_start[c], e = min(s + kChunk, t.last);
    int b, en;
    thread_range(s, e, b, en);
    DBox bb = empty_box(), cb = empty_box();
    for (int i = b; i < en; ++i) {
        const Item it = load_item(B.items[src] + i);
        bb = join(bb, item_box(it));
        cb = join(cb, centroid_box(it));
    }
    bb = block_union(bb, sh);
    cb = block_union(cb, sh);
    if (threadIdx.x == 0) { B.part_b[c] = bb; B.part_cb[c] = cb; }
}

// K2: per task: node bounds, centroid bounds, split axis and mode.
__global__ void k_task_bounds(Bufs B, int n_big) {
    const int ti = blockIdx.x * blockDim.x + threadIdx.x;
    if (ti >= n_big) return;
    const BigTask t = B.big[ti];
    DBox bb = empty_box(), cb = empty_box();
    for (int c = t.chunk0; c < t.chunk0 + t.nchunks; ++c) { bb = join(bb, B.part_b[c]); cb = join(cb, B.part_cb[c]); }
    const int axis = split_axis(cb);
    const float extent = cb.h[axis] - cb.l[axis];
    BigInfo inf{};
    inf.mode = extent == 0.0f ? 1 : 0;
    inf.axis = axis;
    inf.lo_axis = cb.l[axis];
    inf.extent = extent;
    inf.cb = cb;
    B.info[ti] = inf;
    UNode& u = B.unodes[t.unode];
    u.b = bb;
    u.axis = axis;
    u.set = t.set;
}

// K3: per chunk of a binned task: each bin's bounds (ordered) and count.
__global__ void __launch_bounds__(kThreads) k_chunk_bins(Bufs B, int src) {
    __shared__ DBox sh[kWavesPerBlock];
    __shared__ int shn[kWavesPerBlock];
    const int c = blockIdx.x;
    const int ti = B.chunk_task[c];
    const BigInfo inf = B.info[ti];
    if (inf.mode != 0) return;                    // uniform per block
    const BigTask t = B.big[ti];
    const int s = B.chunk_start[c], e = min(s + kChunk, t.last);
    int b, en;
    thread_range(s, e, b, en);
    Item its[kIpt];
    int bins[kIpt];
#pragma unroll
    for (int j = 0; j < kIpt; ++j) {
        if (b + j < en) {
            its[j] = load_item(B.items[src] + b + j);
            bins[j] = bin_of(its[j], inf.axis, inf.lo_axis, inf.extent);
        } else {
            bins[j] = -1;
        }
    }
    for (int k = 0; k < kBins; ++k) {
        DBox bb = empty_box();
        int n = 0;
#pragma unroll
        for (int j = 0; j < kIpt; ++j)
            if (bins[j] == k) { bb = join(bb, item_box(its[j])); ++n; }
        bb = block_union(bb, sh);
        n = block_sum(n, shn);
        if (threadIdx.x == 0) { B.bin_b[(size_t)c * kBins + k] = bb; B.bin_n[(size_t)c * kBins + k] = n; }
    }
}

// K4: per task: SAH sweep (Util/BVH.fs:183-203), chunk scatter bases and the
// two children, each either a task of the next level or a small task.
__global__ void k_task_split(Bufs B, int n_big, int level, int small) {
    const int ti = blockIdx.x * blockDim.x + threadIdx.x;
    if (ti >= n_big) return;
    const BigTask t = B.big[ti];
    BigInfo inf = B.info[ti];
    const int cnt = t.last - t.first;
    int nl;
    if (inf.mode == 0) {
        DBox bb[kBins];
        int bn[kBins];
        for (int k = 0; k < kBins; ++k) { bb[k] = empty_box(); bn[k] = 0; }
        for (int c = t.chunk0; c < t.chunk0 + t.nchunks; ++c)
            for (int k = 0; k < kBins; ++k) {
                bb[k] = join(bb[k], B.bin_b[(size_t)c * kBins + k]);
                bn[k] += B.bin_n[(size_t)c * kBins + k];
            }
        float lcost[kBins - 1], rcost[kBins - 1];
        DBox acc = empty_box();
        int n = 0;
        for (int s = 0; s < kBins - 1; ++s) {
            acc = join(acc, bb[s]);
            n += bn[s];
            lcost[s] = (float)n * area(acc);
        }
        acc = empty_box();
        n = 0;
        for (int s = kBins - 2; s >= 0; --s) {
            acc = join(acc, bb[s + 1]);
            n += bn[s + 1];
            rcost[s] = (float)n * area(acc);
        }
        const float base = (float)cnt * area(inf.cb);
        float best_cost = kInf;
        int best = 0;
        for (int s = 0; s < kBins - 1; ++s) {
            const float cost = (lcost[s] + rcost[s]) + base;
            if (cost < best_cost) { best_cost = cost; best = s; }
        }
        nl = 0;
        for (int k = 0; k <= best; ++k) nl += bn[k];
        // scatter bases: lefts ascend from first, rights descend from last - 1
        int lo = t.first, hi = t.last - 1;
        for (int c = t.chunk0; c < t.chunk0 + t.nchunks; ++c) {
            int l = 0, r = 0;
            for (int k = 0; k < kBins; ++k) (k <= best ? l : r) += B.bin_n[(size_t)c * kBins + k];
            B.chunk_lbase[c] = lo;
            B.chunk_rbase[c] = hi;
            lo += l;
            hi -= r;
        }
        inf.best = best;
    } else {
        nl = cnt / 2;                                          // equal centroids: split in the middle
    }
    inf.nl = nl;
    B.info[ti] = inf;
    const int ranges[2][2] = {{t.first, t.first + nl}, {t.first + nl, t.last}};
    int code[2];
    for (int side = 0; side < 2; ++side) {
        const int f = ranges[side][0], l = ranges[side][1];
        if (l - f > small && t.depth + 1 < kMaxDepth) {
            const int u = atomicAdd(&B.ctl->n_unodes, 1);
            const int k = atomicAdd(&B.ctl->n_next_big, 1);
            if (u >= B.unode_cap || k >= B.big_cap) { atomicOr(&B.ctl->err, 2); code[side] = -1; continue; }
            B.next_big[k] = {f, l, t.depth + 1, u, t.set, 0, 0, 0};
            code[side] = u;
        } else {
            const int k = atomicAdd(&B.ctl->n_small, 1);
            if (k >= B.small_cap) { atomicOr(&B.ctl->err, 2); code[side] = -1; continue; }
            B.small[k] = {f, l, t.depth + 1, (level + 1) & 1, t.set, 0, 0, 0};
            code[side] = ~k;
        }
    }
    B.unodes[t.unode].left = code[0];
    B.unodes[t.unode].right = code[1];
}

// K5: per chunk: move the items to the next level's buffer, lefts and rights
// in order (a middle split keeps every item's position).
__global__ void __launch_bounds__(kThreads) k_chunk_scatter(Bufs B, int src) {
    __shared__ int sh[kWavesPerBlock];
    const int c = blockIdx.x;
    const int ti = B.chunk_task[c];
    const BigInfo inf = B.info[ti];
    const BigTask t = B.big[ti];
    const int s = B.chunk_start[c], e = min(s + kChunk, t.last);
    int b, en;
    thread_range(s, e, b, en);
    const Item* in = B.items[src];
    Item* out = B.items[src ^ 1];
    if (inf.mode != 0) {
        for (int i = b; i < en; ++i) store_item(out + i, load_item(in + i));
        return;
    }
    Item its[kIpt];
    bool left[kIpt];
    int nl = 0;
#pragma unroll
    for (int j = 0; j < kIpt; ++j) {
        left[j] = false;
        if (b + j < en) {
            its[j] = load_item(in + b + j);
            left[j] = bin_of(its[j], inf.axis, inf.lo_axis, inf.extent) <= inf.best;
            nl += left[j];
        }
    }
    const int nr = (en - b) - nl;
    const int pl = block_excl_scan(nl, sh);
    const int pr = block_excl_scan(nr, sh);
    int l = B.chunk_lbase[c] + pl, r = B.chunk_rbase[c] - pr;
#pragma unroll
    for (int j = 0; j < kIpt; ++j) {
        if (b + j >= en) break;
        if (left[j]) store_item(out + l++, its[j]);
        else store_item(out + r--, its[j]);
    }
}

// Next level: chunk lists of the tasks K4 appended (one block).
__global__ void __launch_bounds__(kThreads) k_plan(Bufs B) {
    __shared__ int sh[kWavesPerBlock];
    __shared__ int carry;
    const int n = B.ctl->n_next_big < B.big_cap ? B.ctl->n_next_big : B.big_cap;
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    for (int base = 0; base < n; base += kThreads) {
        const int i = base + threadIdx.x;
        BigTask t{};
        int nc = 0;
        if (i < n) {
            t = B.next_big[i];
            nc = (t.last - t.first + kChunk - 1) / kChunk;
        }
        const int off = carry + block_excl_scan(nc, sh);
        if (i < n) {
            if (off + nc > B.chunk_cap) {
                atomicOr(&B.ctl->err, 2);
            } else {
                B.next_big[i].chunk0 = off;
                B.next_big[i].nchunks = nc;
                for (int k = 0; k < nc; ++k) {
                    B.chunk_task_next[off + k] = i;
                    B.chunk_start_next[off + k] = t.first + k * kChunk;
                }
            }
        }
        __syncthreads();
        if (threadIdx.x == kThreads - 1) carry = off + nc;
        __syncthreads();
    }
    if (threadIdx.x == 0) B.ctl->n_chunks = carry;
}

__device__ __forceinline__ bn_bvh_node make_node(const DBox& b, int32_t word, bool leaf, int axis, int count) {
    bn_bvh_node n;
    n.pmin[0] = b.l[0]; n.pmin[1] = b.l[1]; n.pmin[2] = b.l[2];
    n.pmax[0] = b.h[0]; n.pmax[1] = b.h[1]; n.pmax[2] = b.h[2];
    n.right_or_offset = word;
    n.is_leaf = leaf ? 1 : 0;
    n.split_axis = (int8_t)axis;
    n.count = (int8_t)count;
    n.mask = 0;
    return n;
}

// One wave builds the whole subtree of a small task (Tree.Build, recursive,
// left first) into the task's node region, numbered in pre-order from 0.
// A task has at most 64 * kPer items, so each lane holds its run of at most
// kPer consecutive items in registers while a node is processed.
constexpr int kPer = 8;
constexpr int kSmallMax = 64 * kPer;
__global__ void __launch_bounds__(kThreads) k_small(Bufs B, int n_small) {
    struct Frame { int first, last, depth, parent, buf; };
    __shared__ Frame stack[kWavesPerBlock][kMaxDepth + 2];
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int ti = blockIdx.x * kWavesPerBlock + w;
    if (ti >= n_small) return;
    const SmallTask task = B.small[ti];
    const int off = B.set_off[task.set];
    bn_bvh_node* region = B.region + 2 * (size_t)task.first;
    Frame* st = stack[w];
    int sp = 0, nodes = 0;
    st[sp++] = {task.first, task.last, task.depth, -1, task.buf};
    while (sp > 0) {
        const Frame f = st[--sp];
        const int me = nodes++;
        if (f.parent >= 0 && lane == 0) region[f.parent].right_or_offset = me;
        const int cnt = f.last - f.first;
        const int per = (cnt + 63) / 64;                  // at most kPer
        const int b = min(f.first + lane * per, f.last), n = min(b + per, f.last) - b;
        const Item* in = B.items[f.buf];
        Item its[kPer];
        DBox bb = empty_box(), cb = empty_box();
#pragma unroll
        for (int j = 0; j < kPer; ++j)
            if (j < n) {
                its[j] = load_item(in + b + j);
                bb = join(bb, item_box(its[j]));
                cb = join(cb, centroid_box(its[j]));
            }
        bb = wave_union(bb);
        if (cnt <= kLeaf || f.depth >= kMaxDepth) {
            if (cnt > 127) {
                if (lane == 0) atomicOr(&B.ctl->err, 1);
            } else if (lane == 0) {
                region[me] = make_node(bb, f.first - off, true, 0, cnt);
            }
#pragma unroll
            for (int j = 0; j < kPer; ++j)
                if (j < n) B.perm[b + j] = its[j].id;
            continue;
        }
        cb = wave_union(cb);
        const int axis = split_axis(cb);
        const float extent = cb.h[axis] - cb.l[axis];
        if (lane == 0) region[me] = make_node(bb, 0, false, axis, 0);
        int mid;
        int child_buf = f.buf;
        if (extent == 0.0f) {
            mid = f.first + cnt / 2;                       // items stay where they are
        } else {
            const float lo = cb.l[axis];
            int bin[kPer];
#pragma unroll
            for (int j = 0; j < kPer; ++j) bin[j] = j < n ? bin_of(its[j], axis, lo, extent) : -1;
            DBox bbin[kBins];
            int bn[kBins];
#pragma unroll
            for (int k = 0; k < kBins; ++k) {
                bbin[k] = empty_box();
                bn[k] = 0;
#pragma unroll
                for (int j = 0; j < kPer; ++j)
                    if (bin[j] == k) { bbin[k] = join(bbin[k], item_box(its[j])); ++bn[k]; }
            }
            // the twelve ordered lane reductions, interleaved step by step
#pragma unroll
            for (int s = 1; s < 64; s <<= 1) {
#pragma unroll
                for (int k = 0; k < kBins; ++k) {
                    DBox o;
#pragma unroll
                    for (int q = 0; q < 3; ++q) {
                        o.l[q] = __shfl_down(bbin[k].l[q], s);
                        o.h[q] = __shfl_down(bbin[k].h[q], s);
                    }
                    bbin[k] = join(bbin[k], o);
                    bn[k] += __shfl_down(bn[k], s);
                }
            }
#pragma unroll
            for (int k = 0; k < kBins; ++k) {
#pragma unroll
                for (int q = 0; q < 3; ++q) {
                    bbin[k].l[q] = __shfl(bbin[k].l[q], 0);
                    bbin[k].h[q] = __shfl(bbin[k].h[q], 0);
                }
                bn[k] = __shfl(bn[k], 0);
            }
            float lcost[kBins - 1], rcost[kBins - 1];
            DBox acc = empty_box();
            int c = 0;
#pragma unroll
            for (int s = 0; s < kBins - 1; ++s) {
                acc = join(acc, bbin[s]);
                c += bn[s];
                lcost[s] = (float)c * area(acc);
            }
            acc = empty_box();
            c = 0;
#pragma unroll
            for (int s = kBins - 2; s >= 0; --s) {
                acc = join(acc, bbin[s + 1]);
                c += bn[s + 1];
                rcost[s] = (float)c * area(acc);
            }
            const float base = (float)cnt * area(cb);
            float best_cost = kInf;
            int best = 0;
#pragma unroll
            for (int s = 0; s < kBins - 1; ++s) {
                const float cost = (lcost[s] + rcost[s]) + base;
                if (cost < best_cost) { best_cost = cost; best = s; }
            }
            // partition into the other buffer: lefts from first, rights from last - 1 down
            int nl = 0;
#pragma unroll
            for (int j = 0; j < kPer; ++j) nl += (j < n && bin[j] <= best) ? 1 : 0;
            const int nr = n - nl;
            const int tl = wave_sum(nl);
            int l = f.first + wave_excl_scan(nl), r = f.last - 1 - wave_excl_scan(nr);
            Item* out = B.items[f.buf ^ 1];
#pragma unroll
            for (int j = 0; j < kPer; ++j)
                if (j < n) {
                    if (bin[j] <= best) store_item(out + l++, its[j]);
                    else store_item(out + r--, its[j]);
                }
            __threadfence_block();                 // the children read what other lanes wrote
            mid = f.first + tl;
            child_buf = f.buf ^ 1;
        }
        st[sp++] = {mid, f.last, f.depth + 1, me, child_buf};
        st[sp++] = {f.first, mid, f.depth + 1, -1, child_buf};
    }
    if (lane == 0) B.small_size[ti] = nodes;
}

__global__ void k_write_upper(Bufs B, int n_unodes, const int* udfs, const int* uright, bn_bvh_node* out) {
    const int u = blockIdx.x * blockDim.x + threadIdx.x;
    if (u >= n_unodes) return;
    const UNode n = B.unodes[u];
    out[2 * (size_t)B.set_off[n.set] + udfs[u]] = make_node(n.b, uright[u], false, n.axis, 0);
}

// Copy each region to its place in its set's node array; right-child indices
// move with the region's base.
__global__ void __launch_bounds__(kThreads) k_write_regions(Bufs B, int n_small, const int* rbase, bn_bvh_node* out) {
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int ti = blockIdx.x * kWavesPerBlock + w;
    if (ti >= n_small) return;
    const SmallTask t = B.small[ti];
    const int base = rbase[ti], n = B.small_size[ti];
    const bn_bvh_node* src = B.region + 2 * (size_t)t.first;
    bn_bvh_node* dst = out + 2 * (size_t)B.set_off[t.set] + base;
    for (int k = lane; k < n; k += 64) {
        bn_bvh_node x = src[k];
        if (!x.is_leaf) x.right_or_offset += base;
        dst[k] = x;
    }
}

}  // namespace bvhg

namespace bnx {

using namespace bvhg;
using bnh::Error;

namespace {
// Page-locked staging memory of the context, grown on demand.
void* pinned(Ctx* c, size_t bytes) {
    if (c->g_pinned_bytes < bytes) {
        if (c->g_pinned) (void)hipHostFree(c->g_pinned);
        c->g_pinned = nullptr;
        c->g_pinned_bytes = 0;
        hip_check(hipHostMalloc(&c->g_pinned, bytes, hipHostMallocDefault), "hipHostMalloc");
        c->g_pinned_bytes = bytes;
    }
    return c->g_pinned;
}
}  // namespace

void bvh_build_gpu(Ctx* c, const std::vector<const float*>& sets, const std::vector<int>& counts,
                   std::vector<std::vector<bn_bvh_node>>& nodes_out, std::vector<std::vector<int32_t>>& perm_out) {
    const int S = (int)sets.size();
    nodes_out.assign(S, {});
    perm_out.assign(S, {});
    if (S == 0) return;
    std::vector<int> off(S + 1, 0);
    for (int s = 0; s < S; ++s) {
        if (counts[s] <= 0) throw Error(BN_E_INVALID, "BVH build of an empty item set");
        if ((int64_t)off[s] + counts[s] > (1 << 28)) throw Error(BN_E_CAPACITY, "GPU BVH build: more than 2^28 items");
        off[s + 1] = off[s] + counts[s];
    }
    const int N = off[S];
    const int small = std::min(c->bvh_small, kSmallMax);
    const size_t item_bytes = (size_t)N * sizeof(Item);
    // host: items into page-locked staging, finiteness, the level-0 tasks
    Item* items = (Item*)pinned(c, std::max(item_bytes, (size_t)N * (sizeof(bn_bvh_node) * 2 + 4)));
    for (int s = 0; s < S; ++s) {
        const float* src = sets[s];
        Item* dst = items + off[s];
        bool finite = true;
        for (int i = 0; i < counts[s]; ++i) {
            const float* b = src + 6 * (size_t)i;
            finite = finite && std::isfinite(b[0]) && std::isfinite(b[1]) && std::isfinite(b[2]) &&
                     std::isfinite(b[3]) && std::isfinite(b[4]) && std::isfinite(b[5]);
            dst[i] = {b[0], b[1], b[2], i, b[3], b[4], b[5], 0};
        }
        if (!finite) throw Error(BN_E_INVALID, "GPU BVH build needs finite bounds (bn_bvh_build takes any)");
    }
    const int big_cap = N / small + S + 1;
    const int unode_cap = kMaxDepth * big_cap;
    const int small_cap = (int)std::min((int64_t)N + S, 2 * (int64_t)unode_cap + S) + 1;
    const int chunk_cap = 2 * (N / kChunk + big_cap) + 1;
    std::vector<BigTask> big0;
    std::vector<SmallTask> small0;
    std::vector<int> root(S), ctask, cstart;
    for (int s = 0; s < S; ++s) {
        if (counts[s] > small) {
            const int u = (int)big0.size();
            const int nc = (counts[s] + kChunk - 1) / kChunk;
            big0.push_back({off[s], off[s + 1], 0, u, s, (int)ctask.size(), nc, 0});
            for (int k = 0; k < nc; ++k) { ctask.push_back(u); cstart.push_back(off[s] + k * kChunk); }
            root[s] = u;
        } else {
            root[s] = ~(int)small0.size();
            small0.push_back({off[s], off[s + 1], 0, 0, s, 0, 0, 0});
        }
    }
    // device memory (grow-only buffers of the context)
    c->g_items.ensure(2 * std::max(item_bytes, (size_t)N * sizeof(bn_bvh_node)));   // A | B, then the nodes
    c->g_region.ensure(2 * (size_t)N * sizeof(bn_bvh_node));
    c->g_perm.ensure((size_t)N * 4);
    c->g_rsize.ensure((size_t)small_cap * 4 * 2);         // region sizes, region bases (by small task)
    c->g_big.ensure(2 * (size_t)big_cap * sizeof(BigTask));
    c->g_info.ensure((size_t)big_cap * sizeof(BigInfo));
    c->g_small.ensure((size_t)small_cap * sizeof(SmallTask));
    c->g_chunks.ensure(4 * (size_t)chunk_cap * 4);
    c->g_part.ensure((size_t)chunk_cap * (2 + kBins) * sizeof(DBox) + (size_t)chunk_cap * (kBins + 2) * 4);
    c->g_unodes.ensure((size_t)unode_cap * (sizeof(UNode) + 2 * 4));
    c->g_misc.ensure(sizeof(Ctl) + (size_t)S * 4);
    hipStream_t st = c->stream;
    char* items_base = (char*)c->g_items.p;
    Bufs B{};
    B.items[0] = (Item*)items_base;
    B.items[1] = (Item*)(items_base + item_bytes);
    BigTask* bigs[2] = {(BigTask*)c->g_big.p, (BigTask*)c->g_big.p + big_cap};
    B.info = (BigInfo*)c->g_info.p;
    B.small = (SmallTask*)c->g_small.p;
    int* chunks[4];
    for (int k = 0; k < 4; ++k) chunks[k] = (int*)c->g_chunks.p + (size_t)k * chunk_cap;
    B.part_b = (DBox*)c->g_part.p;
    B.part_cb = B.part_b + chunk_cap;
    B.bin_b = B.part_cb + chunk_cap;
    B.bin_n = (int*)(B.bin_b + (size_t)chunk_cap * kBins);
    B.chunk_lbase = B.bin_n + (size_t)chunk_cap * kBins;
    B.chunk_rbase = B.chunk_lbase + chunk_cap;
    B.unodes = (UNode*)c->g_unodes.p;
    int* udfs = (int*)(B.unodes + unode_cap);
    int* uright = udfs + unode_cap;
    B.region = (bn_bvh_node*)c->g_region.p;
    B.small_size = (int*)c->g_rsize.p;
    int* rbase = B.small_size + small_cap;
    B.perm = (int32_t*)c->g_perm.p;
    B.ctl = (Ctl*)c->g_misc.p;
    int* d_off = (int*)((char*)c->g_misc.p + sizeof(Ctl));
    B.set_off = d_off;
    B.big_cap = big_cap; B.small_cap = small_cap; B.unode_cap = unode_cap; B.chunk_cap = chunk_cap;

    hip_check(hipMemcpyAsync(B.items[0], items, item_bytes, hipMemcpyHostToDevice, st), "bvh items");
    if (!big0.empty()) {
        hip_check(hipMemcpyAsync(bigs[0], big0.data(), big0.size() * sizeof(BigTask), hipMemcpyHostToDevice, st), "bvh tasks");
        hip_check(hipMemcpyAsync(chunks[0], ctask.data(), ctask.size() * 4, hipMemcpyHostToDevice, st), "bvh chunks");
        hip_check(hipMemcpyAsync(chunks[1], cstart.data(), cstart.size() * 4, hipMemcpyHostToDevice, st), "bvh chunks");
    }
    if (!small0.empty())
        hip_check(hipMemcpyAsync(B.small, small0.data(), small0.size() * sizeof(SmallTask), hipMemcpyHostToDevice, st),
                  "bvh small tasks");
    hip_check(hipMemcpyAsync(d_off, off.data(), S * 4, hipMemcpyHostToDevice, st), "bvh sets");
    Ctl ctl{};
    ctl.n_small = (int)small0.size();
    ctl.n_unodes = (int)big0.size();
    hip_check(hipMemcpyAsync(B.ctl, &ctl, sizeof(Ctl), hipMemcpyHostToDevice, st), "bvh ctl");

    int n_big = (int)big0.size(), n_chunks = (int)ctask.size(), cur = 0;
    for (int level = 0; n_big > 0; ++level) {
        if (level >= kMaxDepth) throw Error(BN_E_INVALID, "GPU BVH build: deeper than the depth limit");
        B.big = bigs[cur];
        B.next_big = bigs[cur ^ 1];
        B.chunk_task = chunks[2 * cur];
        B.chunk_start = chunks[2 * cur + 1];
        B.chunk_task_next = chunks[2 * (cur ^ 1)];
        B.chunk_start_next = chunks[2 * (cur ^ 1) + 1];
        const int src = level & 1;
        const int tb = (n_big + 63) / 64;
        hip_check(hipMemsetAsync(&B.ctl->n_next_big, 0, 4, st), "bvh ctl");
        k_chunk_bounds<<<n_chunks, kThreads, 0, st>>>(B, src);
        k_task_bounds<<<tb, 64, 0, st>>>(B, n_big);
        k_chunk_bins<<<n_chunks, kThreads, 0, st>>>(B, src);
        k_task_split<<<tb, 64, 0, st>>>(B, n_big, level, small);
        k_chunk_scatter<<<n_chunks, kThreads, 0, st>>>(B, src);
        k_plan<<<1, kThreads, 0, st>>>(B);
        hip_check(hipGetLastError(), "bvh level launch");
        hip_check(hipMemcpyAsync(&ctl, B.ctl, sizeof(Ctl), hipMemcpyDeviceToHost, st), "bvh ctl");
        hip_check(hipStreamSynchronize(st), "bvh level");
        if (ctl.err) throw Error(BN_E_CAPACITY, "GPU BVH build: task capacity exceeded");
        n_big = ctl.n_next_big;
        n_chunks = ctl.n_chunks;
        cur ^= 1;
    }
    const int n_small = ctl.n_small;
    const int n_unodes = ctl.n_unodes;
    if (n_small > 0) k_small<<<(n_small + kWavesPerBlock - 1) / kWavesPerBlock, kThreads, 0, st>>>(B, n_small);
    hip_check(hipGetLastError(), "bvh small launch");
    // Pre-order numbering of the upper trees on the host (a few nodes per
    // 512 items): subtree sizes, then each upper node's index and right
    // child and each region's base.
    std::vector<UNode> un(n_unodes);
    std::vector<SmallTask> sm(n_small);
    std::vector<int> rsize(n_small);
    if (n_unodes) hip_check(hipMemcpyAsync(un.data(), B.unodes, n_unodes * sizeof(UNode), hipMemcpyDeviceToHost, st), "bvh upper");
    hip_check(hipMemcpyAsync(sm.data(), B.small, n_small * sizeof(SmallTask), hipMemcpyDeviceToHost, st), "bvh small");
    hip_check(hipMemcpyAsync(rsize.data(), B.small_size, n_small * 4, hipMemcpyDeviceToHost, st), "bvh sizes");
    hip_check(hipMemcpyAsync(&ctl, B.ctl, sizeof(Ctl), hipMemcpyDeviceToHost, st), "bvh ctl");
    hip_check(hipStreamSynchronize(st), "bvh subtrees");
    if (ctl.err & 1) throw Error(BN_E_CAPACITY, "BVH leaf larger than the int8 count field");
    if (ctl.err) throw Error(BN_E_CAPACITY, "GPU BVH build: task capacity exceeded");
    std::vector<int> usize(n_unodes), h_udfs(n_unodes), h_uright(n_unodes), h_rbase(n_small), nn(S);
    auto size_of = [&](int code) { return code < 0 ? rsize[~code] : usize[code]; };
    std::vector<std::pair<int, int>> stk;
    for (int s = 0; s < S; ++s) {
        stk.assign(1, {root[s], 0});
        while (!stk.empty()) {                               // post-order: sizes
            auto& [code, vis] = stk.back();
            if (code < 0) { stk.pop_back(); continue; }
            const int u = code;
            if (!vis) {
                vis = 1;
                stk.push_back({un[u].right, 0});
                stk.push_back({un[u].left, 0});
            } else {
                stk.pop_back();
                usize[u] = 1 + size_of(un[u].left) + size_of(un[u].right);
            }
        }
        nn[s] = size_of(root[s]);
        if (nn[s] <= 0 || nn[s] > 2 * counts[s]) throw Error(BN_E_DEVICE, "GPU BVH build: inconsistent node count");
        stk.assign(1, {root[s], 0});
        while (!stk.empty()) {                               // pre-order: indices
            const auto [code, idx] = stk.back();
            stk.pop_back();
            if (code < 0) { h_rbase[~code] = idx; continue; }
            const int sl = size_of(un[code].left);
            h_udfs[code] = idx;
            h_uright[code] = idx + 1 + sl;
            stk.push_back({un[code].right, idx + 1 + sl});
            stk.push_back({un[code].left, idx + 1});
        }
    }
    bn_bvh_node* out = (bn_bvh_node*)items_base;                // the item buffers are done with
    if (n_unodes) {
        hip_check(hipMemcpyAsync(udfs, h_udfs.data(), n_unodes * 4, hipMemcpyHostToDevice, st), "bvh numbering");
        hip_check(hipMemcpyAsync(uright, h_uright.data(), n_unodes * 4, hipMemcpyHostToDevice, st), "bvh numbering");
        k_write_upper<<<(n_unodes + 255) / 256, 256, 0, st>>>(B, n_unodes, udfs, uright, out);
    }
    hip_check(hipMemcpyAsync(rbase, h_rbase.data(), n_small * 4, hipMemcpyHostToDevice, st), "bvh numbering");
    k_write_regions<<<(n_small + kWavesPerBlock - 1) / kWavesPerBlock, kThreads, 0, st>>>(B, n_small, rbase, out);
    hip_check(hipGetLastError(), "bvh flatten launch");
    // results through the staging memory: the nodes of every set, then the permutation
    bn_bvh_node* h_nodes = (bn_bvh_node*)items;
    int32_t* h_perm = (int32_t*)(h_nodes + 2 * (size_t)N);
    for (int s = 0; s < S; ++s)
        hip_check(hipMemcpyAsync(h_nodes + 2 * (size_t)off[s], out + 2 * (size_t)off[s], nn[s] * sizeof(bn_bvh_node),
                                 hipMemcpyDeviceToHost, st), "bvh nodes");
    hip_check(hipMemcpyAsync(h_perm, B.perm, (size_t)N * 4, hipMemcpyDeviceToHost, st), "bvh perm");
    hip_check(hipStreamSynchronize(st), "bvh copy");
    for (int s = 0; s < S; ++s) {
        nodes_out[s].assign(h_nodes + 2 * (size_t)off[s], h_nodes + 2 * (size_t)off[s] + nn[s]);
        perm_out[s].assign(h_perm + off[s], h_perm + off[s + 1]);
    }
}

}  // namespace bnx
