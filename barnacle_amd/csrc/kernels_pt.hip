// kernels_pt.hip -- wavefront path tracing on gfx950.
//
// One pass of PathTracingIntegrator.Li (Extensions/Integrator/PathTracing.fs:14-81)
// is split into per-bounce kernels over compacted queues of path slots.  A
// round streams its paths through a fixed pool of slots; every iteration
//
//   regen -> prep -> trace_closest -> bin -> shade -> trace_shadow
//
// first retires the paths that ended in the previous iteration and refills
// their slots with new camera paths (regen), then advances every active path
// by one bounce whatever its depth (the depth travels with the path).  The
// GPU therefore stays full until the last paths of the round drain, instead
// of idling through the long tail of every bounce loop.  The L update order
// of the reference (emission MIS, then NEE, per bounce) is kept because
// trace_shadow of bounce k runs before shade of bounce k+1.  The radiance of
// every (pixel, sample) is stored by path id when the path retires and folded
// into the film in sample order by k_film, which makes the film bit-identical
// to ProgressiveIntegrator.RenderTile (Base/Integrator.fs:36-44) for any
// schedule.
//
// Every kernel exists in two modes.  Plain: one scene and one film (kernel
// arguments).  Batch: the paths of one round belong to several render items
// (scene, film, depth limits, camera) of a device-side round table; each slot
// records its item, and each lane reads its own scene.  Per-path random
// streams do not depend on scheduling, so an item's film is bit-identical in
// both modes.
#include "kernels.hpp"
#include "shading.hpp"
#include "traverse.hpp"
#include "wave.hpp"

namespace bnd {

// ----------------------------------------------------------- wave helpers
// Wave-uniform fetch of `n` work items: one atomic per wave.
__device__ __forceinline__ uint32_t wave_fetch(uint32_t* ctr, uint32_t n) {
    uint32_t base = 0;
    if (lane_id() == 0) base = atomicAdd(ctr, n);
    return __shfl(base, 0);
}
// Ballot-compacted append: one atomic per wave, lanes with pred write val.
// Must be reached by every lane of the wave.
__device__ __forceinline__ void wave_push(bool pred, uint32_t* ctr, uint32_t* arr, uint32_t val) {
    const uint64_t m = __ballot(pred);
    if (m == 0) return;
    const int leader = __ffsll((unsigned long long)m) - 1;
    uint32_t base = 0;
    if ((int)lane_id() == leader) base = atomicAdd(ctr, (uint32_t)__popcll(m));
    base = __shfl(base, leader);
    if (pred) arr[base + mbcnt(m)] = val;
}
__device__ __forceinline__ void wave_add_stats(unsigned long long* dst, uint32_t v) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    if (lane_id() == 0 && v) atomicAdd(dst, (unsigned long long)v);
}

// ---------------------------------------------------------- work context
// Plain mode: one item, one scene.  Batch mode: round table + scene table.
template <bool BATCH>
struct Work;

template <>
struct Work<false> {
    DevScene S;
    RoundItem it;
    __device__ __forceinline__ int n() const { return 1; }
    __device__ __forceinline__ int item_of_path(uint32_t) const { return 0; }
    __device__ __forceinline__ int item_of(const PathBuf&, uint32_t) const { return 0; }
    __device__ __forceinline__ const RoundItem& item(int) const { return it; }
    __device__ __forceinline__ const DevScene& scene(int) const { return S; }
    __device__ __forceinline__ uint32_t total_paths() const { return (uint32_t)it.path_count; }
    // One scene for the whole launch: traversals bind it once, outside the
    // divergent refill code, so its pointers stay wave-uniform.
    __device__ __forceinline__ void bind(TraceScene& g) const { g.set(S); }
    __device__ __forceinline__ void bind_lane(TraceScene&, int) const {}
};

template <>
struct Work<true> {
    const RoundHeader* hdr;
    const RoundItem* items;
    const DevScene* scenes;
    __device__ __forceinline__ int n() const { return hdr->n_items; }
    // item owning a path id (the round table is ordered by path_base)
    __device__ __forceinline__ int item_of_path(uint32_t id) const {
        int lo = 0, hi = hdr->n_items - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if ((uint32_t)items[mid].path_base <= id) lo = mid; else hi = mid - 1;
        }
        return lo;
    }
    __device__ __forceinline__ int item_of(const PathBuf& P, uint32_t slot) const { return P.item[slot]; }
    __device__ __forceinline__ const RoundItem& item(int k) const { return items[k]; }
    __device__ __forceinline__ const DevScene& scene(int k) const { return scenes[items[k].scene]; }
    __device__ __forceinline__ uint32_t total_paths() const { return hdr->n_paths; }
    __device__ __forceinline__ void bind(TraceScene&) const {}
    __device__ __forceinline__ void bind_lane(TraceScene& g, int k) const { g.set(scene(k)); }
};

// local path id -> (pixel, sample): 8x8 pixel tiles, sample-major.
__device__ __forceinline__ bool path_pixel(const RoundItem& c, uint32_t q_all, int& x, int& y, int& s) {
    const uint32_t q = q_all % (uint32_t)c.p_pad;
    s = c.s0 + (int)(q_all / (uint32_t)c.p_pad);
    const uint32_t tile = q >> 6;
    x = (int)((tile % (uint32_t)c.tiles_x) * 8 + (q & 7));
    y = (int)((tile / (uint32_t)c.tiles_x) * 8 + ((q >> 3) & 7));
    return x < c.width && y < c.height;
}

// ------------------------------------------------------------------- regen
// Free slot j first retires its finished path (radiance stored by path id),
// then takes path id next_pid + j of the round: RenderTile seeding and the
// primary ray (Base/Integrator.fs:34-39), appended to the active queue after
// the paths that continue.  Ids and queue positions follow from j, so the
// kernel needs no atomics; k_prep then advances the counters.  A padding
// pixel, or a path with max-depth 0 (radiance 0), becomes a placeholder with
// depth -1 that misses at once.  Past the last id the slot goes idle.
__device__ __forceinline__ uint32_t new_paths(uint32_t n_free, uint32_t next_pid, uint32_t n_paths) {
    // signed 64-bit: the difference can neither wrap nor be assumed not to
    const int64_t left = (int64_t)n_paths - (int64_t)next_pid;
    return left <= 0 ? 0u : (uint32_t)(left < (int64_t)n_free ? left : (int64_t)n_free);
}

// Adjoint rounds (ADJ) replay the same ids: a new path starts with L = its
// forward radiance (the radiance still to come, decremented as the replay
// passes each contribution) and delta = film adjoint / spp; nothing retires.
template <bool BATCH, bool ADJ>
__global__ __launch_bounds__(256) void k_regen(Work<BATCH> W, PathBuf P, const Ctl* ctl, int cur,
                                               const uint32_t* freel_in, uint32_t* queue, float4* L_out,
                                               AdjBuf A) {
    const uint32_t n_free = ctl->free_count[cur];
    const uint32_t next_pid = ctl->next_pid;
    const uint32_t n_new = new_paths(n_free, next_pid, W.total_paths());
    uint32_t* q = queue + ctl->q_count[cur];
    for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < n_free; j += gridDim.x * blockDim.x) {
        const uint32_t slot = freel_in[j];
        const uint32_t old = P.pid[slot];
        if (!ADJ && old != kNoPath) L_out[old] = P.L[slot];
        if (j >= n_new) {
            P.pid[slot] = kNoPath;
            continue;
        }
        const uint32_t id = next_pid + j;
        const int k = W.item_of_path(id);
        const RoundItem& it = W.item(k);
        int x = 0, y = 0, s = 0;
        bool ok = path_pixel(it, id - (uint32_t)it.path_base, x, y, s) && it.max_depth > 0;
        float4 dl = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        if (ADJ && ok) {
            // a path whose pixel has no adjoint (or whose radiance is not
            // finite) carries no gradient: replay it as a placeholder
            const float* fp = A.d_film + 3 * ((size_t)(it.height - y - 1) * it.width + x);
            dl = make_float4(fp[0] * it.inv_spp, fp[1] * it.inv_spp, fp[2] * it.inv_spp, 0.0f);
            const float4 lf = A.L_fwd[id];
            ok = (dl.x != 0.0f || dl.y != 0.0f || dl.z != 0.0f) && isfinite(lf.x) && isfinite(lf.y) &&
                 isfinite(lf.z);
        }
        if (ok) {
            const DevScene& S = W.scene(k);
            uint32_t rng = xxhash3((uint32_t)x, (uint32_t)y, (uint32_t)(it.frame_id * it.spp + s));
            V2 uPix, uLens;
            uPix.x = lcg(rng); uPix.y = lcg(rng);
            uLens.x = lcg(rng); uLens.y = lcg(rng);
            Ray r = camera_ray(S, x, y, uPix, uLens);
            P.o[slot] = f4(r.o, __int_as_float(0));
            P.d[slot] = f4(r.d, 0.0f);
            P.beta[slot] = make_float4(1.0f, 1.0f, 1.0f, 0.0f);
            P.rng[slot] = rng;
        } else {
            P.o[slot] = make_float4(0.0f, 0.0f, 0.0f, __int_as_float(-1));
            P.d[slot] = make_float4(0.0f, 0.0f, 1.0f, 0.0f);
        }
        if (ADJ) {
            A.delta[slot] = dl;
            P.L[slot] = ok ? A.L_fwd[id] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        } else {
            P.L[slot] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        }
        P.pid[slot] = id;
        if (BATCH) P.item[slot] = k;
        q[j] = slot;
    }
}

// Compact rounds: the new paths take positions after the n_act paths that
// continue from the previous iteration, up to n_slots, with the seeding and
// primary rays of k_regen.  A path's radiance is accumulated in L_out[path
// id] for the whole round, from zero here.
template <bool BATCH>
__global__ __launch_bounds__(256) void k_regen_c(Work<BATCH> W, PathBuf X, const Ctl* ctl, int cur,
                                                 uint32_t n_slots, float4* L_out) {
    const uint32_t n_act = ctl->q_count[cur];
    const uint32_t next_pid = ctl->next_pid;
    const uint32_t n_new = new_paths(n_slots - n_act, next_pid, W.total_paths());
    for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < n_new; j += gridDim.x * blockDim.x) {
        const uint32_t pos = n_act + j;
        const uint32_t id = next_pid + j;
        const int k = W.item_of_path(id);
        const RoundItem& it = W.item(k);
        int x = 0, y = 0, s = 0;
        if (path_pixel(it, id - (uint32_t)it.path_base, x, y, s) && it.max_depth > 0) {
            const DevScene& S = W.scene(k);
            uint32_t rng = xxhash3((uint32_t)x, (uint32_t)y, (uint32_t)(it.frame_id * it.spp + s));
            V2 uPix, uLens;
            uPix.x = lcg(rng); uPix.y = lcg(rng);
            uLens.x = lcg(rng); uLens.y = lcg(rng);
            const Ray r = camera_ray(S, x, y, uPix, uLens);
            X.o[pos] = f4(r.o, __int_as_float(0));
            X.d[pos] = f4(r.d, 0.0f);
            X.rng[pos] = rng;
        } else {
            X.o[pos] = make_float4(0.0f, 0.0f, 0.0f, __int_as_float(-1));
            X.d[pos] = make_float4(0.0f, 0.0f, 1.0f, 0.0f);
        }
        L_out[id] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        X.pid[pos] = id;
        if (BATCH) X.item[pos] = k;
    }
}

// After regen: the new paths join the active queue, the ids advance, and the
// counters of the other parity (filled during this iteration) start at zero.
// compact_slots: 0 for slot rounds (new paths fill the free list), else the
// capacity of a compact round.
// Compact rounds clear only this parity's shadow counters: the shadow rays
// of the previous iteration may still be in flight with the other set.
template <bool BATCH>
__global__ void k_prep(Work<BATCH> W, Ctl* ctl, int cur, uint32_t* fetch_seg, uint32_t compact_slots) {
    const int i = threadIdx.x;
    const int nxt = 1 - cur;
    if (compact_slots) {
        if (i < kFetchSegs) fetch_seg[(size_t)i * kSegStride] = 0;
        else if (i < kFetchSegs + kShadowSegs)
            fetch_seg[(size_t)(i + cur * kShadowSegs) * kSegStride] = 0;
    } else if (i < 2 * kFetchSegs) {
        fetch_seg[(size_t)i * kSegStride] = 0;
    }
    if (i < kBins) ctl->bin_count[i] = 0;
    if (i == 0) {
        const uint32_t n_free = compact_slots ? compact_slots - ctl->q_count[cur] : ctl->free_count[cur];
        const uint32_t n_new = new_paths(n_free, ctl->next_pid, W.total_paths());
        ctl->q_count[cur] += n_new;
        ctl->next_pid += n_new;
        ctl->sh_count = 0;
        ctl->shc[cur] = 0;
        ctl->q_count[nxt] = 0;
        ctl->free_count[nxt] = 0;
    }
}

// Start of a round: every slot free and idle.
__global__ void k_init_slots(PathBuf P, Ctl* ctl, uint32_t* freel, uint32_t n_slots) {
    for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < n_slots; j += gridDim.x * blockDim.x) {
        freel[j] = j;
        P.pid[j] = kNoPath;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) ctl->free_count[0] = n_slots;
}

// Counter clear at the start of a round, and per-iteration reset (kernels,
// so that a captured graph carries no memset nodes).
__global__ void k_ctl_clear(Ctl* ctl) {
    uint32_t* w = reinterpret_cast<uint32_t*>(ctl);
    if (threadIdx.x < sizeof(Ctl) / 4) w[threadIdx.x] = 0;
}
__global__ void k_reset(Ctl* ctl, int nxt, uint32_t* fetch_seg, int par) {
    const int i = threadIdx.x;
    const bool mine = par < 0 ? i < kFetchSegs + 2 * kShadowSegs
                              : i < kFetchSegs || (i >= kFetchSegs + par * kShadowSegs &&
                                                   i < kFetchSegs + (par + 1) * kShadowSegs);
    if (mine) fetch_seg[(size_t)i * kSegStride] = 0;
    if (i < kBins) ctl->bin_count[i] = 0;
    if (i == 0) {
        if (par < 0) ctl->sh_count = 0;
        else ctl->shc[par] = 0;
        ctl->q_count[nxt] = 0;
        ctl->free_count[nxt] = 0;
    }
    if (i < 4) ctl->fetch[i] = 0;
}

// ------------------------------------------------------------ trace closest
// Lane-refilling persistent traversal: a wave keeps up to 64 rays in flight,
// advances each by one item per iteration and refills idle lanes from the
// queue once at least `refill` of them are idle.  The queue is cut into
// segments with one fetch counter each (a counter serialises its atomics, so
// one shared counter would cap the whole GPU at one refill per ~11 ns): a
// wave starts on its own segment and moves on when that one runs dry.
struct SegFetch {
    uint32_t* ctr;    // segs counters, kSegStride words apart, zero at launch
    int segs;
};

template <bool ANY, bool COUNT, int K, typename Setup, typename Init, typename Done>
__device__ __forceinline__ void persistent_trace(SegFetch F, uint32_t count, uint32_t* stk, uint32_t* spill,
                                                 size_t spill_stride, uint32_t refill, TraceCount& tc, Setup setup,
                                                 Init init, Done done) {
    const size_t gid = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    Traversal<K> tr;
    tr.st = LaneStack<K>{stk + threadIdx.x, (int)blockDim.x, spill + gid, spill_stride, 0};
    tr.clear();
    setup(tr);
    const uint32_t seg_len = (count + (uint32_t)F.segs - 1) / (uint32_t)F.segs;
    int seg = (int)((blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6)) % (uint32_t)F.segs);
    int tried = 0;
    bool active = false, more = count > 0;
    uint32_t item = 0;
    for (;;) {
        const uint64_t idle = __ballot(!active);
        const uint32_t n_idle = (uint32_t)__popcll(idle);
        if (more && n_idle >= refill) {
            const uint32_t lo = (uint32_t)seg * seg_len;
            const uint32_t hi = min(count, lo + seg_len);
            uint32_t* c = F.ctr + (size_t)seg * kSegStride;
            uint32_t b = 0;
            if (lane_id() == 0) {
                const uint32_t seen = __hip_atomic_load(c, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                b = lo + seen < hi ? atomicAdd(c, n_idle) : seg_len;
            }
            b = __shfl(b, 0);
            // idle lanes take [start, start + n_idle) clipped to the segment;
            // each lane bounds its own item (start may lie past hi)
            const uint32_t start = lo + b;
            if (start + n_idle >= hi) {
                seg = seg + 1 == F.segs ? 0 : seg + 1;
                if (++tried >= F.segs) more = false;
            }
            if (!active) {
                const uint32_t i = start + mbcnt(idle);
                if (i < hi) {
                    item = i;
                    init(tr, item);
                    active = true;
                }
            }
        }
        if (__ballot(active) == 0) {
            if (!more) break;
            continue;
        }
        // Every lane steps: an idle lane's cleared state has no item, so its
        // step does nothing (and keeps the loop free of a per-lane branch).
        bool found = false;
        const bool going = COUNT ? tr.template step<ANY, COUNT>(tc, found) : tr.template step_fast<ANY>(found);
        const bool fin = active && !going;
        if (__ballot(fin) != 0) {
            done(tr, item, fin, found);
            if (fin) active = false;
            // a finished closest-hit traversal is already empty; an any-hit
            // one may stop with items left
            if (ANY && found) tr.clear();
        }
    }
}

// BVHAggregate.Intersect closest (Extensions/Aggregate/BVH.fs:37-58).  The
// material bin of the hit (kind of the instance's material; light-only
// instances in the last bin, -1 for a miss) travels in the hit record.  The
// rays are queue[i], or position i of a compact round (queue null).
// With a 16-entry LDS stack window seven blocks fit a CU's LDS; the register
// budget is then set for seven waves per SIMD (72 VGPRs; eight would spill),
// six in batches.
template <bool COUNT, bool BATCH, int K>
__global__ __launch_bounds__(256, K == 16 ? (BATCH ? kTraceWavesBatch : kTraceWaves) : 1) void k_trace_closest(Work<BATCH> W, PathBuf P, Ctl* ctl, int cur,
                                                       const uint32_t* queue, uint32_t* spill, size_t spill_stride,
                                                       uint32_t refill, unsigned long long* stats, uint32_t* seg) {
    __shared__ uint32_t stk[K * 256];
    const uint32_t count = ctl->q_count[cur];
    TraceCount tc{0, 0, 0};
    persistent_trace<false, COUNT, K>(
        SegFetch{seg, kFetchSegs}, count, stk, spill, spill_stride, refill, tc,
        [&](Traversal<K>& tr) { W.bind(tr.G); },
        [&](Traversal<K>& tr, uint32_t i) {
            const uint32_t slot = queue ? queue[i] : i;
            const float4 o4 = P.o[slot];
            const int k = W.item_of(P, slot);
            W.bind_lane(tr.G, k);
            tr.start({xyz(o4), xyz(P.d[slot])}, kInf, W.scene(k).exact_only != 0);
            if (!COUNT) tr.take_tlas_root();
            // a placeholder path (padding pixel, max-depth 0) misses at once
            if (__float_as_int(o4.w) < 0) tr.empty();
        },
        [&](Traversal<K>& tr, uint32_t i, bool fin, bool) {
            if (fin) {
                const uint32_t slot = queue ? queue[i] : i;
                int bin = -1;
                if (tr.hit_inst >= 0) {
                    const DevScene& S = W.scene(W.item_of(P, slot));
                    const int4 info = S.insts[tr.hit_inst].info;
                    bin = info.z >= 0 ? __float_as_int(S.mats[2 * info.z].x) : kBins - 1;
                }
                P.hit[slot] = make_int4(tr.hit_inst, tr.hit_prim, __float_as_int(tr.t), bin);
            }
        });
    if (COUNT) {
        wave_add_stats(&stats[0], tc.nodes);
        wave_add_stats(&stats[1], tc.tris);
        wave_add_stats(&stats[2], tc.insts);
    }
}

// Compaction of the closest hits for shading, optionally binned by material
// kind (sort != 0).  Sorting makes waves shade one BSDF kind, but scatters
// the slots each wave touches; on the mesh scenes measured here keeping the
// queue order (ascending slots, coalesced path-state access) is faster, so it
// is the default.  A block handles kBinItems consecutive queue entries:
// positions inside the block come from wave-aggregated LDS counters, and each
// bin costs one global atomic per block.  A miss ends the path: with a free
// list (freel non-null) its slot goes there, as one more bin; the PSSMLT
// driver passes no list and retires misses itself.
constexpr int kBinPer = 16;
constexpr int kBinItems = 256 * kBinPer;

__global__ __launch_bounds__(256) void k_bin(PathBuf P, Ctl* ctl, int cur, const uint32_t* queue, uint32_t* bins,
                                             uint32_t bin_cap, uint32_t* freel, int sort) {
    __shared__ uint32_t cnt[kBins + 1], gbase[kBins + 1];
    const uint32_t count = ctl->q_count[cur];
    const int n_out = freel ? kBins + 1 : kBins;
    for (uint32_t blk = blockIdx.x * (uint32_t)kBinItems; blk < count; blk += gridDim.x * (uint32_t)kBinItems) {
        if (threadIdx.x <= kBins) cnt[threadIdx.x] = 0;
        __syncthreads();
        uint32_t slot[kBinPer], pos[kBinPer];
        int bin[kBinPer];
        for (int k = 0; k < kBinPer; ++k) {
            const uint32_t i = blk + (uint32_t)k * 256u + threadIdx.x;
            bin[k] = -1;
            slot[k] = 0;
            if (i < count) {
                slot[k] = queue[i];
                const int b = P.hit[slot[k]].w;
                bin[k] = b >= 0 ? (sort ? b : 0) : (freel ? kBins : -1);
            }
            pos[k] = 0;
            for (int b = 0; b < n_out; ++b) {
                const uint64_t m = __ballot(bin[k] == b);
                if (m == 0) continue;
                const int leader = __ffsll((unsigned long long)m) - 1;
                uint32_t base = 0;
                if ((int)lane_id() == leader) base = atomicAdd(&cnt[b], (uint32_t)__popcll(m));
                base = __shfl(base, leader);
                if (bin[k] == b) pos[k] = base + mbcnt(m);
            }
        }
        __syncthreads();
        if (threadIdx.x < (uint32_t)n_out && cnt[threadIdx.x]) {
            uint32_t* ctr = threadIdx.x < kBins ? &ctl->bin_count[threadIdx.x] : &ctl->free_count[1 - cur];
            gbase[threadIdx.x] = atomicAdd(ctr, cnt[threadIdx.x]);
        }
        __syncthreads();
        for (int k = 0; k < kBinPer; ++k) {
            if (bin[k] < 0) continue;
            if (bin[k] < kBins) bins[(size_t)bin[k] * bin_cap + gbase[bin[k]] + pos[k]] = slot[k];
            else freel[gbase[kBins] + pos[k]] = slot[k];
        }
        __syncthreads();
    }
}

// Block-aggregated positions in N queues: one global atomic per queue and
// block.  pos[o] is the thread's position in queue o when pred[o] holds.
// Every thread of the block (NW waves) must call it.
template <int N, int NW = 4>
__device__ __forceinline__ void block_pos(const bool (&pred)[N], uint32_t* const (&ctr)[N], uint32_t (&pos)[N]) {
    __shared__ uint32_t wcnt[N][NW + 1];
    const int w = (int)(threadIdx.x >> 6);
    uint64_t m[N];
    for (int o = 0; o < N; ++o) {
        m[o] = __ballot(pred[o]);
        if (lane_id() == 0) wcnt[o][w] = (uint32_t)__popcll(m[o]);
    }
    __syncthreads();
    if (threadIdx.x < (uint32_t)N) {
        const int o = (int)threadIdx.x;
        uint32_t run = 0;
        for (int v = 0; v < NW; ++v) {
            const uint32_t c = wcnt[o][v];
            wcnt[o][v] = run;
            run += c;
        }
        wcnt[o][NW] = run ? atomicAdd(ctr[o], run) : 0u;
    }
    __syncthreads();
    for (int o = 0; o < N; ++o) pos[o] = wcnt[o][NW] + wcnt[o][w] + mbcnt(m[o]);
    __syncthreads();
}

// Block-aggregated append of `val` to N queues.
template <int N, int NW = 4>
__device__ __forceinline__ void block_push(const bool (&pred)[N], uint32_t* const (&ctr)[N],
                                           uint32_t* const (&arr)[N], uint32_t val) {
    uint32_t pos[N];
    block_pos<N, NW>(pred, ctr, pos);
    for (int o = 0; o < N; ++o)
        if (pred[o]) arr[o][pos[o]] = val;
}

// ------------------------------------------------------------------- shade
// One vertex of PathTracing.fs:24-79 after the closest hit: emission with MIS
// (:30-40), next-event estimation set-up (:42-50), BSDF sampling (:61-67) and
// Russian roulette (:69-75).  The NEE contribution is applied by
// k_trace_shadow when the shadow ray is unoccluded (:47-59).  The path's
// bounce depth is o.w.  The state is read at index `at` of P; the results
// are returned for the caller to store.
struct Vertex {
    V3 L;                             // radiance after this vertex's emission term
    bool cont, shadow;                // the path continues / casts a shadow ray
    float4 o, d, beta;                // next ray (o.w depth, d.w BSDF pdf) and throughput, when cont
    uint32_t rng;
    float4 sh_o, sh_d, nee_a, nee_b;  // shadow ray (sh_o.w t max) and NEE factors, when shadow
};
struct ShadeStage;
// One channel of an NEE term a*b that adds exactly nothing: an operand is
// zero and the other finite, so fma(a, b, L) == L for every L the radiance
// can hold (it starts at +0 and no fma of it gives -0).
__device__ __forceinline__ bool dark_term(float a, float b) {
    return (a == 0.0f && __builtin_isfinite(b)) || (b == 0.0f && __builtin_isfinite(a));
}
// Lacc: the path's radiance accumulator (compact rounds), read and written
// only when the vertex emits; null: the radiance is P.L[at], returned in v.L.
// stage: the shadow ray and NEE factors go to LDS slot `slot` as soon as they
// are known (they do not occupy registers through BSDF sampling); null: they
// are returned in v.
__device__ __forceinline__ void stage_shadow(ShadeStage* stage, int slot, const float4& o, const float4& d,
                                             const float4& a, const float4& b);
__device__ __forceinline__ void shade_vertex(const DevScene& S, const RoundItem& it, const PathBuf& P, uint32_t at,
                                             int4 h, Vertex& v, float4* Lacc = nullptr, ShadeStage* stage = nullptr,
                                             int slot = 0, bool skip_dark = false) {
    const float4 o4 = P.o[at], d4 = P.d[at];
    const Ray r = {xyz(o4), xyz(d4)};
    const float prev_pdf = d4.w;
    const int depth = __float_as_int(o4.w);
    const Geom g = hit_geometry(S, r, h.x, h.y, __int_as_float(h.z));
    const int4 info = S.insts[h.x].info;
    V3 L = Lacc ? V3{0.0f, 0.0f, 0.0f} : xyz(P.L[at]);
    // a camera path starts with beta = 1 (regen_c does not store it)
    V3 beta = depth == 0 ? V3{1.0f, 1.0f, 1.0f} : xyz(P.beta[at]);
    v.cont = false;
    v.shadow = false;
    if (info.w >= 0) {
        V3 Le;
        float lpdf;
        light_eval(S, h.x, r.o, g, Le, lpdf);
        const float w = depth == 0 ? 1.0f : prev_pdf * rcp_e(lpdf + prev_pdf);
        if (Lacc) *Lacc = f4(fma3(beta, Le * w, xyz(*Lacc)), 0.0f);
        else L = fma3(beta, Le * w, L);
    }
    if (info.z >= 0) {
        uint32_t rng = P.rng[at];
        const Mat m = load_mat(S, info.z);
        const V3 wo = -r.d;
        const V3 wo_l = to_local(g, wo);
        const float uSel = lcg(rng);
        V2 uL;
        uL.x = lcg(rng); uL.y = lcg(rng);
        const LightSample ls = light_sample(S, g.p, uSel, uL);
        if (ls.pdf != 0.0f) {
            const BSDF e = mat_eval(m, wo_l, to_local(g, ls.wi));
            const float dist = length(ls.p - g.p);
            const float4 so = f4(g.p, dist - 1e-3f), sd = f4(ls.wi, 0.0f);
            const float4 na = f4(beta * e.f, 0.0f), nb = f4(ls.L * rcp_e(e.pdf + ls.pdf), 0.0f);
            if (stage) {
                stage_shadow(stage, slot, so, sd, na, nb);
            } else {
                v.sh_o = so; v.sh_d = sd; v.nee_a = na; v.nee_b = nb;
            }
            v.shadow = !(skip_dark && dark_term(na.x, nb.x) && dark_term(na.y, nb.y) && dark_term(na.z, nb.z));
        }
        const float ul = lcg(rng);
        V2 ub;
        ub.x = lcg(rng); ub.y = lcg(rng);
        V3 wi_l;
        const BSDF bs = mat_sample(m, wo_l, ul, ub, wi_l);
        if (bs.pdf != 0.0f) {
            const V3 wi = to_world(g, wi_l);
            beta = beta * bs.f * rcp_e(bs.pdf);
            v.cont = true;
            if (depth >= it.rr_depth) {
                const float q = min_n(1.0f, max_n(beta.x, max_n(beta.y, beta.z)));
                if (lcg(rng) < q) beta = beta * rcp_e(q);
                else v.cont = false;
            }
            v.o = f4(g.p, __int_as_float(depth + 1));
            v.d = f4(wi, bs.pdf);
            v.beta = f4(beta, 0.0f);
        }
        v.rng = rng;
        v.cont = v.cont && depth + 1 < it.max_depth;
    }
    v.L = L;
}

// Slot rounds: the vertices of the bins, shaded in place; a path that does
// not continue frees its slot for the next iteration (its NEE term still
// lands in this iteration).
template <bool BATCH>
__global__ __launch_bounds__(256) void k_shade(Work<BATCH> W, PathBuf P, Ctl* ctl, const uint32_t* bins,
                                               uint32_t bin_cap, uint32_t* queue_next, int nxt, uint32_t* shq,
                                               uint32_t* freel) {
    uint32_t start[kBins + 1];
    start[0] = 0;
    for (int b = 0; b < kBins; ++b) start[b + 1] = start[b] + ctl->bin_count[b];
    const uint32_t total = start[kBins];
    for (uint32_t base = blockIdx.x * blockDim.x; base < total; base += gridDim.x * blockDim.x) {
        const uint32_t i = base + threadIdx.x;
        Vertex v;
        v.cont = false;
        v.shadow = false;
        uint32_t slot = 0;
        if (i < total) {
            int b = 0;
            while (i >= start[b + 1]) ++b;
            slot = bins[(size_t)b * bin_cap + (i - start[b])];
            const int k = W.item_of(P, slot);
            shade_vertex(W.scene(k), W.item(k), P, slot, P.hit[slot], v);
            if (v.shadow) {
                P.sh_o[slot] = v.sh_o;
                P.sh_d[slot] = v.sh_d;
                P.nee_a[slot] = v.nee_a;
                P.nee_b[slot] = v.nee_b;
            }
            if (v.cont) {
                P.o[slot] = v.o;
                P.d[slot] = v.d;
                P.beta[slot] = v.beta;
                P.rng[slot] = v.rng;
            }
            P.L[slot] = f4(v.L, 0.0f);
        }
        block_push<3>({v.cont, v.shadow, i < total && !v.cont},
                      {&ctl->q_count[nxt], &ctl->sh_count, &ctl->free_count[nxt]}, {queue_next, shq, freel}, slot);
    }
}

// Vertex results of a tile of k_shade_c, staged in LDS.
struct ShadeStage {
    float4 o[256], d[256], beta[256];
    float4 sh_o[256], sh_d[256], nee_a[256], nee_b[256];
    uint32_t rng[256], pid[256];
    int item[256];
    uint16_t inv[2][256];
};
__device__ __forceinline__ void stage_shadow(ShadeStage* stage, int slot, const float4& o, const float4& d,
                                             const float4& a, const float4& b) {
    stage->sh_o[slot] = o;
    stage->sh_d[slot] = d;
    stage->nee_a[slot] = a;
    stage->nee_b[slot] = b;
}
// One tile of positions [base, base + 256) of k_shade_c: shading into LDS
// (the vertex results wait there while the block computes the output
// positions, so they do not occupy registers across that step; the kernel
// would otherwise need about 160 VGPRs, three waves per SIMD), then the
// output positions, then the stores.
//   ORDER: the block's continuing paths and its shadow rays are each grouped
//   by the octant of their ray direction, so that the waves of the next
//   traversal, which take consecutive positions, hold rays of one octant
//   (same near-child order at every node); the stores then read the stage
//   in that order and write consecutive positions.  A path's results do not
//   depend on its position, so the film is the same either way.
template <bool BATCH, bool ORDER>
__device__ __forceinline__ void shade_tile(const Work<BATCH>& W, const PathBuf& X, const PathState& Y, float4* L_out,
                                           ShadeStage& sg, uint32_t base, uint32_t total, uint32_t& shaded,
                                           uint32_t* const (&ctr)[2], bool skip_dark) {
    const int t = (int)threadIdx.x;
    const uint32_t i = base + threadIdx.x;
    bool cont = false, shadow = false;
    uint32_t pid = 0;
    int k = 0;
    if (i < total) {
        pid = X.pid[i];
        k = W.item_of(X, i);
        const int4 h = X.hit[i];
        Vertex v;
        v.cont = false;
        v.shadow = false;
        if (h.x >= 0) {
            ++shaded;
            shade_vertex(W.scene(k), W.item(k), X, i, h, v, &L_out[pid], &sg, t, skip_dark);
        }
        cont = v.cont;
        shadow = v.shadow;
        sg.rng[t] = v.rng;
        if (cont) {
            sg.o[t] = v.o;
            sg.d[t] = v.d;
            sg.beta[t] = v.beta;
        }
    }
    if (!ORDER) {
        uint32_t pos[2];
        block_pos<2>({cont, shadow}, ctr, pos);
        if (i < total) {
            if (cont) {
                const uint32_t j = pos[0];
                Y.o[j] = sg.o[t];
                Y.d[j] = sg.d[t];
                Y.beta[j] = sg.beta[t];
                Y.rng[j] = sg.rng[t];
                Y.pid[j] = pid;
                if (BATCH) Y.item[j] = k;
            }
            if (shadow) {
                const uint32_t q = pos[1];
                const float4 sd = sg.sh_d[t], na = sg.nee_a[t];
                X.sh_o[q] = sg.sh_o[t];
                X.sh_d[q] = make_float4(sd.x, sd.y, sd.z, __int_as_float(k));
                X.nee_a[q] = make_float4(na.x, na.y, na.z, __uint_as_float(pid));
                X.nee_b[q] = sg.nee_b[t];
            }
        }
        return;
    }
    const int key[2] = {cont ? octant(sg.d[t]) : 0, shadow ? octant(sg.sh_d[t]) : 0};
    uint32_t rank[2], bpos[2], n[2];
    block_rank_oct({cont, shadow}, key, ctr, rank, bpos, n);
    sg.pid[t] = pid;
    if (BATCH) sg.item[t] = k;
    if (cont) sg.inv[0][rank[0]] = (uint16_t)t;
    if (shadow) sg.inv[1][rank[1]] = (uint16_t)t;
    __syncthreads();
    if ((uint32_t)t < n[0]) {
        const int u = sg.inv[0][t];
        const uint32_t j = bpos[0] + t;
        Y.o[j] = sg.o[u];
        Y.d[j] = sg.d[u];
        Y.beta[j] = sg.beta[u];
        Y.rng[j] = sg.rng[u];
        Y.pid[j] = sg.pid[u];
        if (BATCH) Y.item[j] = sg.item[u];
    }
    if ((uint32_t)t < n[1]) {
        const int u = sg.inv[1][t];
        const uint32_t q = bpos[1] + t;
        const float4 sd = sg.sh_d[u], na = sg.nee_a[u];
        X.sh_o[q] = sg.sh_o[u];
        X.sh_d[q] = make_float4(sd.x, sd.y, sd.z, __int_as_float(BATCH ? sg.item[u] : 0));
        X.nee_a[q] = make_float4(na.x, na.y, na.z, __uint_as_float(sg.pid[u]));
        X.nee_b[q] = sg.nee_b[u];
    }
    __syncthreads();
}

// Compact rounds: position i of the current copy X for every active path
// (misses included: they end here).  A path that continues moves to its
// compacted position in the other copy Y.  Radiance accumulates in L_out[path
// id] (emission here, NEE terms in k_trace_shadow).  Shadow rays are stored
// by their own position, with the item (batch) in sh_d.w and the path id in
// nee_a.w.  Blocks stride over the tiles and append in completion order (one
// atomic per block and output; an order-preserving scan was measured: the
// traversal of the next bounce ran no faster).  The vertices shaded are
// counted in bin_count[0].  skip_dark: no shadow ray is cast for an NEE term
// that is exactly zero (dark_term), which leaves the film unchanged.
template <bool BATCH, bool ORDER>
__global__ __launch_bounds__(256, 3) void k_shade_c(Work<BATCH> W, PathBuf X, PathState Y, Ctl* ctl, int cur,
                                                    int nxt, float4* L_out, int skip_dark) {
    __shared__ ShadeStage sg;
    const uint32_t total = ctl->q_count[cur];
    uint32_t shaded = 0;
    uint32_t* const ctr[2] = {&ctl->q_count[nxt], &ctl->shc[cur]};
    for (uint32_t base = blockIdx.x * blockDim.x; base < total; base += gridDim.x * blockDim.x)
        shade_tile<BATCH, ORDER>(W, X, Y, L_out, sg, base, total, shaded, ctr, skip_dark != 0);
    for (int off = 32; off > 0; off >>= 1) shaded += __shfl_xor(shaded, off);
    if (lane_id() == 0 && shaded) atomicAdd(&ctl->bin_count[0], shaded);
}

// ----------------------------------------------------------- trace shadow
// aggregate.Intersect(shadowRay, lightDistance - 1e-3) (PathTracing.fs:47-59);
// an unoccluded ray adds beta*f * Le/(pdf_b + pdf_l) to L.  Slot rounds: the
// rays of the slots in shq, added to the slot's L.  Compact rounds (shq
// null): ray i, added to L_out[path id] (nee_a.w).
template <bool COUNT, bool BATCH, int K>
__global__ __launch_bounds__(256, K == 16 ? (BATCH ? kTraceWavesBatch : kTraceWaves) : 1) void k_trace_shadow(Work<BATCH> W, PathBuf P, Ctl* ctl, const uint32_t* shq,
                                                      float4* L_out, int par,
                                                      uint32_t* spill, size_t spill_stride, uint32_t refill,
                                                      unsigned long long* stats, uint32_t* seg) {
    __shared__ uint32_t stk[K * 256];
    const uint32_t count = shq && par < 0 ? ctl->sh_count : ctl->shc[par];
    TraceCount tc{0, 0, 0};
    persistent_trace<true, COUNT, K>(
        SegFetch{seg, kShadowSegs}, count, stk, spill, spill_stride, refill, tc,
        [&](Traversal<K>& tr) { W.bind(tr.G); },
        [&](Traversal<K>& tr, uint32_t i) {
            const uint32_t slot = shq ? shq[i] : i;
            const float4 o4 = P.sh_o[slot], d4 = P.sh_d[slot];
            const int k = shq ? W.item_of(P, slot) : (BATCH ? __float_as_int(d4.w) : 0);
            W.bind_lane(tr.G, k);
            tr.start({xyz(o4), xyz(d4)}, o4.w, W.scene(k).exact_only != 0);
            if (!COUNT) tr.take_tlas_root();
            // a NaN t max fails every slab test of the reference: empty traversal
            if (!(o4.w == o4.w)) tr.empty();
        },
        [&](Traversal<K>& tr, uint32_t i, bool fin, bool found) {
            if (fin && !found) {
                const uint32_t slot = shq ? shq[i] : i;
                const float4 a = P.nee_a[slot];
                const V3 b = xyz(P.nee_b[slot]);
                float4* dst = shq ? &P.L[slot] : &L_out[__float_as_uint(a.w)];
                *dst = f4(fma3(xyz(a), b, xyz(*dst)), 0.0f);
            }
        });
    if (COUNT) {
        wave_add_stats(&stats[3], tc.nodes);
        wave_add_stats(&stats[4], tc.tris);
        wave_add_stats(&stats[5], tc.insts);
    }
}

// ----------------------------------------------------------------- adjoint
// Path-replay backpropagation of an image loss to per-material albedo and
// roughness and per-light emission.  Each path is re-traced from its seed with
// the forward's decisions; L holds the radiance still to come, decremented as
// the replay passes each contribution, so no vertex is stored.  Derivatives
// are taken at fixed sampled directions, light samples and Russian-roulette
// outcomes, with the importance factors 1/pdf of sampled directions held
// constant (detached sampling) and the MIS weights differentiated (they
// depend on roughness through the BSDF pdf): unbiased for the derivative of
// the expected image, visibility changes excluded.  For a contribution
// c = beta * e at a later vertex, d c / d theta of the BSDF sample at vertex j
// is c * (d f_j / d theta) / f_j, so the BSDF term of vertex j is
// delta * (d f_j / f_j) * L_after(j).  The emission MIS weight of vertex k
// depends on the pdf sampled at vertex k-1 (material and derivative carried in
// pend_a.w and beta.w); the NEE weight on the pdf at the light direction.
//
// Gradients are summed per wave for lanes sharing a key (material or light)
// before one atomic per key, and lanes with rare keys add directly.
// Per-block gradient tables in LDS (4 floats per material, then 3 per light),
// flushed to the global tables (4 floats per entry) once per block.
constexpr int kGradLdsThreads = 1024;
__device__ __forceinline__ void lds_grad_clear(float* s, int n_mat, int n_light) {
    for (int k = threadIdx.x; k < 4 * n_mat + 3 * n_light; k += blockDim.x) s[k] = 0.0f;
    __syncthreads();
}
__device__ __forceinline__ void lds_grad_mat(float* s, int key, float4 v, int n_mat) {
    if (key < 0 || key >= n_mat) return;
    float* d = s + 4 * key;
    atomicAdd(d + 0, v.x); atomicAdd(d + 1, v.y); atomicAdd(d + 2, v.z); atomicAdd(d + 3, v.w);
}
__device__ __forceinline__ void lds_grad_light(float* s, int key, float4 v, int n_mat, int n_light) {
    if (key < 0 || key >= n_light) return;
    float* d = s + 4 * n_mat + 3 * key;
    atomicAdd(d + 0, v.x); atomicAdd(d + 1, v.y); atomicAdd(d + 2, v.z);
}
__device__ __forceinline__ void lds_grad_flush(const float* s, float* g_mat, float* g_light, int n_mat, int n_light) {
    __syncthreads();
    for (int k = threadIdx.x; k < 4 * n_mat; k += blockDim.x)
        if (s[k] != 0.0f) unsafeAtomicAdd(&g_mat[k], s[k]);
    const float* sl = s + 4 * n_mat;
    for (int k = threadIdx.x; k < 3 * n_light; k += blockDim.x)
        if (sl[k] != 0.0f) unsafeAtomicAdd(&g_light[4 * (k / 3) + k % 3], sl[k]);
}

__device__ __forceinline__ float wave_sum(float v) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    return v;
}
__device__ __forceinline__ void grad_accum(float* g, int key, float4 v, bool act) {
    bool pending = act && key >= 0;
    for (int it = 0; it < 4; ++it) {
        const uint64_t m = __ballot(pending);
        if (m == 0) return;
        const int leader = __ffsll((unsigned long long)m) - 1;
        const int k = __shfl(key, leader);
        const bool mine = pending && key == k;
        const uint64_t mm = __ballot(mine);
        if (__popcll(mm) < 4) break;                     // rare keys: direct atomics below
        const float sx = wave_sum(mine ? v.x : 0.0f), sy = wave_sum(mine ? v.y : 0.0f);
        const float sz = wave_sum(mine ? v.z : 0.0f), sw = wave_sum(mine ? v.w : 0.0f);
        if ((int)lane_id() == leader) {
            float* d = g + 4 * (size_t)k;
            unsafeAtomicAdd(d + 0, sx); unsafeAtomicAdd(d + 1, sy);
            unsafeAtomicAdd(d + 2, sz); unsafeAtomicAdd(d + 3, sw);
        }
        pending = pending && !mine;
    }
    if (pending) {
        float* d = g + 4 * (size_t)key;
        unsafeAtomicAdd(d + 0, v.x); unsafeAtomicAdd(d + 1, v.y);
        unsafeAtomicAdd(d + 2, v.z); unsafeAtomicAdd(d + 3, v.w);
    }
}
__device__ __forceinline__ float ratio(float d, float f) { return f != 0.0f ? d / f : 0.0f; }
// A gradient term that is not finite (0/0 or inf * 0 of a degenerate
// configuration, e.g. a half vector perpendicular to wo in the pdf's
// derivative) is dropped instead of added: one such term would turn the whole
// parameter's gradient, and the optimiser state after it, into NaN.
__device__ __forceinline__ void drop_nonfinite(int& key, const float4& v) {
    if (!(isfinite(v.x) && isfinite(v.y) && isfinite(v.z) && isfinite(v.w))) key = -1;
}

// k_shade's bounce, replayed: emission and its emission gradient; NEE set-up
// with the terms its gradient needs (applied by k_trace_shadow_adj when the
// shadow ray is unoccluded); the BSDF sample and its term, applied here when
// no shadow ray follows and after the NEE subtraction otherwise.
// LDS: 1024-thread blocks sum the gradients in LDS (see lds_grad_*).
template <bool LDS>
__global__ __launch_bounds__(LDS ? kGradLdsThreads : 256) void k_shade_adj(
        Work<false> W, PathBuf P, AdjBuf A, Ctl* ctl, const uint32_t* bins, uint32_t bin_cap, uint32_t* queue_next,
        int nxt, uint32_t* shq, uint32_t* freel) {
    extern __shared__ float lds_g[];
    if (LDS) lds_grad_clear(lds_g, A.n_mat, A.n_light);
    uint32_t start[kBins + 1];
    start[0] = 0;
    for (int b = 0; b < kBins; ++b) start[b + 1] = start[b] + ctl->bin_count[b];
    const uint32_t total = start[kBins];
    const DevScene& S = W.scene(0);
    const RoundItem& it = W.item(0);
    for (uint32_t base = blockIdx.x * blockDim.x; base < total; base += gridDim.x * blockDim.x) {
        const uint32_t i = base + threadIdx.x;
        bool cont = false, shadow = false, fin = false;
        uint32_t slot = 0;
        int light_key = -1, mat_key = -1, prev_key = -1;
        float4 g_l = make_float4(0, 0, 0, 0), g_m = make_float4(0, 0, 0, 0), g_p = make_float4(0, 0, 0, 0);
        if (i < total) {
            int b = 0;
            while (i >= start[b + 1]) ++b;
            slot = bins[(size_t)b * bin_cap + (i - start[b])];
            const float4 o4 = P.o[slot], d4 = P.d[slot];
            const Ray r = {xyz(o4), xyz(d4)};
            const float prev_pdf = d4.w;
            const int depth = __float_as_int(o4.w);
            const int4 h = P.hit[slot];
            const Geom g = hit_geometry(S, r, h.x, h.y, __int_as_float(h.z));
            const int4 info = S.insts[h.x].info;
            V3 L = xyz(P.L[slot]);
            const float4 b4 = P.beta[slot];
            V3 beta = xyz(b4);
            const V3 dlt = xyz(A.delta[slot]);
            if (info.w >= 0) {
                V3 Le;
                float lpdf;
                light_eval(S, h.x, r.o, g, Le, lpdf);
                const float w = depth == 0 ? 1.0f : prev_pdf * rcp_e(lpdf + prev_pdf);
                L = L - beta * (Le * w);
                const float4 def = S.lightdefs[info.w];
                const float woz = dot(normalize(r.o - g.p), g.n);
                const float facing = fabsf(woz) > 1e-6f && (woz > 0.0f || def.w != 0.0f) ? 1.0f : 0.0f;
                const V3 gl = dlt * beta * (w * facing);
                g_l = f4(gl, 0.0f);
                light_key = info.w;
                if (depth > 0) {
                    // w = prev / (lpdf + prev): d w / d prev = lpdf / (lpdf + prev)^2
                    const float sd = lpdf + prev_pdf;
                    const float dw = lpdf / (sd * sd) * b4.w;
                    prev_key = __float_as_int(A.pend_a[slot].w);
                    g_p = make_float4(0.0f, 0.0f, 0.0f, dot(dlt * beta, Le) * dw);
                }
            }
            if (info.z >= 0) {
                uint32_t rng = P.rng[slot];
                const Mat m = load_mat(S, info.z);
                const float dr = dalpha_drough(m);
                const V3 wo = -r.d;
                const V3 wo_l = to_local(g, wo);
                const float uSel = lcg(rng);
                V2 uL;
                uL.x = lcg(rng); uL.y = lcg(rng);
                const LightSample ls = light_sample(S, g.p, uSel, uL);
                if (ls.pdf != 0.0f) {
                    const BSDFGrad e = mat_eval_grad(m, wo_l, to_local(g, ls.wi));
                    const float dist = length(ls.p - g.p);
                    const float rc = rcp_e(e.pdf + ls.pdf);
                    P.sh_o[slot] = f4(g.p, dist - 1e-3f);
                    P.sh_d[slot] = f4(ls.wi, e.dpdf * dr * rc);        // d pdf_b / d rough / (pdf_b + pdf_l)
                    P.nee_a[slot] = f4(beta * e.f, 0.0f);
                    P.nee_b[slot] = f4(ls.L * rc, rc * ls.facing);
                    A.nee_da[slot] = f4(beta * e.dalb, __int_as_float(info.z));
                    A.nee_dr[slot] = f4(beta * e.dalpha * dr, __int_as_float(ls.ldef));
                    shadow = true;
                }
                const float ul = lcg(rng);
                V2 ub;
                ub.x = lcg(rng); ub.y = lcg(rng);
                V3 wi_l;
                const BSDFGrad bs = mat_sample_grad(m, wo_l, ul, ub, wi_l);
                V3 pa = mk(0, 0, 0), pr = mk(0, 0, 0);
                if (bs.pdf != 0.0f) {
                    const V3 wi = to_world(g, wi_l);
                    beta = beta * bs.f * rcp_e(bs.pdf);
                    cont = true;
                    if (depth >= it.rr_depth) {
                        const float q = min_n(1.0f, max_n(beta.x, max_n(beta.y, beta.z)));
                        if (lcg(rng) < q) beta = beta * rcp_e(q);
                        else cont = false;
                    }
                    P.o[slot] = f4(g.p, __int_as_float(depth + 1));
                    P.d[slot] = f4(wi, bs.pdf);
                    P.beta[slot] = f4(beta, bs.dpdf * dr);                // d prev pdf / d rough
                    pa = {ratio(bs.dalb.x, bs.f.x), ratio(bs.dalb.y, bs.f.y), ratio(bs.dalb.z, bs.f.z)};
                    pr = {ratio(bs.dalpha.x * dr, bs.f.x), ratio(bs.dalpha.y * dr, bs.f.y),
                          ratio(bs.dalpha.z * dr, bs.f.z)};
                }
                P.rng[slot] = rng;
                cont = cont && depth + 1 < it.max_depth;
                const int pend = cont ? info.z : -1;           // no later radiance: no term
                A.pend_a[slot] = f4(pa, __int_as_float(pend));   // w: also the next vertex's prev material
                A.pend_r[slot] = f4(pr, 0.0f);
                if (!shadow && pend >= 0) {
                    const V3 t = dlt * pa * L;
                    g_m = make_float4(t.x, t.y, t.z, dot(dlt * pr, L));
                    mat_key = pend;
                }
            }
            P.L[slot] = f4(L, 0.0f);
            fin = !cont;
        }
        drop_nonfinite(light_key, g_l);
        drop_nonfinite(prev_key, g_p);
        drop_nonfinite(mat_key, g_m);
        if (LDS) {
            lds_grad_light(lds_g, light_key, g_l, A.n_mat, A.n_light);
            lds_grad_mat(lds_g, prev_key, g_p, A.n_mat);
            lds_grad_mat(lds_g, mat_key, g_m, A.n_mat);
        } else {
            grad_accum(A.g_light, light_key, g_l, light_key >= 0);
            grad_accum(A.g_mat, prev_key, g_p, prev_key >= 0);
            grad_accum(A.g_mat, mat_key, g_m, mat_key >= 0);
        }
        block_push<3, LDS ? kGradLdsThreads / 64 : 4>({cont, shadow, fin},
                                                      {&ctl->q_count[nxt], &ctl->sh_count, &ctl->free_count[nxt]},
                                                      {queue_next, shq, freel}, slot);
    }
    if (LDS) lds_grad_flush(lds_g, A.g_mat, A.g_light, A.n_mat, A.n_light);
}

// The shadow rays of a replayed bounce only record their visibility (in
// nee_a.w): gradient sums belong in a kernel whose lanes are all busy, not in
// the traversal loop.
template <int K>
__global__ __launch_bounds__(256, K == 16 ? kTraceWaves : 1) void k_trace_shadow_vis(Work<false> W, PathBuf P, Ctl* ctl, const uint32_t* shq,
                                                          uint32_t* spill, size_t spill_stride, uint32_t refill,
                                                          uint32_t* seg) {
    __shared__ uint32_t stk[K * 256];
    const uint32_t count = ctl->sh_count;
    TraceCount tc{0, 0, 0};
    persistent_trace<true, false, K>(
        SegFetch{seg, kShadowSegs}, count, stk, spill, spill_stride, refill, tc,
        [&](Traversal<K>& tr) { W.bind(tr.G); },
        [&](Traversal<K>& tr, uint32_t i) {
            const uint32_t slot = shq[i];
            const float4 o4 = P.sh_o[slot];
            tr.start({xyz(o4), xyz(P.sh_d[slot])}, o4.w, W.S.exact_only != 0);
            tr.take_tlas_root();
            if (!(o4.w == o4.w)) tr.empty();
        },
        [&](Traversal<K>& tr, uint32_t i, bool fin, bool found) {
            if (fin) P.nee_a[shq[i]].w = found ? 0.0f : 1.0f;
        });
}

// Then, per shadow ray: when unoccluded, the NEE contribution leaves the
// radiance still to come and adds its gradients (the vertex's material
// through f and the NEE MIS weight, the sampled light through its emission);
// then the vertex's BSDF-sample term is applied.
__device__ __forceinline__ void nee_adj_item(const PathBuf& P, const AdjBuf& A, uint32_t slot, int& mat_key,
                                             int& light_key, float4& g_m, float4& g_l) {
    V3 L = xyz(P.L[slot]);
    const V3 dlt = xyz(A.delta[slot]);
    const float4 a = P.nee_a[slot];
    V3 gm = mk(0, 0, 0);
    float gr = 0.0f;
    if (a.w != 0.0f) {
        const float4 b = P.nee_b[slot];
        const float4 da = A.nee_da[slot], dr = A.nee_dr[slot];
        const V3 c = xyz(a) * xyz(b);
        L = L - c;
        gm = dlt * xyz(da) * xyz(b);
        // d (f Le / (pdf_b + pdf_l)) / d rough = f' Le / S - c pdf_b' / S
        gr = dot(dlt, xyz(dr) * xyz(b) - c * P.sh_d[slot].w);
        mat_key = __float_as_int(da.w);
        light_key = __float_as_int(dr.w);
        g_l = f4(dlt * xyz(a) * b.w, 0.0f);
    }
    const float4 pa = A.pend_a[slot];
    const int pend = __float_as_int(pa.w);
    if (pend >= 0) {
        gm = gm + dlt * xyz(pa) * L;
        gr += dot(dlt * xyz(A.pend_r[slot]), L);
        mat_key = pend;
    }
    g_m = f4(gm, gr);
    P.L[slot] = f4(L, 0.0f);
    drop_nonfinite(mat_key, g_m);
    drop_nonfinite(light_key, g_l);
}

__global__ __launch_bounds__(256) void k_nee_adj(PathBuf P, AdjBuf A, const Ctl* ctl, const uint32_t* shq) {
    const uint32_t count = ctl->sh_count;
    for (uint32_t base = blockIdx.x * blockDim.x; base < count; base += gridDim.x * blockDim.x) {
        const uint32_t i = base + threadIdx.x;
        int mat_key = -1, light_key = -1;
        float4 g_m = make_float4(0, 0, 0, 0), g_l = make_float4(0, 0, 0, 0);
        if (i < count) nee_adj_item(P, A, shq[i], mat_key, light_key, g_m, g_l);
        grad_accum(A.g_mat, mat_key, g_m, mat_key >= 0);
        grad_accum(A.g_light, light_key, g_l, light_key >= 0);
    }
}

// The same with the gradients of the whole grid's share summed in LDS first:
// device-scope float atomics on a small table serialise, LDS atomics do not.
// One 1024-thread block per CU; used when the tables fit a workgroup's LDS.
__global__ __launch_bounds__(kGradLdsThreads) void k_nee_adj_lds(PathBuf P, AdjBuf A, const Ctl* ctl,
                                                                const uint32_t* shq) {
    extern __shared__ float lds_g[];
    lds_grad_clear(lds_g, A.n_mat, A.n_light);
    const uint32_t count = ctl->sh_count;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < count; i += gridDim.x * blockDim.x) {
        int mat_key = -1, light_key = -1;
        float4 g_m = make_float4(0, 0, 0, 0), g_l = make_float4(0, 0, 0, 0);
        nee_adj_item(P, A, shq[i], mat_key, light_key, g_m, g_l);
        lds_grad_mat(lds_g, mat_key, g_m, A.n_mat);
        lds_grad_light(lds_g, light_key, g_l, A.n_mat, A.n_light);
    }
    lds_grad_flush(lds_g, A.g_mat, A.g_light, A.n_mat, A.n_light);
}

// -------------------------------------------------------------------- film
// accum = FMA(1/spp, Li / pdf, accum) in sample order (Integrator.fs:41-42);
// the camera pdf is 1.  The film keeps Film.Pixels order (Film.fs:43).  One
// thread handles one pixel of one item of the round (items are ordered by
// their cumulative pixel end pix_end).
template <bool BATCH>
__global__ void k_film(Work<BATCH> W, const float4* L_out, float* film) {
    const int n_items = W.n();
    const uint32_t n_pix = (uint32_t)W.item(n_items - 1).pix_end;
    for (uint32_t p = blockIdx.x * blockDim.x + threadIdx.x; p < n_pix; p += gridDim.x * blockDim.x) {
        int k = 0;
        if (BATCH) {
            int lo = 0, hi = n_items - 1;
            while (lo < hi) {
                const int mid = (lo + hi) >> 1;
                if ((uint32_t)W.item(mid).pix_end <= p) lo = mid + 1; else hi = mid;
            }
            k = lo;
        }
        const RoundItem& c = W.item(k);
        const uint32_t lp = p - (uint32_t)(c.pix_end - c.width * c.height);
        const int x = (int)(lp % (uint32_t)c.width), y = (int)(lp / (uint32_t)c.width);
        const uint32_t q =
            ((uint32_t)((y >> 3) * c.tiles_x + (x >> 3)) << 6) | ((uint32_t)(y & 7) << 3) | (uint32_t)(x & 7);
        float* f = film + c.film_offset + 3 * ((size_t)(c.height - y - 1) * c.width + x);
        float ax = f[0], ay = f[1], az = f[2];
        for (int s = 0; s < c.ns; ++s) {
            const float4 L = L_out[(size_t)c.path_base + (size_t)s * c.p_pad + q];
            ax = fmaf(c.inv_spp, L.x * rcp_e(1.0f), ax);
            ay = fmaf(c.inv_spp, L.y * rcp_e(1.0f), ay);
            az = fmaf(c.inv_spp, L.z * rcp_e(1.0f), az);
        }
        f[0] = ax; f[1] = ay; f[2] = az;
    }
}

// NormalIntegrator.Li (Extensions/Integrator/Normal.fs:10-17) on the bins of
// the first bounce: 0.5 * (n + 1); every path ends here.
__global__ void k_shade_normal(Work<false> W, PathBuf P, Ctl* ctl, const uint32_t* bins, uint32_t bin_cap, int nxt,
                               uint32_t* freel) {
    uint32_t start[kBins + 1];
    start[0] = 0;
    for (int b = 0; b < kBins; ++b) start[b + 1] = start[b] + ctl->bin_count[b];
    const uint32_t total = start[kBins];
    for (uint32_t base = blockIdx.x * blockDim.x; base < total; base += gridDim.x * blockDim.x) {
        const uint32_t i = base + threadIdx.x;
        uint32_t slot = 0;
        if (i < total) {
            int b = 0;
            while (i >= start[b + 1]) ++b;
            slot = bins[(size_t)b * bin_cap + (i - start[b])];
            const Ray r = {xyz(P.o[slot]), xyz(P.d[slot])};
            const int4 h = P.hit[slot];
            const Geom g = hit_geometry(W.S, r, h.x, h.y, __int_as_float(h.z));
            P.L[slot] = f4(0.5f * (g.n + mk(1, 1, 1)), 0.0f);
        }
        block_push<1>({i < total}, {&ctl->free_count[nxt]}, {freel}, slot);
    }
}

// DirectIntegrator.Li (Extensions/Integrator/Direct.fs:10-40) on the bins of
// a slot round.  Depth 0 (camera hit): emission seen directly, L + Le with no
// MIS weight; at a material, a light sample whose unweighted term
// f * Le * (1 / pdf_l) lands through the shadow ray (cast whatever the pdf,
// as the reference does), then a BSDF sample that continues as the depth-1
// ray carrying f in beta.  Depth 1: a light hit adds f * Le * (1 / pdf_l),
// pdf_l the light sampler's Eval at the ray origin; the path ends.  The
// shadow term of depth 0 lands before the depth-1 term, as in the reference.
__global__ void k_shade_direct(Work<false> W, PathBuf P, Ctl* ctl, const uint32_t* bins, uint32_t bin_cap,
                               uint32_t* queue_next, int nxt, uint32_t* shq, uint32_t* freel) {
    uint32_t start[kBins + 1];
    start[0] = 0;
    for (int b = 0; b < kBins; ++b) start[b + 1] = start[b] + ctl->bin_count[b];
    const uint32_t total = start[kBins];
    const DevScene& S = W.S;
    for (uint32_t base = blockIdx.x * blockDim.x; base < total; base += gridDim.x * blockDim.x) {
        const uint32_t i = base + threadIdx.x;
        uint32_t slot = 0;
        bool cont = false, shadow = false;
        if (i < total) {
            int b = 0;
            while (i >= start[b + 1]) ++b;
            slot = bins[(size_t)b * bin_cap + (i - start[b])];
            const float4 o4 = P.o[slot];
            const Ray r = {xyz(o4), xyz(P.d[slot])};
            const int depth = __float_as_int(o4.w);
            const int4 h = P.hit[slot];
            const Geom g = hit_geometry(S, r, h.x, h.y, __int_as_float(h.z));
            const int4 info = S.insts[h.x].info;
            V3 L = xyz(P.L[slot]);
            if (depth == 0) {
                if (info.w >= 0) L = L + light_emit(S.lightdefs[info.w], dot(-r.d, g.n));
                if (info.z >= 0) {
                    uint32_t rng = P.rng[slot];
                    const Mat m = load_mat(S, info.z);
                    const V3 wo_l = to_local(g, -r.d);
                    const float uSel = lcg(rng);
                    V2 uL;
                    uL.x = lcg(rng); uL.y = lcg(rng);
                    const LightSample ls = light_sample(S, g.p, uSel, uL);
                    const BSDF e = mat_eval(m, wo_l, to_local(g, ls.wi));
                    const float dist = length(ls.p - g.p);
                    P.sh_o[slot] = f4(g.p, dist - 1e-3f);
                    P.sh_d[slot] = f4(ls.wi, 0.0f);
                    P.nee_a[slot] = f4(e.f, 0.0f);
                    P.nee_b[slot] = f4(ls.L * (1.0f / ls.pdf), 0.0f);
                    shadow = true;
                    const float ul = lcg(rng);
                    V2 ub;
                    ub.x = lcg(rng); ub.y = lcg(rng);
                    V3 wi_l;
                    const BSDF bs = mat_sample(m, wo_l, ul, ub, wi_l);
                    P.o[slot] = f4(g.p, __int_as_float(1));
                    P.d[slot] = f4(to_world(g, wi_l), 0.0f);
                    P.beta[slot] = f4(bs.f, 0.0f);
                    P.rng[slot] = rng;
                    cont = true;
                }
            } else if (info.w >= 0) {
                V3 Le;
                float lpdf;
                light_eval(S, h.x, r.o, g, Le, lpdf);
                L = fma3(xyz(P.beta[slot]), Le * (1.0f / lpdf), L);
            }
            P.L[slot] = f4(L, 0.0f);
        }
        block_push<3>({cont, shadow, i < total && !cont}, {&ctl->q_count[nxt], &ctl->sh_count, &ctl->free_count[nxt]},
                      {queue_next, shq, freel}, slot);
    }
}

// First hit of one sample per pixel: (instance, primitive, t) in pixel order.
__global__ __launch_bounds__(256) void k_first_hit(DevScene S, int width, int height, int sample_id, int spp,
                                                   int frame_id, int* out_inst, int* out_prim, float* out_t,
                                                   uint32_t* spill, size_t spill_stride, uint32_t* fetch) {
    __shared__ uint32_t stk[kStackK * 256];
    const uint32_t n = (uint32_t)width * (uint32_t)height;
    TraceCount tc{0, 0, 0};
    persistent_trace<false, false, kStackK>(
        SegFetch{fetch, 1}, n, stk, spill, spill_stride, 16u, tc,
        [&](Traversal<kStackK>& tr) { tr.G.set(S); },
        [&](Traversal<kStackK>& tr, uint32_t p) {
            const int x = (int)(p % (uint32_t)width), y = (int)(p / (uint32_t)width);
            uint32_t rng = xxhash3((uint32_t)x, (uint32_t)y, (uint32_t)(frame_id * spp + sample_id));
            V2 uPix, uLens;
            uPix.x = lcg(rng); uPix.y = lcg(rng);
            uLens.x = lcg(rng); uLens.y = lcg(rng);
            tr.start(camera_ray(S, x, y, uPix, uLens), kInf, S.exact_only != 0);
            tr.take_tlas_root();
        },
        [&](Traversal<kStackK>& tr, uint32_t p, bool fin, bool) {
            if (fin) {
                out_inst[p] = tr.hit_inst;
                out_prim[p] = tr.hit_prim;
                out_t[p] = tr.hit_inst >= 0 ? tr.t : kInf;
            }
        });
}

// ------------------------------------------------------------------ launch
static Work<false> plain_work(const Launch& L) {
    Work<false> w;
    w.S = L.scene;
    w.it = L.item;
    return w;
}
static Work<true> batch_work(const Launch& L) {
    Work<true> w;
    w.hdr = L.hdr;
    w.items = L.items;
    w.scenes = L.scenes;
    return w;
}

// Trace launches dispatched over the LDS stack window (16 or 32 entries).
template <bool BATCH>
static void launch_closest_t(bool count, int k, const Work<BATCH>& W, const PathBuf& P, Ctl* ctl, int cur,
                             const uint32_t* queue, const Buffers& B, uint32_t refill, int grid, hipStream_t st) {
#define BN_TC(C, KK)                                                                                            \
    hipLaunchKernelGGL((k_trace_closest<C, BATCH, KK>), dim3(grid), dim3(256), 0, st, W, P, ctl, cur, queue, \
                       B.spill, B.spill_stride, refill, B.stats, B.fetch_seg)
    if (k == 16) { if (count) BN_TC(true, 16); else BN_TC(false, 16); }
    else { if (count) BN_TC(true, 32); else BN_TC(false, 32); }
#undef BN_TC
}
// par: the parity of a compact round's iteration (its shadow count and fetch
// counters); spill: the stack spill area to use.
template <bool BATCH>
static void launch_shadow_t(bool count, int k, const Work<BATCH>& W, const PathBuf& P, Ctl* ctl, const uint32_t* shq,
                            float4* L_out, const Buffers& B, uint32_t refill, int grid, hipStream_t st, int par = -1,
                            uint32_t* spill = nullptr) {
    uint32_t* seg = B.fetch_seg + (size_t)(kFetchSegs + (par < 0 ? 0 : par) * kShadowSegs) * kSegStride;
    if (!spill) spill = B.spill;
#define BN_TS(C, KK)                                                                                            \
    hipLaunchKernelGGL((k_trace_shadow<C, BATCH, KK>), dim3(grid), dim3(256), 0, st, W, P, ctl, shq, L_out, par, \
                       spill, B.spill_stride, refill, B.stats, seg)
    if (k == 16) { if (count) BN_TS(true, 16); else BN_TS(false, 16); }
    else { if (count) BN_TS(true, 32); else BN_TS(false, 32); }
#undef BN_TS
}

struct Bracket {
    LaunchHook* h;
    hipStream_t s;
    Bracket(LaunchHook* hk, const char* name, hipStream_t st_) : h(hk), s(st_) { if (h) h->begin(name, s); }
    ~Bracket() { if (h) h->end(s); }
};

template <bool BATCH>
static void round_begin_t(const Launch& L, const PathBuf& P, Ctl* ctl, const Buffers& B, LaunchHook* hook) {
    hipStream_t st = L.stream;
    { Bracket b(hook, "ctl_clear", st); hipLaunchKernelGGL(k_ctl_clear, dim3(1), dim3(64), 0, st, ctl); }
    { Bracket b(hook, "init_slots", st);
      hipLaunchKernelGGL(k_init_slots, dim3(L.grid_shade), dim3(256), 0, st, P, ctl, B.freel[0], L.n_slots); }
}

// Path-tracing rounds run compact unless material binning is forced.
static bool compact_round(Integrator kind, const Buffers& B) {
    return kind == Integrator::Path && B.alt.o != nullptr && B.sort != 1;
}
static PathBuf with_state(PathBuf P, const PathState& S) {
    P.o = S.o; P.d = S.d; P.beta = S.beta; P.L = S.L;
    P.rng = S.rng; P.item = S.item; P.pid = S.pid;
    return P;
}
static PathState state_of(const PathBuf& P) { return PathState{P.o, P.d, P.beta, P.L, P.rng, P.item, P.pid}; }

template <bool BATCH>
static void compact_iterations_t(const Work<BATCH>& W, const Launch& L, const PathBuf& P, Ctl* ctl,
                                 const Buffers& B, int n, bool count, LaunchHook* hook) {
    hipStream_t st = L.stream;
    const PathBuf copies[2] = {P, with_state(P, B.alt)};
    // The trace kernels read only the geometry of a ray's scene, so a batch
    // round over one scene traces with the single-scene kernels.
    const bool one_scene = BATCH && L.uniform;
    // regen_c and shade_c run one resident set of blocks (three 256-thread
    // blocks per CU, shade_c's launch bound: 139 VGPRs, no spills), each
    // striding over the tiles of positions.  With twice as many blocks the
    // traversal of the next bounce, which reads the positions in the order
    // the blocks appended them, was measured 1.6% (config 1) to 3.6%
    // (32-scene batch) slower; four blocks per CU at four waves per SIMD
    // (6 VGPRs spilled) made shade_c 0.7 ms faster per config-1 frame but
    // the two-stream frame 2.4 ms slower.
    const int grid_c = L.grid_shade * 3 / 8;
    // With a second stream, the shadow rays of iteration i are traced while
    // regen, prep and trace_closest of iteration i + 1 run: shade_c(i + 1)
    // waits for trace_shadow(i) (it reuses the shadow arrays and adds to the
    // same radiance slots), and the block ends with both streams joined.
    const bool overlap = L.stream2 != nullptr;
    hipStream_t st2 = overlap ? L.stream2 : st;
    uint32_t* spill_sh = overlap ? B.spill2 : B.spill;
    for (int i = 0; i < n; ++i) {
        const int cur = i & 1, nxt = 1 - cur;
        const PathBuf& X = copies[cur];
        const PathBuf& Y = copies[nxt];
        { Bracket b(hook, "regen", st);
          hipLaunchKernelGGL(k_regen_c<BATCH>, dim3(grid_c), dim3(256), 0, st, W, X, ctl, cur, L.n_slots,
                             B.L_out); }
        { Bracket b(hook, "prep", st);
          hipLaunchKernelGGL(k_prep<BATCH>, dim3(1), dim3(2 * kFetchSegs), 0, st, W, ctl, cur, B.fetch_seg,
                             L.n_slots); }
        { Bracket b(hook, "trace_closest", st);
          if (one_scene)
              launch_closest_t<false>(count, B.stack_k, plain_work(L), X, ctl, cur, nullptr, B, L.refill,
                                      L.grid_uniform, st);
          else
              launch_closest_t<BATCH>(count, B.stack_k, W, X, ctl, cur, nullptr, B, L.refill, L.grid_trace, st); }
        if (overlap) (void)hipStreamWaitEvent(st, L.ev_shadow, 0);
        { Bracket b(hook, "shade", st);
          const int skip = L.skip_dark == 2 || (L.skip_dark == 1 && !count);
          if (L.octant_order)
              hipLaunchKernelGGL((k_shade_c<BATCH, true>), dim3(grid_c), dim3(256), 0, st, W, X, state_of(Y), ctl,
                                 cur, nxt, B.L_out, skip);
          else
              hipLaunchKernelGGL((k_shade_c<BATCH, false>), dim3(grid_c), dim3(256), 0, st, W, X, state_of(Y), ctl,
                                 cur, nxt, B.L_out, skip); }
        if (overlap) {
            (void)hipEventRecord(L.ev_shade, st);
            (void)hipStreamWaitEvent(st2, L.ev_shade, 0);
        }
        { Bracket b(hook, "trace_shadow", st2);
          if (one_scene)
              launch_shadow_t<false>(count, B.stack_k, plain_work(L), X, ctl, nullptr, B.L_out, B, L.refill,
                                     L.grid_uniform, st2, cur, spill_sh);
          else
              launch_shadow_t<BATCH>(count, B.stack_k, W, X, ctl, nullptr, B.L_out, B, L.refill, L.grid_trace,
                                     st2, cur, spill_sh); }
        if (overlap) (void)hipEventRecord(L.ev_shadow, st2);
        if (hook) hook->after_bounce(cur);
    }
    if (overlap) (void)hipStreamWaitEvent(st, L.ev_shadow, 0);
}

template <bool BATCH>
static void iterations_t(const Work<BATCH>& W, Integrator kind, const Launch& L, const PathBuf& P, Ctl* ctl,
                         const Buffers& B, int n, bool count, LaunchHook* hook) {
    if (compact_round(kind, B)) {
        compact_iterations_t<BATCH>(W, L, P, ctl, B, n, count, hook);
        return;
    }
    hipStream_t st = L.stream;
    for (int i = 0; i < n; ++i) {
        const int cur = i & 1, nxt = 1 - cur;
        { Bracket b(hook, "regen", st);
          if (kind == Integrator::Adjoint)
              hipLaunchKernelGGL((k_regen<BATCH, true>), dim3(L.grid_shade), dim3(256), 0, st, W, P, ctl, cur,
                                 B.freel[cur], B.queue[cur], B.L_out, B.adj);
          else
              hipLaunchKernelGGL((k_regen<BATCH, false>), dim3(L.grid_shade), dim3(256), 0, st, W, P, ctl, cur,
                                 B.freel[cur], B.queue[cur], B.L_out, B.adj); }
        { Bracket b(hook, "prep", st);
          hipLaunchKernelGGL(k_prep<BATCH>, dim3(1), dim3(2 * kFetchSegs), 0, st, W, ctl, cur, B.fetch_seg, 0u); }
        { Bracket b(hook, "trace_closest", st);
          launch_closest_t<BATCH>(count, B.stack_k, W, P, ctl, cur, B.queue[cur], B, L.refill, L.grid_trace, st); }
        { Bracket b(hook, "bin", st);
          hipLaunchKernelGGL(k_bin, dim3(L.grid_shade), dim3(256), 0, st, P, ctl, cur, B.queue[cur], B.bins,
                             B.bin_cap, B.freel[nxt], B.sort < 0 ? kind == Integrator::Adjoint : B.sort); }
        if (kind == Integrator::Normal) {
            if constexpr (!BATCH) {
                Bracket b(hook, "shade_normal", st);
                hipLaunchKernelGGL(k_shade_normal, dim3(L.grid_shade), dim3(256), 0, st, W, P, ctl, B.bins, B.bin_cap,
                                   nxt, B.freel[nxt]);
            }
        } else if (kind == Integrator::Direct) {
            if constexpr (!BATCH) {
                { Bracket b(hook, "shade_direct", st);
                  hipLaunchKernelGGL(k_shade_direct, dim3(L.grid_shade), dim3(256), 0, st, W, P, ctl, B.bins,
                                     B.bin_cap, B.queue[nxt], nxt, B.shq, B.freel[nxt]); }
                { Bracket b(hook, "trace_shadow", st);
                  launch_shadow_t<false>(count, B.stack_k, W, P, ctl, B.shq, nullptr, B, L.refill,
                                         L.grid_trace, st); }
            }
        } else if (kind == Integrator::Adjoint) {
            if constexpr (!BATCH) {
                { Bracket b(hook, "shade_adj", st);
                  if (B.adj.lds_bytes > 0) {
                      hipFuncSetAttribute((const void*)k_shade_adj<true>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                          B.adj.lds_bytes);
                      hipLaunchKernelGGL(k_shade_adj<true>, dim3(L.grid_shade / 8), dim3(kGradLdsThreads),
                                         B.adj.lds_bytes, st, W, P, B.adj, ctl, B.bins, B.bin_cap, B.queue[nxt], nxt,
                                         B.shq, B.freel[nxt]);
                  } else {
                      hipLaunchKernelGGL(k_shade_adj<false>, dim3(L.grid_shade), dim3(256), 0, st, W, P, B.adj, ctl,
                                         B.bins, B.bin_cap, B.queue[nxt], nxt, B.shq, B.freel[nxt]);
                  } }
                { Bracket b(hook, "trace_shadow", st);
                  uint32_t* seg = B.fetch_seg + kFetchSegs * kSegStride;
                  if (B.stack_k == 16)
                      hipLaunchKernelGGL(k_trace_shadow_vis<16>, dim3(L.grid_trace), dim3(256), 0, st, W, P, ctl,
                                         B.shq, B.spill, B.spill_stride, L.refill, seg);
                  else
                      hipLaunchKernelGGL(k_trace_shadow_vis<32>, dim3(L.grid_trace), dim3(256), 0, st, W, P, ctl,
                                         B.shq, B.spill, B.spill_stride, L.refill, seg); }
                { Bracket b(hook, "nee_adj", st);
                  if (B.adj.lds_bytes > 0) {
                      hipFuncSetAttribute((const void*)k_nee_adj_lds, hipFuncAttributeMaxDynamicSharedMemorySize,
                                          B.adj.lds_bytes);
                      hipLaunchKernelGGL(k_nee_adj_lds, dim3(L.grid_shade / 8), dim3(kGradLdsThreads),
                                         B.adj.lds_bytes, st, P, B.adj, ctl, B.shq);
                  } else {
                      hipLaunchKernelGGL(k_nee_adj, dim3(L.grid_shade), dim3(256), 0, st, P, B.adj, ctl, B.shq);
                  } }
            }
        } else {
            { Bracket b(hook, "shade", st);
              hipLaunchKernelGGL(k_shade<BATCH>, dim3(L.grid_shade), dim3(256), 0, st, W, P, ctl, B.bins, B.bin_cap,
                                 B.queue[nxt], nxt, B.shq, B.freel[nxt]); }
            { Bracket b(hook, "trace_shadow", st);
              launch_shadow_t<BATCH>(count, B.stack_k, W, P, ctl, B.shq, nullptr, B, L.refill, L.grid_trace,
                                     st); }
        }
        if (hook) hook->after_bounce(cur);
    }
}

template <bool BATCH>
static void round_end_t(const Work<BATCH>& W, Integrator kind, const Launch& L, const PathBuf& P, Ctl* ctl,
                        const Buffers& B, LaunchHook* hook) {
    hipStream_t st = L.stream;
    // iteration blocks are even: the last paths sit in free list 0 (a compact
    // round has retired every path in its shading and shadow kernels)
    if (kind == Integrator::Adjoint) return;               // nothing to retire, no film
    if (!compact_round(kind, B)) {
        Bracket b(hook, "regen", st);
        hipLaunchKernelGGL((k_regen<BATCH, false>), dim3(L.grid_shade), dim3(256), 0, st, W, P, ctl, 0, B.freel[0],
                           B.queue[0], B.L_out, B.adj);
    }
    { Bracket b(hook, "film", st);
      hipLaunchKernelGGL(k_film<BATCH>, dim3(BATCH ? L.grid_shade : (L.n_pix + 255) / 256), dim3(256), 0, st, W,
                         B.L_out, B.film); }
}

void enqueue_round_begin(bool batch, const Launch& L, const PathBuf& P, Ctl* ctl, const Buffers& B,
                         LaunchHook* hook) {
    if (batch) round_begin_t<true>(L, P, ctl, B, hook);
    else round_begin_t<false>(L, P, ctl, B, hook);
}
void enqueue_iterations(bool batch, Integrator kind, const Launch& L, const PathBuf& P, Ctl* ctl, const Buffers& B,
                        int n, bool count, LaunchHook* hook) {
    if (batch) iterations_t<true>(batch_work(L), kind, L, P, ctl, B, n, count, hook);
    else iterations_t<false>(plain_work(L), kind, L, P, ctl, B, n, count, hook);
}
void enqueue_round_end(bool batch, Integrator kind, const Launch& L, const PathBuf& P, Ctl* ctl, const Buffers& B,
                       LaunchHook* hook) {
    if (batch) round_end_t<true>(batch_work(L), kind, L, P, ctl, B, hook);
    else round_end_t<false>(plain_work(L), kind, L, P, ctl, B, hook);
}

// Single kernels used by the PSSMLT driver (plain scene).
void launch_bin(const PathBuf& P, Ctl* ctl, int cur, const uint32_t* queue, uint32_t* bins, uint32_t bin_cap,
                uint32_t* freel, int sort, int grid, hipStream_t st) {
    hipLaunchKernelGGL(k_bin, dim3(grid), dim3(256), 0, st, P, ctl, cur, queue, bins, bin_cap, freel, sort);
}
void launch_reset(Ctl* ctl, int nxt, uint32_t* fetch_seg, hipStream_t st, int par) {
    hipLaunchKernelGGL(k_reset, dim3(1), dim3(2 * kFetchSegs), 0, st, ctl, nxt, fetch_seg, par);
}
void launch_ctl_clear(Ctl* ctl, hipStream_t st) {
    hipLaunchKernelGGL(k_ctl_clear, dim3(1), dim3(64), 0, st, ctl);
}
void launch_trace_closest(bool count, const DevScene& S, const PathBuf& P, Ctl* ctl, int cur, const uint32_t* queue,
                          const Buffers& B, uint32_t refill, int grid, hipStream_t st) {
    Work<false> W;
    W.S = S;
    W.it = RoundItem{};
    launch_closest_t<false>(count, B.stack_k, W, P, ctl, cur, queue, B, refill, grid, st);
}
void launch_trace_shadow(bool count, const DevScene& S, const PathBuf& P, Ctl* ctl, const uint32_t* shq,
                         const Buffers& B, uint32_t refill, int grid, hipStream_t st, int par, uint32_t* spill) {
    Work<false> W;
    W.S = S;
    W.it = RoundItem{};
    launch_shadow_t<false>(count, B.stack_k, W, P, ctl, shq, nullptr, B, refill, grid, st, par, spill);
}
void launch_first_hit(const DevScene& S, int w, int h, int sample_id, int spp, int frame_id, int* inst, int* prim,
                      float* t, uint32_t* spill, size_t spill_stride, uint32_t* fetch, int grid, hipStream_t st) {
    hipLaunchKernelGGL(k_first_hit, dim3(grid), dim3(256), 0, st, S, w, h, sample_id, spp, frame_id, inst, prim, t,
                       spill, spill_stride, fetch);
}

}  // namespace bnd
