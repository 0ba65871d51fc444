// kernels_mlt.hip -- primary sample space MLT (PSSMLTIntegrator, Extensions/
// Integrator/PSSMLT.fs) as a wavefront over many chains.
//
// Every chain advances one mutation per round; a round is
//   mlt_start (StartIteration + camera samples) -> [trace_closest -> bin ->
//   shade_mlt -> trace_shadow] x max_depth -> mlt_accept
// reusing the path-tracing traversal kernels.  A chain's primary samples live
// in HBM, index-major (sample i of chain c at xs[i * cap + c]) so that a wave
// of consecutive chains touches consecutive 16-byte records.
#include "kernels.hpp"
#include "kernels_mlt.hpp"
#include "shading.hpp"
#include "wave.hpp"

namespace bnd {

__device__ __forceinline__ void wave_push_m(bool pred, uint32_t* ctr, uint32_t* arr, uint32_t val) {
    const uint64_t m = __ballot(pred);
    if (m == 0) return;
    const int leader = __ffsll((unsigned long long)m) - 1;
    uint32_t base = 0;
    if ((int)lane_id() == leader) base = atomicAdd(ctr, (uint32_t)__popcll(m));
    base = __shfl(base, leader);
    if (pred) arr[base + mbcnt(m)] = val;
}

// ErfInv of PSSMLT.fs:69-94 (Giles 2010 single-precision polynomial).
__device__ __forceinline__ float erf_inv(float x) {
    x = min_n(max_n(x, -0.99999f), 0.99999f);
    float w = -log_f(fmaf(x, -x, 1.0f));
    float p;
    if (w < 5.0f) {
        w = w - 2.5f;
        p = 2.81022636e-08f;
        p = fmaf(p, w, 3.43273939e-07f);
        p = fmaf(p, w, -3.5233877e-06f);
        p = fmaf(p, w, -4.39150654e-06f);
        p = fmaf(p, w, 0.00021858087f);
        p = fmaf(p, w, -0.00125372503f);
        p = fmaf(p, w, -0.00417768164f);
        p = fmaf(p, w, 0.246640727f);
        return fmaf(p, w, 1.50140941f) * x;
    }
    w = sqrtf(w) - 3.0f;
    p = -0.000200214257f;
    p = fmaf(p, w, 0.000100950558f);
    p = fmaf(p, w, 0.00134934322f);
    p = fmaf(p, w, -0.00367342844f);
    p = fmaf(p, w, 0.00573950773f);
    p = fmaf(p, w, -0.0076224613f);
    p = fmaf(p, w, 0.00943887047f);
    p = fmaf(p, w, 1.00167406f);
    return fmaf(p, w, 2.83297682f) * x;
}

// MLTSampler of one chain (PSSMLT.fs:35-150), loaded into registers for one
// kernel and written back at the end.  FRESH: the first iteration of a new
// sampler (all samples are large-step draws in index order), which needs no
// sample storage.
template <bool FRESH>
struct ChainSampler {
    const MltBuf* M;
    uint32_t c;             // chain slot
    uint32_t inner;
    int cur, last_large, init;
    bool large;
    __device__ __forceinline__ void load(const MltBuf& m, uint32_t slot) {
        M = &m;
        c = slot;
        inner = m.inner[slot];
        if (!FRESH) {
            const int4 st = m.state[slot];
            cur = st.x; last_large = st.y; init = st.z; large = st.w != 0;
        }
    }
    __device__ __forceinline__ void store() const {
        M->inner[c] = inner;
        if (!FRESH) M->state[c] = make_int4(cur, last_large, init, large ? 1 : 0);
    }
    // StartIteration (PSSMLT.fs:58-61)
    __device__ __forceinline__ void start_iteration(float large_step_prob) {
        large = cur == 0 || lcg(inner) < large_step_prob;
        cur = cur + 1;
    }
    // The records of samples [i0, i0 + N), read together so that their loads
    // overlap: the initialised ones (below `init`; EnsureReady visits them in
    // index order, so `init` only grows past indices not read here).
    template <int N>
    __device__ __forceinline__ void fetch(int i0, float4 (&r)[N]) const {
#pragma unroll
        for (int j = 0; j < N; ++j)
            r[j] = !FRESH && i0 + j < init ? M->xs[(size_t)(i0 + j) * M->cap + c] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    }
    // EnsureReady (PSSMLT.fs:68-136) of sample i, whose record `r` was fetched.
    __device__ __forceinline__ float ensure(int i, float4 r) {
        if (FRESH) return lcg(inner);
        float4* rec = M->xs + (size_t)i * M->cap + c;
        float v, vb;
        int lm, mb;
        if (init <= i) {
            v = 0.0f; vb = 0.0f; lm = 0; mb = 0;
            init = i + 1;
        } else {
            v = r.x; vb = r.y; lm = __float_as_int(r.z); mb = __float_as_int(r.w);
        }
        if (lm < last_large) { v = lcg(inner); lm = last_large; }
        vb = v; mb = lm;
        if (large) {
            v = lcg(inner);
        } else {
            float val;
            if (M->strategy == 0) {
                const float ns = sqrtf(2.0f) * erf_inv(fmaf(2.0f, lcg(inner), -1.0f));
                const float es = M->sigma * sqrtf((float)(cur - lm));
                val = fmaf(ns, es, v);
            } else {
                val = v;
                const float a = log_f(M->eps_max / M->eps_min);
                for (int k = lm; k <= cur - 1; ++k) {
                    const float u1 = lcg(inner) - 0.5f;
                    const float u2 = u1 < 0.0f ? 1.0f + 2.0f * u1 : 2.0f * u1;
                    val = val + copysignf(M->eps_max * exp_f(-a * u2), u1);
                }
            }
            v = val - floorf(val);
        }
        lm = cur;
        *rec = make_float4(v, vb, __int_as_float(lm), __int_as_float(mb));
        return v;
    }
};

// Chain set-up for chains [c0, c0+n): the outer sampler picks a bootstrap
// path through the alias table (PSSMLT.fs:286-290) whose seed starts a fresh
// MLTSampler.
__global__ void k_mlt_chain_init(MltBuf M, MltParamsDev p, const float* aprob, const int* aidx, int c0, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t chain = (uint32_t)(c0 + i);
    uint32_t outer = xxhash2((uint32_t)p.frame_id, chain);
    // AliasTable.Sample (AliasTable.fs:52-62), index clamped to the table.
    const int nb = p.n_bootstrap;
    float u = lcg(outer) * (float)nb;
    int k = sat_int(u);
    k = k < nb - 1 ? k : nb - 1;
    k = k < 0 ? 0 : k;
    u = u - (float)k;
    const int bid = u < aprob[k] ? k : aidx[k];
    M.outer[i] = outer;
    M.bid[i] = bid;
    M.inner[i] = xxhash2((uint32_t)p.frame_id, (uint32_t)bid);
    M.state[i] = make_int4(0, 0, 0, 0);
    M.rad[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    M.accepted[i] = 0;
}

// StartIteration and the camera ray of one MLT path (PSSMLT.fs:254-266,
// 292-304, 333-345).  Bootstrap (FRESH) paths seed their own sampler.
template <bool FRESH>
__global__ __launch_bounds__(256) void k_mlt_start(DevScene S, PathBuf P, MltBuf M, MltParamsDev p, Ctl* ctl,
                                                   uint32_t* queue, int id0, int n) {
    for (int base = blockIdx.x * blockDim.x; base < n; base += gridDim.x * blockDim.x) {
        const int i = base + threadIdx.x;
        const bool ok = i < n;
        if (ok) {
            ChainSampler<FRESH> smp;
            if (FRESH) {
                smp.M = &M;
                smp.c = (uint32_t)i;
                smp.inner = xxhash2((uint32_t)p.frame_id, (uint32_t)(id0 + i));
            } else {
                smp.load(M, (uint32_t)i);
                smp.start_iteration(p.large_step_prob);
            }
            float4 rs[4];
            smp.fetch(0, rs);
            const float u0 = smp.ensure(0, rs[0]), u1 = smp.ensure(1, rs[1]);
            const float upx = u0 * (float)p.width, upy = u1 * (float)p.height;
            const int ix = min(p.width - 1, sat_int(upx));
            const int iy = min(p.height - 1, sat_int(upy));
            V2 frac = {upx - (float)ix, upy - (float)iy};
            V2 uLens;
            uLens.x = smp.ensure(2, rs[2]);
            uLens.y = smp.ensure(3, rs[3]);
            const Ray r = camera_ray(S, ix, iy, frac, uLens);
            P.o[i] = f4(r.o, 0.0f);
            P.d[i] = f4(r.d, 0.0f);
            P.beta[i] = make_float4(1.0f, 1.0f, 1.0f, 0.0f);
            P.L[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            M.pix_new[i] = iy * p.width + ix;
            smp.store();
        }
        wave_push_m(ok, &ctl->q_count[0], queue, (uint32_t)i);
    }
}

// PSSMLTIntegrator.Li after a closest hit (PSSMLT.fs:184-243): emission with
// MIS, then seven primary samples drawn for every hit, then NEE set-up, BSDF
// sampling and Russian roulette.  The NEE term is applied by trace_shadow.
template <bool FRESH>
__global__ __launch_bounds__(256) void k_shade_mlt(DevScene S, PathBuf P, MltBuf M, Ctl* ctl, const uint32_t* bins,
                                                   uint32_t bin_cap, uint32_t* queue_next, int nxt, uint32_t* shq,
                                                   uint32_t* sh_ctr, int depth, int max_depth, int rr_depth) {
    uint32_t start[kBins + 1];
    start[0] = 0;
    for (int b = 0; b < kBins; ++b) start[b + 1] = start[b] + ctl->bin_count[b];
    const uint32_t total = start[kBins];
    for (uint32_t base = blockIdx.x * blockDim.x; base < total; base += gridDim.x * blockDim.x) {
        const uint32_t i = base + threadIdx.x;
        bool cont = false, shadow = false;
        uint32_t slot = 0;
        if (i < total) {
            int b = 0;
            while (i >= start[b + 1]) ++b;
            slot = bins[(size_t)b * bin_cap + (i - start[b])];
            const float4 o4 = P.o[slot], d4 = P.d[slot];
            const Ray r = {xyz(o4), xyz(d4)};
            const float prev_pdf = d4.w;
            const int4 h = P.hit[slot];
            const Geom g = hit_geometry(S, r, h.x, h.y, __int_as_float(h.z));
            const int4 info = S.insts[h.x].info;
            V3 L = xyz(P.L[slot]);
            V3 beta = xyz(P.beta[slot]);
            if (info.w >= 0) {
                V3 Le;
                float lpdf;
                light_eval(S, h.x, r.o, g, Le, lpdf);
                const float w = depth == 0 ? 1.0f : prev_pdf * rcp_e(lpdf + prev_pdf);
                L = fma3(beta, Le * w, L);
            }
            ChainSampler<FRESH> smp;
            smp.load(M, slot);
            const int s0 = 4 + 7 * depth;
            float4 rs[7];
            smp.fetch(s0, rs);
            const float uLight = smp.ensure(s0, rs[0]);
            V2 uEmit;
            uEmit.x = smp.ensure(s0 + 1, rs[1]);
            uEmit.y = smp.ensure(s0 + 2, rs[2]);
            const float uLobe = smp.ensure(s0 + 3, rs[3]);
            V2 uBSDF;
            uBSDF.x = smp.ensure(s0 + 4, rs[4]);
            uBSDF.y = smp.ensure(s0 + 5, rs[5]);
            const float uRR = smp.ensure(s0 + 6, rs[6]);
            smp.store();
            if (info.z >= 0) {
                const Mat m = load_mat(S, info.z);
                const V3 wo_l = to_local(g, -r.d);
                const LightSample ls = light_sample(S, g.p, uLight, uEmit);
                if (ls.pdf != 0.0f) {
                    const BSDF e = mat_eval(m, wo_l, to_local(g, ls.wi));
                    const float dist = length(ls.p - g.p);
                    P.sh_o[slot] = f4(g.p, dist - 1e-3f);
                    P.sh_d[slot] = f4(ls.wi, 0.0f);
                    P.nee_a[slot] = f4(beta * e.f, 0.0f);
                    P.nee_b[slot] = f4(ls.L * rcp_e(e.pdf + ls.pdf), 0.0f);
                    shadow = true;
                }
                V3 wi_l;
                const BSDF bs = mat_sample(m, wo_l, uLobe, uBSDF, wi_l);
                if (bs.pdf != 0.0f) {
                    const V3 wi = to_world(g, wi_l);
                    beta = beta * bs.f * rcp_e(bs.pdf);
                    cont = true;
                    if (depth >= rr_depth) {
                        const float q = min_n(1.0f, max_n(beta.x, max_n(beta.y, beta.z)));
                        if (uRR < q) beta = beta * rcp_e(q);
                        else cont = false;
                    }
                    P.o[slot] = f4(g.p, 0.0f);
                    P.d[slot] = f4(wi, bs.pdf);
                    P.beta[slot] = f4(beta, 0.0f);
                }
                cont = cont && depth + 1 < max_depth;
            }
            P.L[slot] = f4(L, 0.0f);
        }
        wave_push_m(cont, &ctl->q_count[nxt], queue_next, slot);
        wave_push_m(shadow, sh_ctr, shq, slot);
    }
}

__device__ __forceinline__ float luminance(V3 L) { return dot(L, mk(0.2126f, 0.7152f, 0.0722f)); }

// Bootstrap weight y = luminance(L) (PSSMLT.fs:268-273).
__global__ void k_mlt_boot_done(PathBuf P, float* weights, int id0, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const V3 L = xyz(P.L[i]) * rcp_e(1.0f);
    weights[id0 + i] = luminance(L);
}

// End of the replayed bootstrap path of a chain (PSSMLT.fs:306-312): keep
// L, y and the pixel, Accept(), and re-seed the inner sampler.
__global__ void k_mlt_replay_done(PathBuf P, MltBuf M, MltParamsDev p, int c0, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const V3 L = xyz(P.L[i]) * rcp_e(1.0f);
    M.Lcur[i] = f4(L, luminance(L));
    M.pix_cur[i] = M.pix_new[i];
    int4 st = M.state[i];
    if (st.w) st.y = st.x;                                  // Accept(): LastLargeStepIteration
    M.state[i] = st;
    M.inner[i] = xxhash3((uint32_t)(c0 + i), (uint32_t)M.bid[i], (uint32_t)p.frame_id);
}

__device__ __forceinline__ void film_add(float* film, const MltParamsDev& p, int pix, V3 c) {
    const int x = pix % p.width, y = pix / p.width;
    float* f = film + 3 * ((size_t)(p.height - y - 1) * p.width + x);     // Film.fs:50
    atomicAdd(f, c.x);
    atomicAdd(f + 1, c.y);
    atomicAdd(f + 2, c.z);
}

// Metropolis acceptance of one mutation (PSSMLT.fs:347-373).
// PrimarySample.Restore (PSSMLT.fs:31-33): Value = ValueBackUp,
// LastModification = ModificationBackUp.
__device__ __forceinline__ void restore_sample(float4* rec, float4 r) {
    if (__float_as_uint(r.x) != __float_as_uint(r.y) || __float_as_uint(r.z) != __float_as_uint(r.w))
        *rec = make_float4(r.y, r.y, r.w, r.w);
}

__global__ void k_mlt_accept(PathBuf P, MltBuf M, MltParamsDev p, float* film, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const V3 Ln = xyz(P.L[i]) * rcp_e(1.0f);
    const float yn = luminance(Ln);
    const float4 lc = M.Lcur[i];
    const V3 Lc = xyz(lc);
    const float y = lc.w;
    int4 st = M.state[i];
    const bool large = st.w != 0;
    const float a = min_n(1.0f, yn / y);
    const float w_old = (1.0f - a) / fmaf(y, p.inv_b, p.large_step_prob);
    V3 rad = xyz(M.rad[i]) + w_old * Lc;
    const float w_new = (a + (large ? 1.0f : 0.0f)) / fmaf(yn, p.inv_b, p.large_step_prob);
    uint32_t outer = M.outer[i];
    if (lcg(outer) < a) {
        M.accepted[i] = M.accepted[i] + 1;
        film_add(film, p, M.pix_cur[i], rad * p.inv_eff);
        rad = w_new * Ln;
        M.pix_cur[i] = M.pix_new[i];
        M.Lcur[i] = f4(Ln, yn);
        if (large) st.y = st.x;                              // Accept()
    } else {
        if (a > 0.0f) film_add(film, p, M.pix_new[i], (w_new * p.inv_eff) * Ln);
        // Reject(): restore every initialised sample, step the iteration back.
        // Eight records in flight per lane; a record that already equals its
        // backup is left unwritten.
        float4* rec = M.xs + i;
        const size_t stride = M.cap;
        int k = 0;
        for (; k + 8 <= st.z; k += 8) {
            float4 r[8];
#pragma unroll
            for (int j = 0; j < 8; ++j) r[j] = rec[(size_t)(k + j) * stride];
#pragma unroll
            for (int j = 0; j < 8; ++j) restore_sample(rec + (size_t)(k + j) * stride, r[j]);
        }
        for (; k < st.z; ++k) restore_sample(rec + (size_t)k * stride, rec[(size_t)k * stride]);
        st.x = st.x - 1;
    }
    M.outer[i] = outer;
    M.rad[i] = f4(rad, 0.0f);
    M.state[i] = st;
}

// Final Film.Accumulate of a chain (PSSMLT.fs:375).
__global__ void k_mlt_finish(MltBuf M, MltParamsDev p, float* film, int* accepted_out, int c0, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    film_add(film, p, M.pix_cur[i], xyz(M.rad[i]) * p.inv_eff);
    accepted_out[c0 + i] = M.accepted[i];
}

// ------------------------------------------------------------------ launch
static int blocks(int n) { return (n + 255) / 256; }
void launch_mlt_chain_init(const MltBuf& M, const MltParamsDev& p, const float* aprob, const int* aidx, int c0, int n,
                           hipStream_t st) {
    hipLaunchKernelGGL(k_mlt_chain_init, dim3(blocks(n)), dim3(256), 0, st, M, p, aprob, aidx, c0, n);
}
void launch_mlt_start(bool fresh, const DevScene& S, const PathBuf& P, const MltBuf& M, const MltParamsDev& p,
                      Ctl* ctl, uint32_t* queue, int id0, int n, int grid, hipStream_t st) {
    const int g = min(grid, blocks(n));
    if (fresh) hipLaunchKernelGGL(k_mlt_start<true>, dim3(g), dim3(256), 0, st, S, P, M, p, ctl, queue, id0, n);
    else hipLaunchKernelGGL(k_mlt_start<false>, dim3(g), dim3(256), 0, st, S, P, M, p, ctl, queue, id0, n);
}
void launch_shade_mlt(bool fresh, const DevScene& S, const PathBuf& P, const MltBuf& M, Ctl* ctl,
                      const uint32_t* bins, uint32_t bin_cap, uint32_t* queue_next, int nxt, uint32_t* shq,
                      int depth, int max_depth, int rr_depth, int grid, hipStream_t st, int par) {
    uint32_t* sh_ctr = par < 0 ? &ctl->sh_count : &ctl->shc[par];
    if (fresh)
        hipLaunchKernelGGL(k_shade_mlt<true>, dim3(grid), dim3(256), 0, st, S, P, M, ctl, bins, bin_cap, queue_next,
                           nxt, shq, sh_ctr, depth, max_depth, rr_depth);
    else
        hipLaunchKernelGGL(k_shade_mlt<false>, dim3(grid), dim3(256), 0, st, S, P, M, ctl, bins, bin_cap, queue_next,
                           nxt, shq, sh_ctr, depth, max_depth, rr_depth);
}
void launch_mlt_boot_done(const PathBuf& P, float* weights, int id0, int n, hipStream_t st) {
    hipLaunchKernelGGL(k_mlt_boot_done, dim3(blocks(n)), dim3(256), 0, st, P, weights, id0, n);
}
void launch_mlt_replay_done(const PathBuf& P, const MltBuf& M, const MltParamsDev& p, int c0, int n, hipStream_t st) {
    hipLaunchKernelGGL(k_mlt_replay_done, dim3(blocks(n)), dim3(256), 0, st, P, M, p, c0, n);
}
void launch_mlt_accept(const PathBuf& P, const MltBuf& M, const MltParamsDev& p, float* film, int n, hipStream_t st) {
    hipLaunchKernelGGL(k_mlt_accept, dim3(blocks(n)), dim3(256), 0, st, P, M, p, film, n);
}
void launch_mlt_finish(const MltBuf& M, const MltParamsDev& p, float* film, int* accepted, int c0, int n,
                       hipStream_t st) {
    hipLaunchKernelGGL(k_mlt_finish, dim3(blocks(n)), dim3(256), 0, st, M, p, film, accepted, c0, n);
}

}  // namespace bnd
