// kernels_mlt.hpp -- device buffers and launchers of the PSSMLT wavefront.
#pragma once
#include "kernels.hpp"

namespace bnd {

// Per-chain state, struct of arrays over chain slots [0, cap).
struct MltBuf {
    uint32_t* inner;      // MLTSampler.InnerSampler state
    uint32_t* outer;      // the chain's own Sampler (acceptance draws)
    int4* state;          // CurrentIteration, LastLargeStepIteration, InitializedSampleCount, LargeStep
    int* bid;             // bootstrap path index
    int* pix_new;         // pixel (y*W + x) of the proposal
    int* pix_cur;         // pixel of the current state
    float4* Lcur;         // current L, w = luminance y
    float4* rad;          // accumulated radiance of the current state
    int* accepted;
    float4* xs;           // PrimarySample (Value, ValueBackUp, LastModification, ModificationBackUp)
    uint32_t cap;
    int strategy;
    float sigma, eps_min, eps_max;
};

struct MltParamsDev {
    int width, height, frame_id, n_bootstrap;
    float large_step_prob;
    float inv_b;          // ReciprocalEstimate(B)
    float inv_eff;        // ReciprocalEstimate(mutations * chains / pixels)
};

void launch_mlt_chain_init(const MltBuf& M, const MltParamsDev& p, const float* aprob, const int* aidx, int c0, int n,
                           hipStream_t st);
void launch_mlt_start(bool fresh, const DevScene& S, const PathBuf& P, const MltBuf& M, const MltParamsDev& p,
                      Ctl* ctl, uint32_t* queue, int id0, int n, int grid, hipStream_t st);
void launch_shade_mlt(bool fresh, const DevScene& S, const PathBuf& P, const MltBuf& M, Ctl* ctl,
                      const uint32_t* bins, uint32_t bin_cap, uint32_t* queue_next, int nxt, uint32_t* shq,
                      int depth, int max_depth, int rr_depth, int grid, hipStream_t st, int par = -1);
void launch_mlt_boot_done(const PathBuf& P, float* weights, int id0, int n, hipStream_t st);
void launch_mlt_replay_done(const PathBuf& P, const MltBuf& M, const MltParamsDev& p, int c0, int n, hipStream_t st);
void launch_mlt_accept(const PathBuf& P, const MltBuf& M, const MltParamsDev& p, float* film, int n, hipStream_t st);
void launch_mlt_finish(const MltBuf& M, const MltParamsDev& p, float* film, int* accepted, int c0, int n,
                       hipStream_t st);

}  // namespace bnd
