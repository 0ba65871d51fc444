// kernels.hpp -- shared device-buffer layouts and kernel launchers.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "device.hpp"

namespace bnd {

constexpr int kBins = 5;          // material kinds 0..3 + light-only instances
constexpr int kStackK = 32;       // LDS window of the per-lane traversal stack
constexpr int kStackMax = 144;    // TLAS + BLAS depth <= 2 * (64 + 1) entries
constexpr int kBlock = 256;

// Path state, struct of arrays indexed by slot.
struct PathBuf {
    float4* o;        // ray origin
    float4* d;        // ray direction, w = pdf of the last BSDF sample
    float4* beta;
    float4* L;
    uint32_t* rng;    // Sampler state (Base/Sampler.fs)
    int4* hit;        // instance, primitive, t bits
    float4* sh_o;     // shadow ray origin, w = t max
    float4* sh_d;
    float4* nee_a;    // beta * f
    float4* nee_b;    // Le / (pdf_b + pdf_l)
};

// Device-side counters of one bounce loop.
struct Ctl {
    uint32_t q_count[2];
    uint32_t bin_count[kBins];
    uint32_t sh_count;
    uint32_t fetch[4];
    uint32_t pad[4];
};

struct ChunkInfo {
    int width, height;
    int tiles_x, tiles_y;
    int p_pad;        // tiles_x * tiles_y * 64
    int s0, ns;       // samples [s0, s0 + ns) of this chunk
    int spp, frame_id;
};

void launch_raygen(const DevScene& S, const PathBuf& P, Ctl* ctl, uint32_t* queue, const ChunkInfo& c,
                   int grid, hipStream_t st);
void launch_reset(Ctl* ctl, int nxt, hipStream_t st);
void launch_trace_closest(bool count, const DevScene& S, const PathBuf& P, Ctl* ctl, int cur, const uint32_t* queue,
                          uint32_t* bins, uint32_t bin_cap, uint32_t* spill, size_t spill_stride, uint32_t refill,
                          unsigned long long* stats, int grid, hipStream_t st);
void launch_bin(const PathBuf& P, Ctl* ctl, int cur, const uint32_t* queue, uint32_t* bins, uint32_t bin_cap,
                int grid, hipStream_t st);
void launch_shade(const DevScene& S, const PathBuf& P, Ctl* ctl, const uint32_t* bins, uint32_t bin_cap,
                  uint32_t* queue_next, int nxt, uint32_t* shq, const int4* lrec, int depth, int max_depth,
                  int rr_depth, int grid, hipStream_t st);
void launch_trace_shadow(bool count, const DevScene& S, const PathBuf& P, Ctl* ctl, const uint32_t* shq,
                         uint32_t* spill, size_t spill_stride, uint32_t refill, unsigned long long* stats, int grid,
                         hipStream_t st);
void launch_film(const PathBuf& P, float* film, const ChunkInfo& c, float inv_spp, hipStream_t st);
void launch_shade_normal(const DevScene& S, const PathBuf& P, Ctl* ctl, const uint32_t* bins, uint32_t bin_cap,
                         int grid, hipStream_t st);
void launch_first_hit(const DevScene& S, int w, int h, int sample_id, int spp, int frame_id, int* inst, int* prim,
                      float* t, uint32_t* spill, size_t spill_stride, uint32_t* fetch, int grid, hipStream_t st);

}  // namespace bnd
