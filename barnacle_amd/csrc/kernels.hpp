// kernels.hpp -- shared device-buffer layouts and kernel launchers.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "device.hpp"

namespace bnd {

constexpr int kBins = 5;          // material kinds 0..3 + light-only instances
constexpr int kStackK = 32;       // default LDS window of the per-lane traversal stack
constexpr int kStackKMin = 16;    // smallest window (sizes the spill area)
constexpr int kStackMax = 144;    // TLAS + BLAS depth <= 2 * (64 + 1) entries
constexpr int kBlock = 256;
constexpr int kFetchSegs = 128;   // fetch counters per trace queue (closest-hit traversal)
constexpr int kShadowSegs = 64;   // of which the any-hit traversals use the first 64
constexpr int kSegStride = 64;    // words between two fetch counters (own cache lines)

constexpr uint32_t kNoPath = 0xffffffffu;

// Waves per SIMD the 16-entry-window trace kernels are built for: seven for
// one scene (72 VGPRs), six in batches, whose per-lane scene pointers need
// the 80 VGPRs of six waves to avoid spilling.
constexpr int kTraceWaves = 7;
constexpr int kTraceWavesBatch = 6;

// Path state, struct of arrays indexed by slot.
struct PathBuf {
    float4* o;        // ray origin, w = bounce depth of the path (int bits)
    float4* d;        // ray direction, w = pdf of the last BSDF sample
    float4* beta;
    float4* L;
    uint32_t* rng;    // Sampler state (Base/Sampler.fs)
    int4* hit;        // instance, primitive, t bits, material bin
    float4* sh_o;     // shadow ray origin, w = t max
    float4* sh_d;
    float4* nee_a;    // beta * f
    float4* nee_b;    // Le / (pdf_b + pdf_l)
    int* item;        // render item of the slot's path (batched rounds)
    uint32_t* pid;    // path id of the slot within its round, kNoPath if idle
};

// The second copy of the per-path state of compact rounds.  There the active
// paths of an iteration are dense in positions [0, count) of one copy, and
// shading writes the paths that continue, compacted, into the other copy
// (hit, shadow-ray and NEE arrays are shared and indexed by position).
struct PathState {
    float4 *o, *d, *beta, *L;
    uint32_t* rng;
    int* item;
    uint32_t* pid;
};

// Device-side counters of the iterations of a round.
struct Ctl {
    uint32_t q_count[2];      // active queues (this iteration, next iteration)
    uint32_t bin_count[kBins];
    uint32_t sh_count;
    uint32_t fetch[4];
    uint32_t free_count[2];   // free-slot lists (read by this iteration's regen, filled for the next)
    uint32_t next_pid;        // next path id of the round to issue (may run past the end)
    uint32_t pad;
    uint32_t shc[2];          // compact rounds: shadow rays of the iterations of each parity
};

// One render item of a round: samples [s0, s0 + ns) of every pixel of one
// film, path ids [path_base, path_base + ns * p_pad) of the round.
struct RoundItem {
    int path_base, path_count;
    int scene;                // index into the scene table (batch)
    int width, height, tiles_x, p_pad;
    int s0, ns, spp, frame_id;
    int max_depth, rr_depth;
    float inv_spp;            // ReciprocalEstimate(spp)
    int film_offset;          // float offset of the item's film
    int pix_end;              // cumulative pixel count up to and including this item
};

// Sizes of a batched round, read by the kernels from device memory so that a
// captured block of iterations can be replayed for every round.
struct RoundHeader {
    int n_items;
    uint32_t n_paths;
    uint32_t n_pix;
    int pad;
};

// Everything a round needs besides buffers.
struct Launch {
    DevScene scene;           // plain mode
    RoundItem item;           // plain mode
    const RoundHeader* hdr;   // batch mode (device)
    const RoundItem* items;   // batch mode (device)
    const DevScene* scenes;   // batch mode (device)
    uint32_t n_pix;           // plain mode
    uint32_t n_slots;         // path slots in flight
    int grid_trace, grid_shade;
    // Batch mode: every item of the round traces the geometry of `scene`, so
    // its compact rounds launch the single-scene trace kernels (the scene
    // bound once per launch) on grid_uniform blocks.
    bool uniform;
    int grid_uniform;
    uint32_t refill;
    hipStream_t stream;
    // Compact path-tracing rounds with stream2 set trace the shadow rays of
    // iteration i on stream2 while stream runs regen, prep and trace_closest
    // of iteration i + 1; ev_shade / ev_shadow order the two streams.  Not
    // for rounds captured into a graph.
    hipStream_t stream2;
    hipEvent_t ev_shade, ev_shadow;
    // Compact shading groups each block's continuing paths and shadow rays
    // by the octant of their ray direction.
    bool octant_order;
    // Compact shading casts no shadow ray whose NEE term is exactly zero:
    // 0 never, 1 in plain frames (counting frames cast the reference's
    // rays), 2 in every frame (counting frames count the rays traced).
    int skip_dark;
};

// Path-replay adjoint state (section "adjoint" of kernels_pt.hip).
struct AdjBuf {
    float4* delta;            // dLoss/dLi of the slot's path: film adjoint / spp
    float4* pend_a;           // BSDF-sample term: (d f / d albedo) / f per channel, w = material bits (-1: none)
    float4* pend_r;           // (d f / d roughness) / f per channel
    float4* nee_da;           // beta * d f_nee / d albedo per channel, w = material bits
    float4* nee_dr;           // beta * d f_nee / d roughness per channel, w = light definition bits
    const float4* L_fwd;      // radiance of every path id of the round from the forward pass
    const float* d_film;      // W*H*3 film adjoint, Film.Pixels order
    float* g_mat;             // 4 per material: albedo rgb, roughness
    float* g_light;           // 4 per light definition: emission rgb, unused
    int n_mat, n_light;       // table sizes
    int lds_bytes;            // > 0: NEE gradients summed per block in this much LDS first
};

struct Buffers {
    uint32_t* queue[2];
    uint32_t* freel[2];
    uint32_t* shq;
    uint32_t* bins;
    uint32_t bin_cap;
    uint32_t* spill;
    uint32_t* spill2;         // a second spill area, for a shadow traversal running beside trace_closest
    size_t spill_stride;
    int stack_k;              // LDS stack window: 16 or 32 entries
    unsigned long long* stats;
    float* film;
    float4* L_out;            // radiance per path id of the round
    uint32_t* fetch_seg;      // 2 x kFetchSegs fetch counters: closest, then the shadow
                              // traversal (compact rounds: two sets, one per parity)
    int sort;                 // bin the shade queue by material kind (1), not (0), adjoint only (-1)
    AdjBuf adj;               // used by Integrator::Adjoint only
    PathState alt;            // second state copy: Integrator::Path rounds run compact when it is set
};

// Called around every launch of a round (timing, counters).
struct LaunchHook {
    virtual void begin(const char* name, hipStream_t st) = 0;
    virtual void end(hipStream_t st) = 0;
    virtual void after_bounce(int cur) = 0;
    virtual ~LaunchHook() = default;
};

// A round streams its paths through the slots: every iteration refills the
// free slots with new paths (regen), then traces, bins, shades and traces
// shadow rays for all active paths, whatever their depth.
//   begin:      counter clear, every slot free
//   iterations: n (even) x [regen, prep, trace_closest, bin, shade, trace_shadow]
//   end:        retire the last paths, fold the round into the films (not for
//               adjoint rounds, which only accumulate gradients)
// Path-tracing rounds run compact when a second state copy is given (and
// material binning is not forced): no free lists, queues or bins; regen
// appends new paths after the continuing ones, shading moves the continuing
// paths to the other copy and retires the ended ones by path id.
//   iterations: n (even) x [regen, prep, trace_closest, shade, trace_shadow]
enum class Integrator { Path, Normal, Adjoint, Direct };
void enqueue_round_begin(bool batch, const Launch& L, const PathBuf& P, Ctl* ctl, const Buffers& B,
                         LaunchHook* hook);
void enqueue_iterations(bool batch, Integrator kind, const Launch& L, const PathBuf& P, Ctl* ctl, const Buffers& B,
                        int n, bool count, LaunchHook* hook);
void enqueue_round_end(bool batch, Integrator kind, const Launch& L, const PathBuf& P, Ctl* ctl, const Buffers& B,
                       LaunchHook* hook);

void launch_bin(const PathBuf& P, Ctl* ctl, int cur, const uint32_t* queue, uint32_t* bins, uint32_t bin_cap,
                uint32_t* freel, int sort, int grid, hipStream_t st);
// Counters of one slot-round bounce.  par < 0: every fetch counter and
// sh_count; par 0/1: the closest-hit counters, the any-hit counters of that
// parity and shc[par], leaving those of the other parity to a shadow pass
// still running on a second stream.
void launch_reset(Ctl* ctl, int nxt, uint32_t* fetch_seg, hipStream_t st, int par = -1);
void launch_ctl_clear(Ctl* ctl, hipStream_t st);
void launch_trace_closest(bool count, const DevScene& S, const PathBuf& P, Ctl* ctl, int cur, const uint32_t* queue,
                          const Buffers& B, uint32_t refill, int grid, hipStream_t st);
// par < 0: the rays counted in sh_count; par 0/1: in shc[par], with that
// parity's fetch counters and, when given, its own stack spill area.
void launch_trace_shadow(bool count, const DevScene& S, const PathBuf& P, Ctl* ctl, const uint32_t* shq,
                         const Buffers& B, uint32_t refill, int grid, hipStream_t st, int par = -1,
                         uint32_t* spill = nullptr);
void launch_first_hit(const DevScene& S, int w, int h, int sample_id, int spp, int frame_id, int* inst, int* prim,
                      float* t, uint32_t* spill, size_t spill_stride, uint32_t* fetch, int grid, hipStream_t st);

}  // namespace bnd
