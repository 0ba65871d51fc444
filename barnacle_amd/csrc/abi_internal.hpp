// abi_internal.hpp -- context and scene records behind the C ABI.
#pragma once
#include <hip/hip_runtime.h>

#include <map>
#include <memory>
#include <string>
#include <utility>
#include <vector>

#include "../../include/barnacle_hip.h"
#include "host_scene.hpp"
#include "kernels.hpp"
#include "kernels_mlt.hpp"

namespace bnx {

void hip_check(hipError_t e, const char* what);

// Device allocation that only grows.
struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf();
    void ensure(size_t n);
    void upload(const void* src, size_t n, hipStream_t s);
};

// A bottom-level acceleration structure as MeshPrimitive holds it
// (Mesh.fs:157-186): nodes, triangles gathered in BVH order, area alias table.
struct BlasRec {
    std::vector<bn_bvh_node> nodes;
    std::vector<float> tris;     // 9 floats per triangle
    bnh::AliasTable alias;
    bool live = true;            // false once released (its host copy is freed)
    int owner = -1;              // scene whose bn_scene_build made it (released with
                                 // that scene); -1: bn_upload_blas, released by
                                 // bn_release_blas or bn_destroy
};

struct SceneRec {
    bool live = false;
    DevBuf tlas, blas, tris, insts, mats, lightdefs, lrec, aprob, apdf, aidx;
    int n_lights = 0, n_insts = 0;
    bn_camera camera{};
    float c2w[16] = {};
    bnd::DevScene dev{};
    uint64_t gen = 0;                        // bumped by bn_set_params
    std::vector<bn_material> mats_host;     // what bn_set_params edits
    std::vector<bn_light> lights_host;
    std::shared_ptr<bnh::HostScene> host;   // set by bn_scene_build
};

// Event pairs around kernel launches, resolved after the stream syncs.
struct KernelTimer {
    bool on = false;
    std::vector<hipEvent_t> pool;
    size_t used = 0;
    std::vector<std::pair<std::string, size_t>> pending;   // name, first event index
    std::vector<std::pair<std::string, std::pair<int64_t, double>>> totals;
    hipEvent_t next();
    void begin(const char* name, hipStream_t s);
    void end(hipStream_t s);
    void resolve();                                        // after a stream sync
    ~KernelTimer();
};

struct GraphExec {
    hipGraphExec_t e = nullptr;
    explicit GraphExec(hipGraphExec_t x) : e(x) {}
    GraphExec(GraphExec&& o) noexcept : e(o.e) { o.e = nullptr; }
    GraphExec(const GraphExec&) = delete;
    ~GraphExec() { if (e) (void)hipGraphExecDestroy(e); }
};

struct Ctx {
    int device = 0;
    int n_cu = 1;
    hipStream_t stream = nullptr;
    hipStream_t stream2 = nullptr;        // shadow rays of compact rounds, beside `stream`
    hipEvent_t ev_shade = nullptr, ev_shadow = nullptr;
    bool octant_order = true;             // compact shading groups each block's outputs by ray octant
    int skip_dark = 1;                    // no shadow ray for an exactly-zero NEE term: 0, 1 plain frames, 2 all
    bool overlap_shadow = true;           // trace shadow rays on stream2 in plain path-tracing frames
    int trace_blocks_per_cu = 7;
    int refill = 40;                      // idle lanes that trigger a refill
    size_t chunk_slots = (size_t)64 << 20;       // path slots in flight
    size_t window_paths = 0;                     // paths per round; 0: as many as slots
    std::vector<BlasRec> blas;
    std::vector<std::unique_ptr<SceneRec>> scenes;
    // path state
    size_t path_cap = 0;
    DevBuf p_o, p_d, p_beta, p_L, p_rng, p_hit, p_sh_o, p_sh_d, p_nee_a, p_nee_b, p_item, p_pid;
    DevBuf p2_o, p2_d, p2_beta, p2_L, p2_rng, p2_item, p2_pid;   // second state copy (compact rounds)
    DevBuf queue[2], freel[2], shq, bins, spill, film, ctl, stats, L_out, fetch_seg;
    // path-replay adjoint: per-slot terms, gradients, film adjoint
    DevBuf a_delta, a_pend_a, a_pend_r, a_nee_da, a_nee_dr, a_gmat, a_glight, a_dfilm;
    size_t adj_cap = 0;
    // L_out holds the radiance of every path of this single-round frame
    // (what bn_backward_pt replays against), when valid.
    struct { bool valid; const SceneRec* scene; uint64_t gen; bn_pt_params p; size_t slots; } fwd{};
    size_t spill_stride = 0;
    DevBuf out_a, out_b, out_c;
    KernelTimer timer;
    // batched rounds
    bool use_graph = true;
    int stack_k = 16;                      // LDS traversal stack window
    int batch_pack = 0;                    // 0 item-major, 1 fair share
    int material_sort = -1;                // bin shading by material kind; -1: adjoint rounds only
    DevBuf b_scenes, b_items, b_hdr;
    std::map<std::string, GraphExec> graphs;   // captured bounce loops by configuration
    // PSSMLT chain state
    size_t mlt_cap = 0;
    int mlt_nx = 0;
    size_t mlt_chunk = (size_t)1 << 20;   // chains (or bootstrap paths) in flight
    DevBuf m_inner, m_outer, m_state, m_bid, m_pix_new, m_pix_cur, m_Lcur, m_rad, m_acc, m_xs;
    DevBuf m_weights, m_aprob, m_aidx, m_accepted_out;
    // GPU BVH builder (bvh_gpu.hip): items (two buffers, then the output
    // nodes), node regions, permutation, task lists, chunk partials
    bool bvh_gpu = true;                   // bn_scene_build builds finite meshes' BLAS on the GPU
    int bvh_small = 512;                   // largest task one wave builds whole (at most 512)
    DevBuf g_items, g_region, g_perm, g_rsize, g_big, g_info, g_small, g_chunks, g_part, g_unodes, g_misc;
    void* g_pinned = nullptr;              // page-locked staging (hipHostMalloc), grown on demand
    size_t g_pinned_bytes = 0;
    std::vector<float> last_bootstrap;     // of the last bn_render_pssmlt
    std::vector<int32_t> last_chain_accepts;
    ~Ctx() {
        if (g_pinned) (void)hipHostFree(g_pinned);
        if (ev_shade) (void)hipEventDestroy(ev_shade);
        if (ev_shadow) (void)hipEventDestroy(ev_shadow);
        if (stream2) (void)hipStreamDestroy(stream2);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

// Viewport constants of the scene's camera for a film size (abi.hip).
void camera_constants(SceneRec& sc, int W, int H);

// Allocation-free first hit (noalloc.hip): a static device array holds its
// scratch; noalloc_init records the array's address on a device (called by
// bn_create with that device current).
constexpr size_t kNoAllocMaxPix = (size_t)1 << 20;
void noalloc_init(int device);
void first_hit_noalloc(Ctx* c, SceneRec& s, const bn_pt_params& p, int sample_id, int32_t* inst, int32_t* prim,
                       float* t);

}  // namespace bnx
