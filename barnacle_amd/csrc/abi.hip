// abi.hip -- the extern "C" entry points of include/barnacle_hip.h.
//
// Host side of the drop-in boundary: owns device memory and the stream,
// uploads scenes in the SoA layout the kernels read, and drives the bounce
// loop.  Errors become status codes with a thread-local message; no C++
// exception crosses the ABI.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <functional>
#include <memory>
#include <string>
#include <vector>

#include "../../include/barnacle_hip.h"
#include "abi_internal.hpp"
#include "bvh_gpu.hpp"
#include "host_scene.hpp"
#include "kernels.hpp"
#include "kernels_mlt.hpp"

using bnh::Error;

namespace {
thread_local std::string g_error;

int guard(const std::function<void()>& f) {
    try {
        f();
        return BN_OK;
    } catch (const Error& e) {
        g_error = e.what();
        return e.code;
    } catch (const std::exception& e) {
        g_error = e.what();
        return BN_E_INVALID;
    }
}
}  // namespace

namespace bnx {

void hip_check(hipError_t e, const char* what) {
    if (e != hipSuccess) throw Error(BN_E_DEVICE, std::string(what) + ": " + hipGetErrorString(e));
}

DevBuf::~DevBuf() { if (p) (void)hipFree(p); }
void DevBuf::ensure(size_t n) {
    if (n <= bytes) return;
    if (p) hip_check(hipFree(p), "hipFree");
    p = nullptr;
    bytes = 0;
    hip_check(hipMalloc(&p, n), "hipMalloc");
    bytes = n;
}
void DevBuf::upload(const void* src, size_t n, hipStream_t s) {
    ensure(std::max<size_t>(n, 16));
    if (n) hip_check(hipMemcpyAsync(p, src, n, hipMemcpyHostToDevice, s), "hipMemcpyAsync H2D");
}

hipEvent_t KernelTimer::next() {
    if (used == pool.size()) {
        hipEvent_t e;
        hip_check(hipEventCreate(&e), "hipEventCreate");
        pool.push_back(e);
    }
    return pool[used++];
}
void KernelTimer::begin(const char* name, hipStream_t s) {
    if (!on) return;
    pending.push_back({name, used});
    hip_check(hipEventRecord(next(), s), "hipEventRecord");
}
void KernelTimer::end(hipStream_t s) {
    if (!on) return;
    hip_check(hipEventRecord(next(), s), "hipEventRecord");
}
void KernelTimer::resolve() {
    for (const auto& pr : pending) {
        float ms = 0.0f;
        hip_check(hipEventElapsedTime(&ms, pool[pr.second], pool[pr.second + 1]), "hipEventElapsedTime");
        bool found = false;
        for (auto& t : totals)
            if (t.first == pr.first) { t.second.first += 1; t.second.second += ms; found = true; break; }
        if (!found) totals.push_back({pr.first, {1, (double)ms}});
    }
    pending.clear();
    used = 0;
}
KernelTimer::~KernelTimer() {
    for (hipEvent_t e : pool) (void)hipEventDestroy(e);
}

// Brackets one kernel launch with timing events when timing is on.
struct Timed {
    KernelTimer& t;
    hipStream_t s;
    Timed(KernelTimer& timer, const char* name, hipStream_t st) : t(timer), s(st) { t.begin(name, s); }
    ~Timed() { t.end(s); }
};

// True when the row-major 4x4 map has 1.0f on the diagonal and +0.0f (the bit
// pattern 0) everywhere else.  Vector3.Transform by such a map returns
// x*1 + y*0 + z*0 + 0 = x + 0.0f per coordinate: the input with -0 made +0.
static bool is_identity_bits(const float* m) {
    for (int r = 0; r < 4; ++r)
        for (int k = 0; k < 4; ++k) {
            uint32_t u;
            std::memcpy(&u, &m[4 * r + k], 4);
            if (u != (r == k ? 0x3f800000u : 0u)) return false;
        }
    return true;
}

static bnd::Xf to_xf(const float* m) {
    bnd::Xf x;
    x.c0 = make_float4(m[0], m[4], m[8], m[12]);
    x.c1 = make_float4(m[1], m[5], m[9], m[13]);
    x.c2 = make_float4(m[2], m[6], m[10], m[14]);
    return x;
}

// A tree the kernels may traverse: interior children at i+1 and at a larger
// index, leaves inside [0, n_items), depth <= 64 (BVHBuildConfig.MaxDepth),
// which bounds the traversal stack (kStackMax) and rules out cycles.
void validate_bvh(const std::vector<bn_bvh_node>& nodes, int n_items, const char* what) {
    const int n = (int)nodes.size();
    if (n == 0) throw Error(BN_E_INVALID, std::string(what) + ": empty BVH");
    std::vector<std::pair<int, int>> work{{0, 0}};
    int64_t seen = 0;
    while (!work.empty()) {
        auto [i, depth] = work.back();
        work.pop_back();
        if (i < 0 || i >= n || depth > 64 || ++seen > n)
            throw Error(BN_E_INVALID, std::string(what) + ": malformed BVH");
        const bn_bvh_node& b = nodes[i];
        if (b.is_leaf) {
            if (b.right_or_offset < 0 || b.count < 0 || (int64_t)b.right_or_offset + b.count > n_items)
                throw Error(BN_E_INVALID, std::string(what) + ": BVH leaf outside the item range");
        } else {
            if (b.split_axis < 0 || b.split_axis > 2 || b.right_or_offset <= i + 1)
                throw Error(BN_E_INVALID, std::string(what) + ": malformed BVH");
            work.push_back({i + 1, depth + 1});
            work.push_back({b.right_or_offset, depth + 1});
        }
    }
}

// A released BLAS keeps its id (ids are never reused) and frees its host copy.
void release_blas(BlasRec& b) {
    b.live = false;
    std::vector<bn_bvh_node>().swap(b.nodes);
    std::vector<float>().swap(b.tris);
    b.alias = bnh::AliasTable{};
}

// Concatenate the referenced BLAS and upload the whole scene.
// Device material records: (kind bits, albedo) (ior, alpha, metallic,
// roughness).  PBRMaterial: Alpha = Max(roughness^2, 1e-3), Metallic =
// Clamp(m, 0, 1) (PBR.fs:11-12); the raw roughness is kept for d alpha.
std::vector<float4> pack_materials(const std::vector<bn_material>& mats) {
    std::vector<float4> d(2 * std::max<size_t>(1, mats.size()));
    for (size_t i = 0; i < mats.size(); ++i) {
        const bn_material& m = mats[i];
        const bool given_alpha = m.kind == BN_MAT_PBR_ALPHA;
        int kind = given_alpha ? BN_MAT_PBR : m.kind;
        float kf;
        std::memcpy(&kf, &kind, 4);
        float r2 = m.roughness * m.roughness;
        float alpha = (r2 != r2) ? r2 : (r2 > 1e-3f ? r2 : 1e-3f);
        // BN_MAT_PBR_ALPHA: the field holds PBRMaterial.Alpha itself; the
        // roughness kept for d alpha / d roughness is its square root
        const float rough = given_alpha ? std::sqrt(m.roughness) : m.roughness;
        if (given_alpha) alpha = m.roughness;
        float met = m.metallic < 0.0f ? 0.0f : (m.metallic > 1.0f ? 1.0f : m.metallic);
        d[2 * i] = make_float4(kf, m.albedo[0], m.albedo[1], m.albedo[2]);
        d[2 * i + 1] = make_float4(m.ior, alpha, met, rough);
    }
    return d;
}
std::vector<float4> pack_lights(const std::vector<bn_light>& lights) {
    std::vector<float4> d(std::max<size_t>(1, lights.size()));
    for (size_t i = 0; i < lights.size(); ++i)
        d[i] = make_float4(lights[i].emission[0], lights[i].emission[1], lights[i].emission[2],
                           lights[i].two_sided ? 1.0f : 0.0f);
    return d;
}

void upload_scene(Ctx* c, SceneRec& sc, const std::vector<bn_bvh_node>& tlas,
                  const std::vector<bn_instance_upload>& insts, const std::vector<bn_material>& mats,
                  const std::vector<bn_light>& lights, const bn_camera& cam, const float* c2w) {
    for (const auto& in : insts) {
        if (in.blas >= (int)c->blas.size() || (in.blas >= 0 && !c->blas[in.blas].live) ||
            (in.blas < 0 && !(in.sphere_radius == in.sphere_radius)))
            throw Error(BN_E_INVALID, "instance references an unknown BLAS");
        if (in.material >= (int)mats.size() || in.light >= (int)lights.size() || in.material < -1 || in.light < -1)
            throw Error(BN_E_INVALID, "instance references an unknown material/light");
    }
    if (tlas.empty()) throw Error(BN_E_INVALID, "empty TLAS");
    validate_bvh(tlas, (int)insts.size(), "TLAS");
    // BLAS pool: nodes keep their BLAS-local child indices; each instance
    // records where its BLAS starts in the node and triangle pools.
    std::map<int, std::pair<int, int>> base;      // BLAS id -> (first node, first triangle)
    std::vector<bn_bvh_node> nodes;
    std::vector<float> tris, aprob, apdf;
    std::vector<int> aidx;
    for (const auto& in : insts) {
        if (in.blas < 0 || base.count(in.blas)) continue;
        const BlasRec& b = c->blas[in.blas];
        base[in.blas] = {(int)nodes.size(), (int)(tris.size() / 9)};
        nodes.insert(nodes.end(), b.nodes.begin(), b.nodes.end());
        tris.insert(tris.end(), b.tris.begin(), b.tris.end());
        aprob.insert(aprob.end(), b.alias.prob.begin(), b.alias.prob.end());
        apdf.insert(apdf.end(), b.alias.pdf.begin(), b.alias.pdf.end());
        aidx.insert(aidx.end(), b.alias.alias.begin(), b.alias.alias.end());
    }
    std::vector<bnd::DevInst> di(insts.size());
    std::vector<int4> lrec;
    for (size_t i = 0; i < insts.size(); ++i) {
        const bn_instance_upload& in = insts[i];
        bnd::DevInst& d = di[i];
        d.w2o = to_xf(in.w2o);
        d.o2w = to_xf(in.o2w);
        float pdf_const = 0.0f;
        int ntri = 0;
        if (in.blas >= 0) {
            const BlasRec& b = c->blas[in.blas];
            ntri = (int)(b.tris.size() / 9);
            if (in.light >= 0) {
                // MeshInstance.EvalPDF (Mesh.fs:300-304) with the tag reset to 0.
                bnh::M4 m;
                std::memcpy(m.m, in.o2w, 64);
                const float* p = b.tris.data();
                bnh::V3 p0 = bnh::transform_point({p[0], p[1], p[2]}, m);
                bnh::V3 p1 = bnh::transform_point({p[3], p[4], p[5]}, m);
                bnh::V3 p2 = bnh::transform_point({p[6], p[7], p[8]}, m);
                pdf_const = b.alias.pdf[0] / bnh::tri_area(p0, p1, p2);
            }
        }
        d.lo = make_float4(in.bounds[0], in.bounds[1], in.bounds[2], in.sphere_radius);
        d.hi = make_float4(in.bounds[3], in.bounds[4], in.bounds[5], pdf_const);
        d.info = make_int4(in.blas >= 0 ? base[in.blas].first : -1, in.blas >= 0 ? base[in.blas].second : 0,
                           in.material, in.light);
        d.aux = make_int4(is_identity_bits(in.w2o) ? 1 : 0, 0, 0, 0);
        if (in.light >= 0) lrec.push_back(make_int4((int)i, d.info.y, ntri, in.light));
    }
    for (const auto& m : mats)
        if (m.kind < 0 || m.kind > BN_MAT_PBR_ALPHA) throw Error(BN_E_INVALID, "unknown material kind");
    sc.mats_host = mats;
    sc.lights_host = lights;
    const std::vector<float4> dmats = pack_materials(mats);
    const std::vector<float4> dl = pack_lights(lights);
    hipStream_t s = c->stream;
    sc.tlas.upload(tlas.data(), tlas.size() * sizeof(bn_bvh_node), s);
    sc.blas.upload(nodes.data(), nodes.size() * sizeof(bn_bvh_node), s);
    sc.tris.upload(tris.data(), tris.size() * sizeof(float), s);
    sc.insts.upload(di.data(), di.size() * sizeof(bnd::DevInst), s);
    sc.mats.upload(dmats.data(), dmats.size() * sizeof(float4), s);
    sc.lightdefs.upload(dl.data(), dl.size() * sizeof(float4), s);
    sc.lrec.upload(lrec.data(), lrec.size() * sizeof(int4), s);
    sc.aprob.upload(aprob.data(), aprob.size() * sizeof(float), s);
    sc.apdf.upload(apdf.data(), apdf.size() * sizeof(float), s);
    sc.aidx.upload(aidx.data(), aidx.size() * sizeof(int), s);
    hip_check(hipStreamSynchronize(s), "upload");
    sc.n_lights = (int)lrec.size();
    sc.n_insts = (int)insts.size();
    sc.camera = cam;
    std::memcpy(sc.c2w, c2w, 64);
    sc.dev = bnd::DevScene{};
    sc.dev.tlas = (const float4*)sc.tlas.p;
    sc.dev.blas = (const float4*)sc.blas.p;
    sc.dev.tris = (const float*)sc.tris.p;
    sc.dev.insts = (const bnd::DevInst*)sc.insts.p;
    sc.dev.mats = (const float4*)sc.mats.p;
    sc.dev.lightdefs = (const float4*)sc.lightdefs.p;
    sc.dev.lrec = (const int4*)sc.lrec.p;
    sc.dev.n_lights = sc.n_lights;
    sc.dev.alias_prob = (const float*)sc.aprob.p;
    sc.dev.alias_idx = (const int*)sc.aidx.p;
    sc.dev.alias_pdf = (const float*)sc.apdf.p;
    sc.dev.c2w = to_xf(c2w);
    sc.dev.cam_kind = cam.kind;
    bool finite = true;
    auto fin_nodes = [&](const std::vector<bn_bvh_node>& v) {
        for (const auto& n : v)
            for (int k = 0; k < 3; ++k) finite = finite && std::isfinite(n.pmin[k]) && std::isfinite(n.pmax[k]);
    };
    fin_nodes(tlas);
    fin_nodes(nodes);
    for (float v : tris) finite = finite && std::isfinite(v);
    for (const auto& in : insts)
        for (int k = 0; k < 6; ++k) finite = finite && std::isfinite(in.bounds[k]);
    sc.dev.exact_only = finite ? 0 : 1;
    sc.live = true;
}

// Viewport of PinholeCamera.GenerateRay (Pinhole.fs:12-21) for a film size.
void camera_constants(SceneRec& sc, int W, int H) {
    const bn_camera& c = sc.camera;
    const float vh = 2.0f * (float)std::tan((double)(c.fov * bnh::kPi / 360.0f));
    const float vw = vh * c.aspect;
    const bnh::V3 U = vw * bnh::V3{1, 0, 0}, V = vh * bnh::V3{0, 1, 0};
    const bnh::V3 du = U / (float)W, dv = V / (float)H;
    const bnh::V3 ul = -bnh::V3{0, 0, 1} - 0.5f * (U + V);
    sc.dev.ul = make_float4(ul.x, ul.y, ul.z, c.push_forward);
    sc.dev.du = make_float4(du.x, du.y, du.z, c.aperture);
    sc.dev.dv = make_float4(dv.x, dv.y, dv.z, c.focus_distance);
}

SceneRec& get_scene(Ctx* c, int id) {
    if (!c) throw Error(BN_E_INVALID, "null context");
    if (id < 0 || id >= (int)c->scenes.size() || !c->scenes[id]->live) throw Error(BN_E_INVALID, "unknown scene id");
    return *c->scenes[id];
}

int grid_trace(Ctx* c) { return c->n_cu * c->trace_blocks_per_cu; }
// Batch rounds: no more blocks per CU than their trace kernels keep resident.
int grid_trace_batch(Ctx* c) { return c->n_cu * std::min(c->trace_blocks_per_cu, bnd::kTraceWavesBatch); }

void ensure_paths(Ctx* c, size_t cap) {
    if (cap <= c->path_cap) return;
    auto f4 = [&](DevBuf& b) { b.ensure(cap * sizeof(float4)); };
    f4(c->p_o); f4(c->p_d); f4(c->p_beta); f4(c->p_L); f4(c->p_sh_o); f4(c->p_sh_d); f4(c->p_nee_a); f4(c->p_nee_b);
    f4(c->p2_o); f4(c->p2_d); f4(c->p2_beta); f4(c->p2_L);
    c->p_rng.ensure(cap * 4);
    c->p_item.ensure(cap * 4);
    c->p_pid.ensure(cap * 4);
    c->p2_rng.ensure(cap * 4);
    c->p2_item.ensure(cap * 4);
    c->p2_pid.ensure(cap * 4);
    c->p_hit.ensure(cap * sizeof(int4));
    c->queue[0].ensure(cap * 4);
    c->queue[1].ensure(cap * 4);
    c->freel[0].ensure(cap * 4);
    c->freel[1].ensure(cap * 4);
    c->shq.ensure(cap * 4);
    c->bins.ensure(cap * 4 * bnd::kBins);
    c->path_cap = cap;
    c->graphs.clear();                   // captured graphs point at the old buffers
}

bnd::PathBuf path_buf(Ctx* c) {
    bnd::PathBuf P;
    P.o = (float4*)c->p_o.p; P.d = (float4*)c->p_d.p; P.beta = (float4*)c->p_beta.p; P.L = (float4*)c->p_L.p;
    P.rng = (uint32_t*)c->p_rng.p; P.hit = (int4*)c->p_hit.p;
    P.sh_o = (float4*)c->p_sh_o.p; P.sh_d = (float4*)c->p_sh_d.p;
    P.nee_a = (float4*)c->p_nee_a.p; P.nee_b = (float4*)c->p_nee_b.p;
    P.item = (int*)c->p_item.p;
    P.pid = (uint32_t*)c->p_pid.p;
    return P;
}

// Two spill areas: a shadow traversal may run beside trace_closest.
void ensure_spill(Ctx* c) {
    const size_t lanes = (size_t)grid_trace(c) * bnd::kBlock;
    const size_t need = 2 * lanes * (bnd::kStackMax - bnd::kStackKMin) * 4;
    if (need > c->spill.bytes) c->graphs.clear();
    c->spill.ensure(need);
    c->spill_stride = lanes;
}

bnd::Buffers buffers(Ctx* c) {
    bnd::Buffers B;
    B.queue[0] = (uint32_t*)c->queue[0].p;
    B.queue[1] = (uint32_t*)c->queue[1].p;
    B.freel[0] = (uint32_t*)c->freel[0].p;
    B.freel[1] = (uint32_t*)c->freel[1].p;
    B.shq = (uint32_t*)c->shq.p;
    B.bins = (uint32_t*)c->bins.p;
    B.bin_cap = (uint32_t)c->path_cap;
    B.spill = (uint32_t*)c->spill.p;
    B.spill2 = B.spill + c->spill.bytes / 8;              // the second half
    B.spill_stride = c->spill_stride;
    B.stack_k = c->stack_k;
    B.stats = (unsigned long long*)c->stats.p;
    B.film = (float*)c->film.p;
    B.L_out = (float4*)c->L_out.p;
    c->fetch_seg.ensure((size_t)2 * bnd::kFetchSegs * bnd::kSegStride * 4);
    B.fetch_seg = (uint32_t*)c->fetch_seg.p;
    B.sort = c->material_sort;
    B.alt = bnd::PathState{(float4*)c->p2_o.p, (float4*)c->p2_d.p, (float4*)c->p2_beta.p, (float4*)c->p2_L.p,
                           (uint32_t*)c->p2_rng.p, (int*)c->p2_item.p, (uint32_t*)c->p2_pid.p};
    return B;
}

// Timing (per launch) and exact per-bounce counters of a bounce loop.
struct Hook : bnd::LaunchHook {
    Ctx* c;
    bnd::Ctl* ctl;
    bn_trace_stats* acc;    // non-null: read the counters back after every bounce
    Hook(Ctx* cx, bnd::Ctl* ct, bn_trace_stats* a) : c(cx), ctl(ct), acc(a) {}
    void begin(const char* name, hipStream_t st) override { c->timer.begin(name, st); }
    void end(hipStream_t st) override { c->timer.end(st); }
    void after_bounce(int cur) override {
        if (!acc) return;
        bnd::Ctl h;
        hip_check(hipMemcpyAsync(&h, ctl, sizeof(h), hipMemcpyDeviceToHost, c->stream), "ctl readback");
        hip_check(hipStreamSynchronize(c->stream), "sync");
        acc->rays_closest += h.q_count[cur];
        acc->rays_shadow += h.sh_count + h.shc[cur];     // slot rounds, compact rounds
        for (int b = 0; b < bnd::kBins; ++b) acc->vertices += h.bin_count[b];
    }
};

bnd::RoundItem plain_item(const bn_pt_params& p) {
    bnd::RoundItem it{};
    it.width = p.width; it.height = p.height;
    it.tiles_x = (p.width + 7) / 8;
    it.p_pad = it.tiles_x * ((p.height + 7) / 8) * 64;
    it.spp = p.spp; it.frame_id = p.frame_id;
    it.max_depth = p.max_depth; it.rr_depth = p.rr_depth;
    it.inv_spp = 1.0f / (float)p.spp;                     // MathF.ReciprocalEstimate(spp)
    it.pix_end = p.width * p.height;
    return it;
}

void check_pt_params(const bn_pt_params& p) {
    if (p.width <= 0 || p.height <= 0) throw Error(BN_E_INVALID, "film size must be positive");
    if (p.spp < 0 || p.max_depth < 0) throw Error(BN_E_INVALID, "spp and max-depth must be >= 0");
    if ((int64_t)p.width * p.height > ((int64_t)1 << 28)) throw Error(BN_E_CAPACITY, "film larger than 2^28 pixels");
}

constexpr int kIterBlock = 8;     // iterations between two drain checks (even)
constexpr size_t kMaxGraphs = 64;  // captured graphs kept per context

// Paths per round.  By default a round holds no more paths than there are
// slots, so every path of a round is in flight from the first iteration:
// reusing slots within a round scatters new paths over the slot arrays and
// costs more in uncoalesced path-state access than it saves in bounce tails.
size_t window_paths(const Ctx* c) { return c->window_paths ? c->window_paths : c->chunk_slots; }

// One round: its paths stream through L.n_slots slots until every path has
// retired, checked on the host after every block of kIterBlock iterations
// (a captured graph of one block when `block` is set).
void run_round(Ctx* c, bool batch, bnd::Integrator kind, const bnd::Launch& L, const bnd::PathBuf& P, bnd::Ctl* ctl,
               const bnd::Buffers& B, uint32_t n_paths, int max_depth, bool count, Hook* hook,
               hipGraphExec_t block) {
    hipStream_t s = L.stream;
    bnd::enqueue_round_begin(batch, L, P, ctl, B, hook);
    // every path needs at most max_depth + 1 iterations once it has a slot
    const int64_t limit = ((int64_t)n_paths / L.n_slots + 2) * (int64_t)(max_depth + 2) * 2 + 4 * kIterBlock;
    int64_t iters = 0;
    // A round whose paths all fit its slots issues every path in its first
    // iteration, and a path retires within max_depth iterations of its
    // issue: the first check waits until then (the iterations in between run
    // back to back, with no host synchronisation between blocks).
    const bool all_at_once = n_paths <= L.n_slots;
    for (;;) {
        int n = kIterBlock;
        if (iters == 0 && all_at_once)
            n = std::max(kIterBlock, block ? (max_depth + kIterBlock - 1) / kIterBlock * kIterBlock
                                           : (max_depth + 1) / 2 * 2);
        if (block) {
            for (int r = 0; r < n / kIterBlock; ++r) {
                c->timer.begin("iteration_block", s);
                hip_check(hipGraphLaunch(block, s), "graph launch");
                c->timer.end(s);
            }
        } else {
            bnd::enqueue_iterations(batch, kind, L, P, ctl, B, n, count, hook);
        }
        iters += n;
        bnd::Ctl h;
        hip_check(hipMemcpyAsync(&h, ctl, sizeof(h), hipMemcpyDeviceToHost, s), "ctl readback");
        hip_check(hipStreamSynchronize(s), "iteration block");
        if (h.next_pid >= n_paths && h.q_count[0] == 0) break;
        if (iters > limit) throw Error(BN_E_DEVICE, "round did not drain");
    }
    bnd::enqueue_round_end(batch, kind, L, P, ctl, B, hook);
}

// One frame of PathTracingIntegrator (or the Normal view, or the Direct
// integrator), in rounds of whole samples of all pixels, round after round in
// sample order.
void render_frames(Ctx* c, SceneRec& sc, const bn_pt_params& p, bnd::Integrator kind, float* film_host,
                   bn_trace_stats* stats) {
    const bool normal = kind == bnd::Integrator::Normal;
    check_pt_params(p);
    if (sc.n_lights == 0) throw Error(BN_E_NO_LIGHT, "LightSamplerBase: No light primitives found.");
    hipStream_t s = c->stream;
    camera_constants(sc, p.width, p.height);
    bnd::RoundItem it = plain_item(p);
    if (normal) it.max_depth = 1;                          // one bounce, whatever max-depth says
    // DirectIntegrator ignores max-depth: camera hit + one BSDF-sampled hit
    if (kind == bnd::Integrator::Direct) { it.max_depth = 2; it.rr_depth = 2; }
    const size_t want = std::max<size_t>((size_t)it.p_pad, window_paths(c));
    const int ns = (int)std::max<size_t>(1, std::min<size_t>((size_t)std::max(p.spp, 1), want / it.p_pad));
    const size_t round_paths = (size_t)ns * it.p_pad;
    const size_t slots = std::min(round_paths, c->chunk_slots);
    ensure_paths(c, slots);
    ensure_spill(c);
    c->L_out.ensure(round_paths * sizeof(float4));
    const size_t film_n = (size_t)p.width * p.height * 3;
    c->film.ensure(film_n * 4);
    c->ctl.ensure(sizeof(bnd::Ctl));
    c->stats.ensure(8 * sizeof(unsigned long long));
    hip_check(hipMemsetAsync(c->film.p, 0, film_n * 4, s), "memset film");
    hip_check(hipMemsetAsync(c->stats.p, 0, 8 * sizeof(unsigned long long), s), "memset stats");
    bn_trace_stats acc{};
    bnd::PathBuf P = path_buf(c);
    bnd::Ctl* ctl = (bnd::Ctl*)c->ctl.p;
    Hook hook(c, ctl, stats ? &acc : nullptr);
    bnd::Launch L{};
    L.scene = sc.dev;
    L.grid_trace = grid_trace(c);
    L.grid_shade = c->n_cu * 8;
    L.refill = (uint32_t)c->refill;
    L.octant_order = c->octant_order;
    L.skip_dark = c->skip_dark;
    L.stream = s;
    if (c->overlap_shadow) {
        L.stream2 = c->stream2;
        L.ev_shade = c->ev_shade;
        L.ev_shadow = c->ev_shadow;
    }
    L.n_pix = (uint32_t)(p.width * p.height);
    L.n_slots = (uint32_t)slots;
    const bnd::Buffers B = buffers(c);
    c->fwd.valid = false;
    for (int s0 = 0; s0 < p.spp; s0 += ns) {
        it.s0 = s0;
        it.ns = std::min(ns, p.spp - s0);
        it.path_count = it.ns * it.p_pad;
        L.item = it;
        run_round(c, false, kind, L, P, ctl, B, (uint32_t)it.path_count, it.max_depth, stats != nullptr, &hook,
                  nullptr);
    }
    if (kind == bnd::Integrator::Path && ns >= p.spp) {     // one round: L_out holds the frame
        c->fwd.valid = true;
        c->fwd.scene = &sc;
        c->fwd.gen = sc.gen;
        c->fwd.p = p;
        c->fwd.slots = slots;
    }
    hip_check(hipGetLastError(), "kernel launch");
    if (film_host) hip_check(hipMemcpyAsync(film_host, c->film.p, film_n * 4, hipMemcpyDeviceToHost, s), "film D2H");
    hip_check(hipStreamSynchronize(s), "render");
    c->timer.resolve();
    if (stats) {
        unsigned long long h[8];
        hip_check(hipMemcpy(h, c->stats.p, sizeof(h), hipMemcpyDeviceToHost), "stats D2H");
        acc.nodes_closest = (int64_t)h[0]; acc.tris_closest = (int64_t)h[1]; acc.insts_closest = (int64_t)h[2];
        acc.nodes_shadow = (int64_t)h[3]; acc.tris_shadow = (int64_t)h[4]; acc.insts_shadow = (int64_t)h[5];
        acc.paths = (int64_t)p.width * p.height * p.spp;
        // placeholder paths (padding pixels; every path at max-depth 0) are
        // queued for one iteration but trace nothing
        const int64_t pix = (int64_t)p.width * p.height;
        acc.rays_closest -= (int64_t)p.spp * (it.p_pad - pix) + (p.max_depth <= 0 ? (int64_t)p.spp * pix : 0);
        *stats = acc;
    }
}

// Differentiable-rendering parameters: albedo and roughness of every
// material, emission of every light definition.  Geometry, light records
// and alias tables do not depend on them.
void set_params(Ctx* c, SceneRec& sc, const float* mp, int n_mat, const float* le, int n_light) {
    if (mp) {
        if (n_mat != (int)sc.mats_host.size()) throw Error(BN_E_INVALID, "material parameter count mismatch");
        for (int i = 0; i < n_mat; ++i) {
            bn_material& m = sc.mats_host[i];
            m.albedo[0] = mp[4 * i]; m.albedo[1] = mp[4 * i + 1]; m.albedo[2] = mp[4 * i + 2];
            m.roughness = mp[4 * i + 3];
        }
        const std::vector<float4> d = pack_materials(sc.mats_host);
        sc.mats.upload(d.data(), d.size() * sizeof(float4), c->stream);
        sc.dev.mats = (const float4*)sc.mats.p;
    }
    if (le) {
        if (n_light != (int)sc.lights_host.size()) throw Error(BN_E_INVALID, "light parameter count mismatch");
        for (int i = 0; i < n_light; ++i)
            for (int k = 0; k < 3; ++k) sc.lights_host[i].emission[k] = le[3 * i + k];
        const std::vector<float4> d = pack_lights(sc.lights_host);
        sc.lightdefs.upload(d.data(), d.size() * sizeof(float4), c->stream);
        sc.dev.lightdefs = (const float4*)sc.lightdefs.p;
    }
    hip_check(hipStreamSynchronize(c->stream), "set params");
    c->graphs.clear();
    ++sc.gen;
}

// Gradient of sum(d_film * film) for the frame bn_render_pt renders with the
// same parameters: per round, the forward pass (radiance of every path id),
// then the path replay (kernels_pt.hip, adjoint section).
void backward_pt(Ctx* c, SceneRec& sc, const bn_pt_params& p, const float* d_film, float* d_mat, float* d_light) {
    check_pt_params(p);
    if (!d_film) throw Error(BN_E_INVALID, "null film adjoint");
    if (sc.n_lights == 0) throw Error(BN_E_NO_LIGHT, "LightSamplerBase: No light primitives found.");
    hipStream_t s = c->stream;
    camera_constants(sc, p.width, p.height);
    bnd::RoundItem it = plain_item(p);
    const size_t want = std::max<size_t>((size_t)it.p_pad, window_paths(c));
    const int ns = (int)std::max<size_t>(1, std::min<size_t>((size_t)std::max(p.spp, 1), want / it.p_pad));
    const size_t round_paths = (size_t)ns * it.p_pad;
    const size_t slots = std::min(round_paths, c->chunk_slots);
    ensure_paths(c, slots);
    ensure_spill(c);
    if (slots > c->adj_cap) {
        for (DevBuf* b : {&c->a_delta, &c->a_pend_a, &c->a_pend_r, &c->a_nee_da, &c->a_nee_dr})
            b->ensure(slots * sizeof(float4));
        c->adj_cap = slots;
    }
    c->L_out.ensure(round_paths * sizeof(float4));
    const size_t film_n = (size_t)p.width * p.height * 3;
    const int n_mat = (int)sc.mats_host.size(), n_light = (int)sc.lights_host.size();
    c->film.ensure(film_n * 4);
    c->ctl.ensure(sizeof(bnd::Ctl));
    c->stats.ensure(8 * sizeof(unsigned long long));
    c->a_gmat.ensure((size_t)std::max(1, n_mat) * 16);
    c->a_glight.ensure((size_t)std::max(1, n_light) * 16);
    c->a_dfilm.upload(d_film, film_n * 4, s);
    hip_check(hipMemsetAsync(c->film.p, 0, film_n * 4, s), "memset film");
    hip_check(hipMemsetAsync(c->a_gmat.p, 0, (size_t)std::max(1, n_mat) * 16, s), "memset gradients");
    hip_check(hipMemsetAsync(c->a_glight.p, 0, (size_t)std::max(1, n_light) * 16, s), "memset gradients");
    bnd::PathBuf P = path_buf(c);
    bnd::Ctl* ctl = (bnd::Ctl*)c->ctl.p;
    Hook hook(c, ctl, nullptr);
    bnd::Launch L{};
    L.scene = sc.dev;
    L.grid_trace = grid_trace(c);
    L.grid_shade = c->n_cu * 8;
    L.refill = (uint32_t)c->refill;
    L.octant_order = c->octant_order;
    L.skip_dark = c->skip_dark;
    L.stream = s;
    L.n_pix = (uint32_t)(p.width * p.height);
    L.n_slots = (uint32_t)slots;
    bnd::Buffers B = buffers(c);
    B.adj.delta = (float4*)c->a_delta.p;
    B.adj.pend_a = (float4*)c->a_pend_a.p;
    B.adj.pend_r = (float4*)c->a_pend_r.p;
    B.adj.nee_da = (float4*)c->a_nee_da.p;
    B.adj.nee_dr = (float4*)c->a_nee_dr.p;
    B.adj.L_fwd = (const float4*)c->L_out.p;
    B.adj.d_film = (const float*)c->a_dfilm.p;
    B.adj.g_mat = (float*)c->a_gmat.p;
    B.adj.g_light = (float*)c->a_glight.p;
    B.adj.n_mat = n_mat;
    B.adj.n_light = n_light;
    {
        // per-block gradient sums in LDS when the tables fit a workgroup's 160 KB
        const size_t bytes = (4 * (size_t)n_mat + 3 * (size_t)n_light) * sizeof(float);
        B.adj.lds_bytes = bytes <= 160 * 1024 - 1024 ? (int)bytes : 0;
    }
    // The forward pass of the same frame may still be in L_out (the last
    // bn_render_pt, one round, same scene parameters): then only replay.
    const bool cached = c->fwd.valid && c->fwd.scene == &sc && c->fwd.gen == sc.gen && ns >= p.spp &&
                        c->fwd.slots == slots && std::memcmp(&c->fwd.p, &p, sizeof(p)) == 0;
    for (int s0 = 0; s0 < p.spp; s0 += ns) {
        it.s0 = s0;
        it.ns = std::min(ns, p.spp - s0);
        it.path_count = it.ns * it.p_pad;
        L.item = it;
        if (!cached)
            run_round(c, false, bnd::Integrator::Path, L, P, ctl, B, (uint32_t)it.path_count, it.max_depth, false,
                      &hook, nullptr);
        run_round(c, false, bnd::Integrator::Adjoint, L, P, ctl, B, (uint32_t)it.path_count, it.max_depth, false,
                  &hook, nullptr);
    }
    c->fwd.valid = ns >= p.spp;                             // L_out still holds this frame's forward
    c->fwd.scene = &sc;
    c->fwd.gen = sc.gen;
    c->fwd.p = p;
    c->fwd.slots = slots;
    hip_check(hipGetLastError(), "kernel launch");
    std::vector<float> gm((size_t)std::max(1, n_mat) * 4), gl((size_t)std::max(1, n_light) * 4);
    hip_check(hipMemcpyAsync(gm.data(), c->a_gmat.p, gm.size() * 4, hipMemcpyDeviceToHost, s), "gradient D2H");
    hip_check(hipMemcpyAsync(gl.data(), c->a_glight.p, gl.size() * 4, hipMemcpyDeviceToHost, s), "gradient D2H");
    hip_check(hipStreamSynchronize(s), "backward");
    c->timer.resolve();
    if (d_mat) std::memcpy(d_mat, gm.data(), (size_t)n_mat * 16);
    if (d_light)
        for (int i = 0; i < n_light; ++i)
            for (int k = 0; k < 3; ++k) d_light[3 * i + k] = gl[4 * (size_t)i + k];
}

// Several renders in one launch sequence per round (see kernels_pt.hip).
// Rounds pack whole samples of the items, every item in sample order, up to
// window_paths paths; all paths of a round share the slot pool, so the items
// fill each other's bounce tails.  A block of iterations is captured once as
// a graph and replayed for every block of every round.
void render_batch(Ctx* c, const bn_batch_item* items, int n, float* const* films, bn_batch_stats* stats) {
    if (!items || n <= 0) throw Error(BN_E_INVALID, "empty batch");
    if (n > 4096) throw Error(BN_E_CAPACITY, "more than 4096 batch items");
    hipStream_t s = c->stream;
    std::vector<bnd::RoundItem> base(n);
    std::vector<bnd::DevScene> scenes(n);
    size_t film_total = 0, paths_total = 0;
    size_t max_pad = 0;
    for (int k = 0; k < n; ++k) {
        const bn_pt_params& p = items[k].params;
        check_pt_params(p);
        SceneRec& sc = get_scene(c, items[k].scene);
        if (sc.n_lights == 0) throw Error(BN_E_NO_LIGHT, "LightSamplerBase: No light primitives found.");
        camera_constants(sc, p.width, p.height);
        scenes[k] = sc.dev;                                   // camera constants of this item's film
        base[k] = plain_item(p);
        base[k].scene = k;
        base[k].film_offset = (int)film_total;
        film_total += (size_t)p.width * p.height * 3;
        if (film_total > ((size_t)1 << 31)) throw Error(BN_E_CAPACITY, "batch films exceed 2^31 floats");
        max_pad = std::max(max_pad, (size_t)base[k].p_pad);
        paths_total += (size_t)base[k].p_pad * (size_t)std::max(p.spp, 0);
    }
    const size_t cap = std::max(max_pad, std::min(window_paths(c), paths_total));
    // path ids and RoundItem path counts are int32: rounds stay below 2^31 paths
    if (cap > ((size_t)1 << 30)) throw Error(BN_E_CAPACITY, "round larger than 2^30 paths");
    const size_t slots = std::min(cap, c->chunk_slots);
    ensure_paths(c, slots);
    ensure_spill(c);
    c->L_out.ensure(cap * sizeof(float4));
    c->film.ensure(film_total * 4);
    c->ctl.ensure(sizeof(bnd::Ctl));
    c->stats.ensure(8 * sizeof(unsigned long long));
    c->fwd.valid = false;                                   // L_out is reused below
    c->b_scenes.upload(scenes.data(), scenes.size() * sizeof(bnd::DevScene), s);
    c->b_items.ensure((size_t)n * sizeof(bnd::RoundItem));
    c->b_hdr.ensure(sizeof(bnd::RoundHeader));
    hip_check(hipMemsetAsync(c->film.p, 0, film_total * 4, s), "memset film");
    bnd::PathBuf P = path_buf(c);
    bnd::Ctl* ctl = (bnd::Ctl*)c->ctl.p;
    bnd::Launch L{};
    L.hdr = (const bnd::RoundHeader*)c->b_hdr.p;
    L.items = (const bnd::RoundItem*)c->b_items.p;
    L.scenes = (const bnd::DevScene*)c->b_scenes.p;
    L.grid_trace = grid_trace_batch(c);
    L.grid_uniform = grid_trace(c);
    L.grid_shade = c->n_cu * 8;
    L.refill = (uint32_t)c->refill;
    L.octant_order = c->octant_order;
    L.skip_dark = c->skip_dark;
    L.stream = s;
    L.n_slots = (uint32_t)slots;
    const bnd::Buffers B = buffers(c);
    // One captured block of iterations per trace binding: per ray (mixed
    // rounds), or one scene's geometry in the launch arguments (rounds over
    // one scene; keyed by everything of the scene that the trace kernels read).
    auto block_graph = [&](const bnd::Launch& Lr) -> hipGraphExec_t {
        if (!c->use_graph) return nullptr;
        std::string key = "block:" + std::to_string(L.grid_trace) + ":" + std::to_string(L.refill) + ":" +
                          std::to_string(B.stack_k) + ":" + std::to_string(L.n_slots) + ":" +
                          std::to_string((uintptr_t)c->b_items.p) + ":" + std::to_string((uintptr_t)c->b_scenes.p) +
                          ":" + std::to_string((uintptr_t)c->L_out.p) + ":" + std::to_string((int)L.octant_order) + ":" + std::to_string(L.skip_dark);
        if (Lr.uniform) {
            const bnd::DevScene& d = Lr.scene;
            for (const void* q : {(const void*)d.tlas, (const void*)d.blas, (const void*)d.tris, (const void*)d.insts,
                                  (const void*)d.mats})
                key += ":" + std::to_string((uintptr_t)q);
            key += ":" + std::to_string(d.exact_only) + ":" + std::to_string(Lr.grid_uniform);
        }
        auto f = c->graphs.find(key);
        if (f == c->graphs.end()) {
            if (c->graphs.size() >= kMaxGraphs) c->graphs.clear();   // no earlier round's graph is in flight
            hipGraph_t g;
            hip_check(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal), "begin capture");
            bnd::enqueue_iterations(true, bnd::Integrator::Path, Lr, P, ctl, B, kIterBlock, false, nullptr);
            hip_check(hipStreamEndCapture(s, &g), "end capture");
            hipGraphExec_t e;
            hip_check(hipGraphInstantiate(&e, g, nullptr, nullptr, 0), "graph instantiate");
            hip_check(hipGraphDestroy(g), "graph destroy");
            f = c->graphs.emplace(key, GraphExec(e)).first;
        }
        return f->second.e;
    };
    // An item with at least an eighth of a round of paths gets rounds over
    // its own scene; smaller items share rounds whatever their scenes, so
    // their bounce tails fill each other.
    auto big = [&](int k) { return (size_t)base[k].p_pad * (size_t)items[k].params.spp * 8 >= cap; };
    bool used_graph = false;
    int scene_rounds = 0;
    Hook hook(c, ctl, nullptr);
    std::vector<int> next(n, 0);
    std::vector<bnd::RoundItem> round;
    int rounds = 0;
    for (;;) {
        round.clear();
        size_t used = 0;
        int active = 0;
        for (int k = 0; k < n; ++k) active += next[k] < items[k].params.spp;
        if (active == 0) break;
        // Item-major packing, from the first unfinished item: the samples of
        // its scene's items when it is big, else of the small items; rounds
        // hold whole samples.  Fair-share packing gives every unfinished item
        // a slice.
        std::vector<int> take(n, 0);
        if (c->batch_pack == 0) {
            int first = 0;
            while (next[first] >= items[first].params.spp) ++first;
            const bool own = big(first);
            for (int k = 0; k < n; ++k) {
                const int left = items[k].params.spp - next[k];
                if (left <= 0) continue;
                if (own ? items[k].scene != items[first].scene : big(k)) continue;
                const int t = (int)std::min<size_t>((size_t)left, (cap - used) / (size_t)base[k].p_pad);
                if (t <= 0) break;
                take[k] = t;
                used += (size_t)t * base[k].p_pad;
            }
        } else {
            const size_t share = std::max<size_t>(1, cap / active);
            for (int pass = 0; pass < 2; ++pass)
                for (int k = 0; k < n; ++k) {
                    const int left = items[k].params.spp - next[k] - take[k];
                    if (left <= 0) continue;
                    const size_t pad = (size_t)base[k].p_pad;
                    size_t fit = (cap - used) / pad;
                    if (pass == 0) fit = std::min(fit, std::max<size_t>(1, share / pad));
                    const int t = (int)std::min<size_t>((size_t)left, fit);
                    if (t <= 0) continue;
                    take[k] += t;
                    used += (size_t)t * pad;
                }
        }
        int path = 0, pix = 0, round_depth = 0;
        L.uniform = true;
        int k0 = -1;
        for (int k = 0; k < n; ++k) {
            if (take[k] == 0) continue;
            if (k0 < 0) k0 = k;
            L.uniform = L.uniform && items[k].scene == items[k0].scene;
            bnd::RoundItem it = base[k];
            it.s0 = next[k];
            it.ns = take[k];
            it.path_base = path;
            it.path_count = take[k] * it.p_pad;
            pix += it.width * it.height;
            it.pix_end = pix;
            path += it.path_count;
            next[k] += take[k];
            round_depth = std::max(round_depth, it.max_depth);
            round.push_back(it);
        }
        bnd::RoundHeader h{(int)round.size(), (uint32_t)path, (uint32_t)pix, 0};
        // the stream orders these copies before the round's kernels; run_round
        // synchronises before the host tables are reused
        hip_check(hipMemcpyAsync(c->b_items.p, round.data(), round.size() * sizeof(bnd::RoundItem),
                                 hipMemcpyHostToDevice, s), "round table");
        hip_check(hipMemcpyAsync(c->b_hdr.p, &h, sizeof(h), hipMemcpyHostToDevice, s), "round header");
        L.scene = scenes[k0];
        scene_rounds += L.uniform;
        // A large round over one scene runs eagerly with its shadow rays on the
        // second stream (as a plain frame does); other rounds replay a graph.
        const bool overlap = c->overlap_shadow && L.uniform && path >= (1 << 20);
        L.stream2 = overlap ? c->stream2 : nullptr;
        L.ev_shade = c->ev_shade;
        L.ev_shadow = c->ev_shadow;
        const hipGraphExec_t exec = overlap ? nullptr : block_graph(L);
        used_graph = used_graph || exec;
        run_round(c, true, bnd::Integrator::Path, L, P, ctl, B, (uint32_t)path, round_depth, false, &hook, exec);
        hip_check(hipStreamSynchronize(s), "round");
        ++rounds;
    }
    hip_check(hipGetLastError(), "kernel launch");
    for (int k = 0; k < n; ++k)
        if (films && films[k])
            hip_check(hipMemcpyAsync(films[k], (float*)c->film.p + base[k].film_offset,
                                     (size_t)items[k].params.width * items[k].params.height * 12,
                                     hipMemcpyDeviceToHost, s), "film D2H");
    hip_check(hipStreamSynchronize(s), "batch");
    c->timer.resolve();
    if (stats) {
        stats->rounds = rounds;
        stats->graph = used_graph ? 1 : 0;
        stats->scene_rounds = scene_rounds;
    }
}

// Launch-parameter autotuner: coordinate descent over the persistent trace
// grid (blocks per CU), the lane-refill threshold, the LDS stack window and
// the path slots in flight, timing the caller's own workload (plain renders
// of every item in turn, or one batched render).  Each candidate runs once to
// warm up (allocation, graph capture) and once timed; the best setting stays
// on the context.  Results do not depend on any of these parameters.
void autotune(Ctx* c, const bn_batch_item* items, int n, int flags, bn_tune_result* out) {
    if (!items || n <= 0) throw Error(BN_E_INVALID, "empty workload");
    const bool batched = (flags & BN_TUNE_BATCH) != 0;
    size_t slots_total = 0;
    for (int k = 0; k < n; ++k) {
        check_pt_params(items[k].params);
        const bnd::RoundItem it = plain_item(items[k].params);
        slots_total += (size_t)it.p_pad * (size_t)std::max(items[k].params.spp, 0);
    }
    size_t free_b = 0, total_b = 0;
    hip_check(hipMemGetInfo(&free_b, &total_b), "hipMemGetInfo");
    // path state (two copies of the moving part) + queues + bins per slot
    const size_t per_slot = 12 * sizeof(float4) + sizeof(int4) + 4 * 11 + 4 * bnd::kBins;
    const size_t slot_limit = (free_b + c->path_cap * per_slot) / 4 / per_slot;

    auto run = [&]() -> double {
        const auto t0 = std::chrono::steady_clock::now();
        if (batched) {
            render_batch(c, items, n, nullptr, nullptr);
        } else {
            for (int k = 0; k < n; ++k)
                render_frames(c, get_scene(c, items[k].scene), items[k].params, bnd::Integrator::Path, nullptr,
                              nullptr);
        }
        return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    };
    auto measure = [&]() { run(); return run(); };

    struct Knob { int64_t* cur; std::vector<int64_t> values; };
    int64_t blocks = c->trace_blocks_per_cu, refill = c->refill, stack_k = c->stack_k;
    int64_t chunk = (int64_t)c->chunk_slots;
    auto apply = [&]() {
        c->trace_blocks_per_cu = (int)blocks;
        c->refill = (int)refill;
        c->stack_k = (int)stack_k;
        c->chunk_slots = (size_t)chunk;
    };
    std::vector<int64_t> chunks;
    for (int64_t v = (int64_t)1 << 22; v <= ((int64_t)1 << 26); v <<= 1) {
        if ((size_t)v > slot_limit) break;
        chunks.push_back(v);
        if ((size_t)v >= slots_total) break;          // larger chunks change nothing
    }
    if (chunks.empty()) chunks.push_back(chunk);
    std::vector<Knob> knobs = {{&blocks, {4, 5, 6, 7, 8}}, {&refill, {24, 32, 40, 48}}, {&stack_k, {16, 32}},
                               {&chunk, chunks}};
    apply();
    double best = measure();
    const double baseline = best;
    int trials = 1;
    for (int pass = 0; pass < 2; ++pass) {
        bool changed = false;
        for (Knob& kb : knobs) {
            const int64_t keep = *kb.cur;
            int64_t arg = keep;
            for (int64_t v : kb.values) {
                if (v == keep) continue;
                *kb.cur = v;
                apply();
                const double t = measure();
                ++trials;
                if (t < best * 0.99) { best = t; arg = v; }
            }
            *kb.cur = arg;
            apply();
            changed |= arg != keep;
        }
        if (!changed) break;
    }
    if (out) {
        out->trace_blocks_per_cu = (int32_t)blocks;
        out->refill = (int32_t)refill;
        out->stack_k = (int32_t)stack_k;
        out->trials = trials;
        out->chunk_slots = chunk;
        out->baseline_ms = (float)baseline;
        out->best_ms = (float)best;
    }
}

// PSSMLTIntegrator.Render (Extensions/Integrator/PSSMLT.fs:379-414): the
// bootstrap paths and the chains run as wavefronts on the device; the mean
// bootstrap luminance B (Array.average, sequential) and the bootstrap alias
// table are formed on the host between the two phases, as the reference does.
void render_pssmlt(Ctx* c, SceneRec& sc, const bn_mlt_params& p, float* film_host, bn_mlt_stats* stats) {
    if (p.width <= 0 || p.height <= 0) throw Error(BN_E_INVALID, "film size must be positive");
    if (p.n_bootstrap <= 0 || p.n_chains <= 0 || p.spp < 0 || p.max_depth < 0)
        throw Error(BN_E_INVALID, "n-bootstrap and n-chains must be positive, spp and max-depth >= 0");
    if (p.strategy != BN_MLT_GAUSSIAN && p.strategy != BN_MLT_KELEMEN) throw Error(BN_E_INVALID, "unknown strategy");
    if (sc.n_lights == 0) throw Error(BN_E_NO_LIGHT, "LightSamplerBase: No light primitives found.");
    hipStream_t s = c->stream;
    camera_constants(sc, p.width, p.height);
    const int nx = 4 + 7 * p.max_depth;
    const size_t cap = std::min<size_t>(std::max(p.n_bootstrap, p.n_chains), c->mlt_chunk);
    ensure_paths(c, cap);
    ensure_spill(c);
    if (cap > c->mlt_cap || nx > c->mlt_nx) {
        const size_t k = std::max(cap, c->mlt_cap);
        c->m_inner.ensure(k * 4); c->m_outer.ensure(k * 4); c->m_state.ensure(k * 16); c->m_bid.ensure(k * 4);
        c->m_pix_new.ensure(k * 4); c->m_pix_cur.ensure(k * 4); c->m_Lcur.ensure(k * 16); c->m_rad.ensure(k * 16);
        c->m_acc.ensure(k * 4);
        c->m_xs.ensure(k * (size_t)std::max(nx, c->mlt_nx) * 16);
        c->mlt_cap = k;
        c->mlt_nx = std::max(nx, c->mlt_nx);
    }
    bnd::MltBuf M;
    M.inner = (uint32_t*)c->m_inner.p; M.outer = (uint32_t*)c->m_outer.p; M.state = (int4*)c->m_state.p;
    M.bid = (int*)c->m_bid.p; M.pix_new = (int*)c->m_pix_new.p; M.pix_cur = (int*)c->m_pix_cur.p;
    M.Lcur = (float4*)c->m_Lcur.p; M.rad = (float4*)c->m_rad.p; M.accepted = (int*)c->m_acc.p;
    M.xs = (float4*)c->m_xs.p; M.cap = (uint32_t)cap;
    M.strategy = p.strategy; M.sigma = p.sigma; M.eps_min = p.eps_min; M.eps_max = p.eps_max;
    const size_t film_n = (size_t)p.width * p.height * 3;
    c->film.ensure(film_n * 4);
    c->ctl.ensure(sizeof(bnd::Ctl));
    c->m_weights.ensure((size_t)p.n_bootstrap * 4);
    c->m_accepted_out.ensure((size_t)p.n_chains * 4);
    hip_check(hipMemsetAsync(c->film.p, 0, film_n * 4, s), "memset film");     // Film.Clear
    bnd::PathBuf P = path_buf(c);
    bnd::Ctl* ctl = (bnd::Ctl*)c->ctl.p;
    const int gt = grid_trace(c), gs = c->n_cu * 8;
    c->stats.ensure(8 * sizeof(unsigned long long));
    const bnd::Buffers BF = buffers(c);
    KernelTimer& T = c->timer;
    bnd::MltParamsDev dp{p.width, p.height, p.frame_id, p.n_bootstrap, p.large_step_prob, 0.0f, 0.0f};
    const uint32_t bin_cap = (uint32_t)c->path_cap;

    // One MLT path per slot [0, n): start, bounce loop.  With a second
    // stream the shadow rays of bounce d are traced on it while bounce d + 1
    // resets, traces and bins on the first; shade(d + 1) waits for them (it
    // adds to the same radiance and reuses the shadow arrays), so every path
    // adds its terms in the reference's order.  Shadow counts and any-hit
    // fetch counters are kept per bounce parity, the any-hit traversal spills
    // to its own area, and the streams are joined after the last bounce.
    const bool overlap = c->overlap_shadow;
    hipStream_t s2 = overlap ? c->stream2 : s;
    auto run_paths = [&](bool fresh, int id0, int n) {
        bnd::launch_ctl_clear(ctl, s);
        { Timed tm(T, "mlt_start", s);
          bnd::launch_mlt_start(fresh, sc.dev, P, M, dp, ctl, (uint32_t*)c->queue[0].p, id0, n, gs, s); }
        int cur = 0;
        for (int depth = 0; depth < p.max_depth; ++depth) {
            const int nxt = 1 - cur;
            const int par = overlap ? (depth & 1) : -1;
            { Timed tm(T, "reset", s); bnd::launch_reset(ctl, nxt, BF.fetch_seg, s, par); }
            { Timed tm(T, "trace_closest", s);
              bnd::launch_trace_closest(false, sc.dev, P, ctl, cur, (const uint32_t*)c->queue[cur].p, BF,
                                        (uint32_t)c->refill, gt, s); }
            { Timed tm(T, "bin", s);
              bnd::launch_bin(P, ctl, cur, (const uint32_t*)c->queue[cur].p, (uint32_t*)c->bins.p, bin_cap, nullptr,
                              c->material_sort > 0, gs, s); }
            if (overlap) hip_check(hipStreamWaitEvent(s, c->ev_shadow, 0), "wait shadow");
            { Timed tm(T, "shade_mlt", s);
              bnd::launch_shade_mlt(fresh, sc.dev, P, M, ctl, (const uint32_t*)c->bins.p, bin_cap,
                                    (uint32_t*)c->queue[nxt].p, nxt, (uint32_t*)c->shq.p,
                                    depth, p.max_depth, p.rr_depth, gs, s, par); }
            if (overlap) {
                hip_check(hipEventRecord(c->ev_shade, s), "record shade");
                hip_check(hipStreamWaitEvent(s2, c->ev_shade, 0), "wait shade");
            }
            { Timed tm(T, "trace_shadow", s2);
              bnd::launch_trace_shadow(false, sc.dev, P, ctl, (const uint32_t*)c->shq.p, BF, (uint32_t)c->refill,
                                       gt, s2, par, overlap ? BF.spill2 : nullptr); }
            if (overlap) hip_check(hipEventRecord(c->ev_shadow, s2), "record shadow");
            cur = nxt;
        }
        if (overlap) hip_check(hipStreamWaitEvent(s, c->ev_shadow, 0), "join shadow");
    };

    // PSSMLT.fs:382-392: bootstrap
    for (int id0 = 0; id0 < p.n_bootstrap; id0 += (int)cap) {
        const int n = std::min<int>((int)cap, p.n_bootstrap - id0);
        run_paths(true, id0, n);
        bnd::launch_mlt_boot_done(P, (float*)c->m_weights.p, id0, n, s);
    }
    c->last_bootstrap.resize(p.n_bootstrap);
    hip_check(hipMemcpyAsync(c->last_bootstrap.data(), c->m_weights.p, (size_t)p.n_bootstrap * 4,
                             hipMemcpyDeviceToHost, s), "weights D2H");
    hip_check(hipStreamSynchronize(s), "bootstrap");
    float sum = 0.0f;
    for (float w : c->last_bootstrap) sum = sum + w;              // Array.average (PSSMLT.fs:394)
    const float B = sum / (float)p.n_bootstrap;
    const uint64_t mpc64 = ((uint64_t)p.spp * (uint64_t)p.width * (uint64_t)p.height + (uint64_t)p.n_chains - 1) /
                           (uint64_t)p.n_chains;
    if (mpc64 > 0x7fffffffull) throw Error(BN_E_CAPACITY, "mutations per chain exceed int32");
    const int mpc = (int)mpc64;
    c->last_chain_accepts.assign(p.n_chains, 0);
    int64_t accepted = 0;
    if (B != 0.0f) {                                              // PSSMLT.fs:396-397
        const bnh::AliasTable tab = bnh::build_alias(c->last_bootstrap.data(), p.n_bootstrap);   // :399
        c->m_aprob.upload(tab.prob.data(), tab.prob.size() * 4, s);
        c->m_aidx.upload(tab.alias.data(), tab.alias.size() * 4, s);
        dp.inv_b = 1.0f / B;
        dp.inv_eff = 1.0f / ((float)mpc * (float)p.n_chains / (float)(p.width * p.height));
        for (int c0 = 0; c0 < p.n_chains; c0 += (int)cap) {
            const int n = std::min<int>((int)cap, p.n_chains - c0);
            bnd::launch_mlt_chain_init(M, dp, (const float*)c->m_aprob.p, (const int*)c->m_aidx.p, c0, n, s);
            run_paths(false, c0, n);                                 // replay of the bootstrap path
            bnd::launch_mlt_replay_done(P, M, dp, c0, n, s);
            for (int m = 0; m < mpc; ++m) {
                run_paths(false, c0, n);
                { Timed tm(T, "mlt_accept", s); bnd::launch_mlt_accept(P, M, dp, (float*)c->film.p, n, s); }
            }
            bnd::launch_mlt_finish(M, dp, (float*)c->film.p, (int*)c->m_accepted_out.p, c0, n, s);
        }
        hip_check(hipMemcpyAsync(c->last_chain_accepts.data(), c->m_accepted_out.p, (size_t)p.n_chains * 4,
                                 hipMemcpyDeviceToHost, s), "accepts D2H");
    }
    hip_check(hipGetLastError(), "kernel launch");
    if (film_host) hip_check(hipMemcpyAsync(film_host, c->film.p, film_n * 4, hipMemcpyDeviceToHost, s), "film D2H");
    hip_check(hipStreamSynchronize(s), "pssmlt");
    c->timer.resolve();
    for (int32_t a : c->last_chain_accepts) accepted += a;
    if (stats) {
        stats->accepted = accepted;
        stats->proposed = B != 0.0f ? (int64_t)mpc * p.n_chains : 0;
        stats->b = B;
        stats->mutations_per_chain = mpc;
    }
}

}  // namespace bnx

using namespace bnx;

extern "C" {

const char* bn_last_error(void) { return g_error.c_str(); }
int bn_abi_version(void) { return 1; }

int bn_struct_sizes(int32_t* out, int32_t n) {
    const int32_t v[] = {(int32_t)sizeof(bn_primitive), (int32_t)sizeof(bn_material), (int32_t)sizeof(bn_light),
                         (int32_t)sizeof(bn_instance), (int32_t)sizeof(bn_transform), (int32_t)sizeof(bn_node),
                         (int32_t)sizeof(bn_camera), (int32_t)sizeof(bn_scene_desc), (int32_t)sizeof(bn_pt_params),
                         (int32_t)sizeof(bn_mlt_params), (int32_t)sizeof(bn_mlt_stats),
                         (int32_t)sizeof(bn_trace_stats), (int32_t)sizeof(bn_bvh_node),
                         (int32_t)sizeof(bn_instance_upload), (int32_t)sizeof(bn_batch_item),
                         (int32_t)sizeof(bn_batch_stats), (int32_t)sizeof(bn_kernel_time),
                         (int32_t)sizeof(bn_tune_result)};
    const int m = (int)(sizeof(v) / sizeof(v[0]));
    for (int i = 0; i < n && i < m; ++i) out[i] = v[i];
    return m;
}

int bn_create(int device, bn_ctx** out) {
    return guard([&] {
        if (!out) throw Error(BN_E_INVALID, "null output pointer");
        int n = 0;
        hip_check(hipGetDeviceCount(&n), "hipGetDeviceCount");
        if (device < 0 || device >= n) throw Error(BN_E_DEVICE, "no such HIP device");
        hip_check(hipSetDevice(device), "hipSetDevice");
        auto c = std::make_unique<Ctx>();
        c->device = device;
        hipDeviceProp_t prop;
        hip_check(hipGetDeviceProperties(&prop, device), "hipGetDeviceProperties");
        c->n_cu = prop.multiProcessorCount;
        hip_check(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking), "hipStreamCreate");
        hip_check(hipStreamCreateWithFlags(&c->stream2, hipStreamNonBlocking), "hipStreamCreate");
        hip_check(hipEventCreateWithFlags(&c->ev_shade, hipEventDisableTiming), "hipEventCreate");
        hip_check(hipEventCreateWithFlags(&c->ev_shadow, hipEventDisableTiming), "hipEventCreate");
        noalloc_init(device);
        *out = reinterpret_cast<bn_ctx*>(c.release());
    });
}

void bn_destroy(bn_ctx* ctx) {
    Ctx* c = reinterpret_cast<Ctx*>(ctx);
    if (!c) return;
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    delete c;
}

int bn_bvh_build(const float* b, int32_t n, bn_bvh_node* out_nodes, int32_t* out_n_nodes, int32_t* out_perm) {
    return guard([&] {
        if (n <= 0 || !b || !out_nodes || !out_n_nodes || !out_perm) throw Error(BN_E_INVALID, "bad arguments");
        std::vector<bnh::Box> items(n);
        for (int i = 0; i < n; ++i)
            items[i] = {{b[6 * i], b[6 * i + 1], b[6 * i + 2]}, {b[6 * i + 3], b[6 * i + 4], b[6 * i + 5]}};
        std::vector<bn_bvh_node> nodes;
        std::vector<int32_t> perm;
        bnh::build_bvh(items, nodes, perm);
        if ((int)nodes.size() > *out_n_nodes) throw Error(BN_E_CAPACITY, "node buffer too small");
        std::memcpy(out_nodes, nodes.data(), nodes.size() * sizeof(bn_bvh_node));
        *out_n_nodes = (int)nodes.size();
        std::memcpy(out_perm, perm.data(), perm.size() * 4);
    });
}

int bn_bvh_build_gpu(bn_ctx* ctx, const float* b, int32_t n, bn_bvh_node* out_nodes, int32_t* out_n_nodes,
                     int32_t* out_perm) {
    return guard([&] {
        Ctx* c = reinterpret_cast<Ctx*>(ctx);
        if (!c || n <= 0 || !b || !out_nodes || !out_n_nodes || !out_perm) throw Error(BN_E_INVALID, "bad arguments");
        hip_check(hipSetDevice(c->device), "hipSetDevice");
        std::vector<std::vector<bn_bvh_node>> nodes;
        std::vector<std::vector<int32_t>> perm;
        bvh_build_gpu(c, {b}, {n}, nodes, perm);
        if ((int)nodes[0].size() > *out_n_nodes) throw Error(BN_E_CAPACITY, "node buffer too small");
        std::memcpy(out_nodes, nodes[0].data(), nodes[0].size() * sizeof(bn_bvh_node));
        *out_n_nodes = (int)nodes[0].size();
        std::memcpy(out_perm, perm[0].data(), perm[0].size() * 4);
    });
}

int bn_upload_blas(bn_ctx* ctx, const bn_bvh_node* nodes, int32_t n_nodes, const float* verts, int32_t n_verts,
                   const int32_t* tri_idx, int32_t n_tris, const float* alias_pdf, int32_t* out_blas) {
    return guard([&] {
        Ctx* c = reinterpret_cast<Ctx*>(ctx);
        if (!c || !nodes || n_nodes <= 0 || !verts || n_verts <= 0 || !tri_idx || n_tris <= 0 || !out_blas)
            throw Error(BN_E_INVALID, "bad arguments");
        BlasRec b;
        b.nodes.assign(nodes, nodes + n_nodes);
        b.tris.resize(9 * (size_t)n_tris);
        for (int i = 0; i < n_tris; ++i)
            for (int j = 0; j < 3; ++j) {
                int v = tri_idx[3 * i + j];
                if (v < 0 || v >= n_verts) throw Error(BN_E_INVALID, "triangle index out of range");
                for (int k = 0; k < 3; ++k) b.tris[9 * (size_t)i + 3 * j + k] = verts[3 * (size_t)v + k];
            }
        std::vector<float> w(n_tris);
        if (alias_pdf) {
            // The F# host passes AliasTable.Table[i].pdf (Mesh.fs:181-186),
            // kept verbatim: EvalPDF reads Table[0].pdf (Mesh.fs:300-304) and
            // Sample returns Table[i].pdf.  The reference's table over
            // non-negative weights is degenerate (AliasTable.fs:20,45-50: every
            // prob <= 1, so every entry ends with prob 1 and alias = itself).
            b.alias.pdf.assign(alias_pdf, alias_pdf + n_tris);
            b.alias.prob.assign(n_tris, 1.0f);
            b.alias.alias.resize(n_tris);
            for (int i = 0; i < n_tris; ++i) b.alias.alias[i] = i;
        } else {
            for (int i = 0; i < n_tris; ++i) {
                const float* t = &b.tris[9 * (size_t)i];
                w[i] = bnh::tri_area({t[0], t[1], t[2]}, {t[3], t[4], t[5]}, {t[6], t[7], t[8]});
            }
            b.alias = bnh::build_alias(w.data(), n_tris);
        }
        validate_bvh(b.nodes, n_tris, "BLAS");
        c->blas.push_back(std::move(b));
        *out_blas = (int)c->blas.size() - 1;
    });
}

int bn_release_blas(bn_ctx* ctx, int32_t blas) {
    return guard([&] {
        Ctx* c = reinterpret_cast<Ctx*>(ctx);
        if (!c) throw Error(BN_E_INVALID, "null context");
        if (blas < 0 || blas >= (int)c->blas.size() || !c->blas[blas].live || c->blas[blas].owner >= 0)
            throw Error(BN_E_INVALID, "not a live BLAS from bn_upload_blas");
        release_blas(c->blas[blas]);
    });
}

int bn_upload_scene(bn_ctx* ctx, const bn_bvh_node* tlas, int32_t n_tlas, const bn_instance_upload* instances,
                    int32_t n_instances, const bn_material* materials, int32_t n_materials, const bn_light* lights,
                    int32_t n_lights, const bn_camera* camera, const float* c2w, int32_t* out_scene) {
    return guard([&] {
        Ctx* c = reinterpret_cast<Ctx*>(ctx);
        if (!c || !tlas || n_tlas <= 0 || !instances || n_instances <= 0 || !camera || !c2w || !out_scene ||
            n_materials < 0 || n_lights < 0 || (n_materials > 0 && !materials) || (n_lights > 0 && !lights))
            throw Error(BN_E_INVALID, "bad arguments");
        hip_check(hipSetDevice(c->device), "hipSetDevice");
        auto sc = std::make_unique<SceneRec>();
        upload_scene(c, *sc, std::vector<bn_bvh_node>(tlas, tlas + n_tlas),
                     std::vector<bn_instance_upload>(instances, instances + n_instances),
                     n_materials ? std::vector<bn_material>(materials, materials + n_materials)
                                 : std::vector<bn_material>(),
                     n_lights ? std::vector<bn_light>(lights, lights + n_lights) : std::vector<bn_light>(), *camera,
                     c2w);
        c->scenes.push_back(std::move(sc));
        *out_scene = (int)c->scenes.size() - 1;
    });
}

int bn_scene_build(bn_ctx* ctx, const bn_scene_desc* desc, int32_t* out_scene) {
    return guard([&] {
        Ctx* c = reinterpret_cast<Ctx*>(ctx);
        if (!c || !desc || !out_scene) throw Error(BN_E_INVALID, "bad arguments");
        hip_check(hipSetDevice(c->device), "hipSetDevice");
        auto hs = std::make_shared<bnh::HostScene>();
        const bnh::BatchBvhBuild gpu_build = [c](const std::vector<const float*>& sets, const std::vector<int>& counts,
                                                 std::vector<std::vector<bn_bvh_node>>& nodes,
                                                 std::vector<std::vector<int32_t>>& perms) {
            bvh_build_gpu(c, sets, counts, nodes, perms);
        };
        bnh::build_scene(*desc, *hs, c->bvh_gpu ? &gpu_build : nullptr);
        std::vector<int> blas_of(hs->meshes.size(), -1);
        for (size_t i = 0; i < hs->meshes.size(); ++i) {
            const bnh::HostMesh& m = hs->meshes[i];
            if (m.sphere) continue;
            BlasRec b;
            b.nodes = m.nodes;
            const int nt = (int)(m.tri.size() / 3);
            b.tris.resize(9 * (size_t)nt);
            for (int k = 0; k < nt; ++k)
                for (int j = 0; j < 3; ++j) {
                    const bnh::V3 v = m.verts[m.tri[3 * k + j]];
                    b.tris[9 * (size_t)k + 3 * j] = v.x;
                    b.tris[9 * (size_t)k + 3 * j + 1] = v.y;
                    b.tris[9 * (size_t)k + 3 * j + 2] = v.z;
                }
            b.alias = m.alias;
            b.owner = (int)c->scenes.size();      // the id this scene gets below
            validate_bvh(b.nodes, nt, "BLAS");
            c->blas.push_back(std::move(b));
            blas_of[i] = (int)c->blas.size() - 1;
        }
        std::vector<bn_instance_upload> ups(hs->insts.size());
        for (size_t i = 0; i < hs->insts.size(); ++i) {
            const bnh::HostInstance& h = hs->insts[i];
            bn_instance_upload& u = ups[i];
            std::memcpy(u.o2w, h.o2w.m, 64);
            std::memcpy(u.w2o, h.w2o.m, 64);
            u.bounds[0] = h.bounds.lo.x; u.bounds[1] = h.bounds.lo.y; u.bounds[2] = h.bounds.lo.z;
            u.bounds[3] = h.bounds.hi.x; u.bounds[4] = h.bounds.hi.y; u.bounds[5] = h.bounds.hi.z;
            u.blas = blas_of[h.prim];
            u.sphere_radius = hs->meshes[h.prim].sphere ? hs->meshes[h.prim].radius : 0.0f;
            u.material = h.material;
            u.light = h.light;
        }
        auto sc = std::make_unique<SceneRec>();
        try {
            upload_scene(c, *sc, hs->tlas, ups, hs->materials, hs->light_defs, hs->camera, hs->c2w.m);
        } catch (...) {
            for (int b : blas_of)
                if (b >= 0) release_blas(c->blas[b]);
            throw;
        }
        sc->host = hs;
        c->scenes.push_back(std::move(sc));
        *out_scene = (int)c->scenes.size() - 1;
    });
}

int bn_scene_release(bn_ctx* ctx, int32_t scene) {
    return guard([&] {
        Ctx* c = reinterpret_cast<Ctx*>(ctx);
        SceneRec& s = get_scene(c, scene);
        hip_check(hipStreamSynchronize(c->stream), "sync");
        if (c->fwd.scene == &s) c->fwd.valid = false;
        c->scenes[scene] = std::make_unique<SceneRec>();
        // the BLAS this scene's build made; its device copies went with the scene
        for (BlasRec& b : c->blas)
            if (b.live && b.owner == scene) release_blas(b);
    });
}

int bn_scene_export(bn_ctx* ctx, int32_t scene, bn_bvh_node* tlas, int32_t* n_tlas, int32_t* inst_order,
                    int32_t* n_inst, float* o2w, float* w2o, int32_t* light_insts, int32_t* n_lights,
                    float* c2w) {
    return guard([&] {
        Ctx* c = reinterpret_cast<Ctx*>(ctx);
        SceneRec& s = get_scene(c, scene);
        if (!s.host) throw Error(BN_E_INVALID, "scene was not built by bn_scene_build");
        const bnh::HostScene& h = *s.host;
        if ((int)h.tlas.size() > *n_tlas || (int)h.insts.size() > *n_inst || (int)h.lights.size() > *n_lights)
            throw Error(BN_E_CAPACITY, "export buffer too small");
        std::memcpy(tlas, h.tlas.data(), h.tlas.size() * sizeof(bn_bvh_node));
        *n_tlas = (int)h.tlas.size();
        for (size_t i = 0; i < h.insts.size(); ++i) {
            inst_order[i] = h.inst_desc[i];
            std::memcpy(o2w + 16 * i, h.insts[i].o2w.m, 64);
            std::memcpy(w2o + 16 * i, h.insts[i].w2o.m, 64);
        }
        *n_inst = (int)h.insts.size();
        std::memcpy(light_insts, h.lights.data(), h.lights.size() * 4);
        *n_lights = (int)h.lights.size();
        std::memcpy(c2w, h.c2w.m, 64);
    });
}

int bn_scene_export_blas(bn_ctx* ctx, int32_t scene, int32_t prim, bn_bvh_node* nodes, int32_t* n_nodes,
                         int32_t* tri, int32_t* n_tris, float* pdf) {
    return guard([&] {
        Ctx* c = reinterpret_cast<Ctx*>(ctx);
        SceneRec& s = get_scene(c, scene);
        if (!s.host) throw Error(BN_E_INVALID, "scene was not built by bn_scene_build");
        const bnh::HostScene& h = *s.host;
        if (prim < 0 || prim >= (int)h.meshes.size() || h.meshes[prim].sphere)
            throw Error(BN_E_INVALID, "not a mesh primitive");
        const bnh::HostMesh& m = h.meshes[prim];
        const int nt = (int)(m.tri.size() / 3);
        if ((int)m.nodes.size() > *n_nodes || nt > *n_tris) throw Error(BN_E_CAPACITY, "export buffer too small");
        std::memcpy(nodes, m.nodes.data(), m.nodes.size() * sizeof(bn_bvh_node));
        *n_nodes = (int)m.nodes.size();
        std::memcpy(tri, m.tri.data(), m.tri.size() * 4);
        std::memcpy(pdf, m.alias.pdf.data(), nt * 4);
        *n_tris = nt;
    });
}

int bn_scene_export_instances(bn_ctx* ctx, int32_t scene, bn_instance_upload* out, int32_t* n) {
    return guard([&] {
        Ctx* c = reinterpret_cast<Ctx*>(ctx);
        SceneRec& s = get_scene(c, scene);
        if (!s.host) throw Error(BN_E_INVALID, "scene was not built by bn_scene_build");
        if (!out || !n) throw Error(BN_E_INVALID, "bad arguments");
        const bnh::HostScene& h = *s.host;
        if ((int)h.insts.size() > *n) throw Error(BN_E_CAPACITY, "export buffer too small");
        for (size_t i = 0; i < h.insts.size(); ++i) {
            const bnh::HostInstance& hi = h.insts[i];
            bn_instance_upload& u = out[i];
            std::memcpy(u.o2w, hi.o2w.m, 64);
            std::memcpy(u.w2o, hi.w2o.m, 64);
            u.bounds[0] = hi.bounds.lo.x; u.bounds[1] = hi.bounds.lo.y; u.bounds[2] = hi.bounds.lo.z;
            u.bounds[3] = hi.bounds.hi.x; u.bounds[4] = hi.bounds.hi.y; u.bounds[5] = hi.bounds.hi.z;
            const bool sphere = h.meshes[hi.prim].sphere;
            u.blas = sphere ? -1 : hi.prim;
            u.sphere_radius = sphere ? h.meshes[hi.prim].radius : 0.0f;
            u.material = hi.material;
            u.light = hi.light;
        }
        *n = (int)h.insts.size();
    });
}

int bn_film_postprocess(const float* film, int64_t n_pixels, int32_t tone_mapping, float* out_rgb,
                        uint8_t* out_rgba8) {
    return guard([&] {
        if (!film || n_pixels < 0 || (!out_rgb && !out_rgba8)) throw Error(BN_E_INVALID, "bad arguments");
        if (tone_mapping < BN_TONE_IDENTITY || tone_mapping > BN_TONE_GAMMA)
            throw Error(BN_E_INVALID, "Unknown tone mapping type");
        bnh::postprocess(film, n_pixels, tone_mapping, out_rgb, out_rgba8);
    });
}

int bn_render_pt(bn_ctx* ctx, int32_t scene, const bn_pt_params* params, float* film, bn_trace_stats* stats) {
    return guard([&] {
        Ctx* c = reinterpret_cast<Ctx*>(ctx);
        if (!params) throw Error(BN_E_INVALID, "null params");
        SceneRec& s = get_scene(c, scene);
        hip_check(hipSetDevice(c->device), "hipSetDevice");
        render_frames(c, s, *params, bnd::Integrator::Path, film, stats);
    });
}

int bn_render_normal(bn_ctx* ctx, int32_t scene, const bn_pt_params* params, float* film) {
    return guard([&] {
        Ctx* c = reinterpret_cast<Ctx*>(ctx);
        if (!params) throw Error(BN_E_INVALID, "null params");
        SceneRec& s = get_scene(c, scene);
        hip_check(hipSetDevice(c->device), "hipSetDevice");
        render_frames(c, s, *params, bnd::Integrator::Normal, film, nullptr);
    });
}

int bn_render_direct(bn_ctx* ctx, int32_t scene, const bn_pt_params* params, float* film) {
    return guard([&] {
        Ctx* c = reinterpret_cast<Ctx*>(ctx);
        if (!params) throw Error(BN_E_INVALID, "null params");
        SceneRec& s = get_scene(c, scene);
        hip_check(hipSetDevice(c->device), "hipSetDevice");
        render_frames(c, s, *params, bnd::Integrator::Direct, film, nullptr);
    });
}

int bn_first_hit(bn_ctx* ctx, int32_t scene, const bn_pt_params* p, int32_t sample_id, int32_t* inst,
                 int32_t* prim, float* t) {
    return guard([&] {
        Ctx* c = reinterpret_cast<Ctx*>(ctx);
        if (!p || !inst || !prim || !t) throw Error(BN_E_INVALID, "bad arguments");
        if (p->width <= 0 || p->height <= 0) throw Error(BN_E_INVALID, "film size must be positive");
        SceneRec& s = get_scene(c, scene);
        hip_check(hipSetDevice(c->device), "hipSetDevice");
        camera_constants(s, p->width, p->height);
        ensure_spill(c);
        const size_t n = (size_t)p->width * p->height;
        c->out_a.ensure(n * 4);
        c->out_b.ensure(n * 4);
        c->out_c.ensure(n * 4);
        c->ctl.ensure(sizeof(bnd::Ctl));
        hip_check(hipMemsetAsync(c->ctl.p, 0, sizeof(bnd::Ctl), c->stream), "memset ctl");
        bnd::launch_first_hit(s.dev, p->width, p->height, sample_id, p->spp, p->frame_id, (int*)c->out_a.p,
                              (int*)c->out_b.p, (float*)c->out_c.p, (uint32_t*)c->spill.p, c->spill_stride,
                              &((bnd::Ctl*)c->ctl.p)->fetch[0], grid_trace(c), c->stream);
        hip_check(hipGetLastError(), "kernel launch");
        hip_check(hipMemcpyAsync(inst, c->out_a.p, n * 4, hipMemcpyDeviceToHost, c->stream), "D2H");
        hip_check(hipMemcpyAsync(prim, c->out_b.p, n * 4, hipMemcpyDeviceToHost, c->stream), "D2H");
        hip_check(hipMemcpyAsync(t, c->out_c.p, n * 4, hipMemcpyDeviceToHost, c->stream), "D2H");
        hip_check(hipStreamSynchronize(c->stream), "first hit");
    });
}

int bn_first_hit_noalloc(bn_ctx* ctx, int32_t scene, const bn_pt_params* p, int32_t sample_id, int32_t* inst,
                         int32_t* prim, float* t) {
    return guard([&] {
        Ctx* c = reinterpret_cast<Ctx*>(ctx);
        if (!p || !inst || !prim || !t) throw Error(BN_E_INVALID, "bad arguments");
        SceneRec& s = get_scene(c, scene);
        hip_check(hipSetDevice(c->device), "hipSetDevice");
        first_hit_noalloc(c, s, *p, sample_id, inst, prim, t);
    });
}

}  // extern "C"

extern "C" int bn_set_option(bn_ctx* ctx, const char* key, int64_t value) {
    return guard([&] {
        Ctx* c = reinterpret_cast<Ctx*>(ctx);
        if (!c || !key) throw Error(BN_E_INVALID, "bad arguments");
        const std::string k(key);
        if (k == "timing") {
            c->timer.on = value != 0;
        } else if (k == "chunk_slots") {
            if (value < 64 || value > ((int64_t)1 << 30)) throw Error(BN_E_INVALID, "chunk_slots out of range");
            c->chunk_slots = (size_t)value;
        } else if (k == "trace_blocks_per_cu") {
            if (value < 1 || value > 8) throw Error(BN_E_INVALID, "trace_blocks_per_cu must be in [1, 8]");
            c->trace_blocks_per_cu = (int)value;
        } else if (k == "mlt_chunk") {
            if (value < 64 || value > ((int64_t)1 << 26)) throw Error(BN_E_INVALID, "mlt_chunk out of range");
            c->mlt_chunk = (size_t)value;
        } else if (k == "graph") {
            c->use_graph = value != 0;
        } else if (k == "window_paths") {
            if (value != 0 && (value < 64 || value > ((int64_t)1 << 30)))
                throw Error(BN_E_INVALID, "window_paths out of range [64, 2^30]");
            c->window_paths = (size_t)value;
        } else if (k == "material_sort") {
            if (value < -1 || value > 1) throw Error(BN_E_INVALID, "material_sort must be -1, 0 or 1");
            c->material_sort = (int)value;
            c->graphs.clear();
        } else if (k == "stack_k") {
            if (value != 16 && value != 32) throw Error(BN_E_INVALID, "stack_k must be 16 or 32");
            c->stack_k = (int)value;
        } else if (k == "batch_pack") {
            if (value < 0 || value > 1) throw Error(BN_E_INVALID, "batch_pack must be 0 or 1");
            c->batch_pack = (int)value;
        } else if (k == "bvh_gpu") {
            if (value < 0 || value > 1) throw Error(BN_E_INVALID, "bvh_gpu must be 0 or 1");
            c->bvh_gpu = value != 0;
        } else if (k == "skip_dark") {
            if (value < 0 || value > 2) throw Error(BN_E_INVALID, "skip_dark must be 0, 1 or 2");
            c->skip_dark = (int)value;
        } else if (k == "octant_order") {
            if (value < 0 || value > 1) throw Error(BN_E_INVALID, "octant_order must be 0 or 1");
            c->octant_order = value != 0;
        } else if (k == "overlap_shadow") {
            if (value < 0 || value > 1) throw Error(BN_E_INVALID, "overlap_shadow must be 0 or 1");
            c->overlap_shadow = value != 0;
        } else if (k == "bvh_small") {
            if (value < 128 || value > 512) throw Error(BN_E_INVALID, "bvh_small must be in [128, 512]");
            c->bvh_small = (int)value;
        } else if (k == "refill") {
            if (value < 1 || value > 64) throw Error(BN_E_INVALID, "refill must be in [1, 64]");
            c->refill = (int)value;
        } else {
            throw Error(BN_E_INVALID, "unknown option " + k);
        }
    });
}

extern "C" int bn_kernel_times(bn_ctx* ctx, bn_kernel_time* out, int32_t* n) {
    return guard([&] {
        Ctx* c = reinterpret_cast<Ctx*>(ctx);
        if (!c || !out || !n) throw Error(BN_E_INVALID, "bad arguments");
        const auto& t = c->timer.totals;
        if ((int)t.size() > *n) throw Error(BN_E_CAPACITY, "kernel time buffer too small");
        for (size_t i = 0; i < t.size(); ++i) {
            std::memset(out[i].name, 0, sizeof(out[i].name));
            std::strncpy(out[i].name, t[i].first.c_str(), sizeof(out[i].name) - 1);
            out[i].launches = t[i].second.first;
            out[i].total_ms = t[i].second.second;
        }
        *n = (int)t.size();
    });
}

extern "C" int bn_reset_kernel_times(bn_ctx* ctx) {
    return guard([&] {
        Ctx* c = reinterpret_cast<Ctx*>(ctx);
        if (!c) throw Error(BN_E_INVALID, "null context");
        c->timer.totals.clear();
    });
}

extern "C" int bn_render_pssmlt(bn_ctx* ctx, int32_t scene, const bn_mlt_params* params, float* film,
                                bn_mlt_stats* stats) {
    return guard([&] {
        Ctx* c = reinterpret_cast<Ctx*>(ctx);
        if (!params) throw Error(BN_E_INVALID, "null params");
        SceneRec& s = get_scene(c, scene);
        hip_check(hipSetDevice(c->device), "hipSetDevice");
        render_pssmlt(c, s, *params, film, stats);
    });
}

extern "C" int bn_pssmlt_last(bn_ctx* ctx, float* bootstrap, int32_t n_bootstrap, int32_t* chain_accepts,
                              int32_t n_chains) {
    return guard([&] {
        Ctx* c = reinterpret_cast<Ctx*>(ctx);
        if (!c) throw Error(BN_E_INVALID, "null context");
        if (bootstrap) {
            if (n_bootstrap != (int32_t)c->last_bootstrap.size()) throw Error(BN_E_INVALID, "bootstrap size mismatch");
            std::memcpy(bootstrap, c->last_bootstrap.data(), c->last_bootstrap.size() * 4);
        }
        if (chain_accepts) {
            if (n_chains != (int32_t)c->last_chain_accepts.size()) throw Error(BN_E_INVALID, "chain count mismatch");
            std::memcpy(chain_accepts, c->last_chain_accepts.data(), c->last_chain_accepts.size() * 4);
        }
    });
}

extern "C" int bn_render_batch(bn_ctx* ctx, const bn_batch_item* items, int32_t n, float* const* films,
                               bn_batch_stats* stats) {
    return guard([&] {
        Ctx* c = reinterpret_cast<Ctx*>(ctx);
        if (!c) throw Error(BN_E_INVALID, "null context");
        hip_check(hipSetDevice(c->device), "hipSetDevice");
        render_batch(c, items, n, films, stats);
    });
}

extern "C" int bn_autotune(bn_ctx* ctx, const bn_batch_item* items, int32_t n, int32_t flags, bn_tune_result* out) {
    return guard([&] {
        Ctx* c = reinterpret_cast<Ctx*>(ctx);
        if (!c) throw Error(BN_E_INVALID, "null context");
        hip_check(hipSetDevice(c->device), "hipSetDevice");
        autotune(c, items, n, flags, out);
    });
}

extern "C" int bn_set_params(bn_ctx* ctx, int32_t scene, const float* mat_params, int32_t n_mat,
                             const float* light_emission, int32_t n_light) {
    return guard([&] {
        Ctx* c = reinterpret_cast<Ctx*>(ctx);
        SceneRec& sc = get_scene(c, scene);
        hip_check(hipSetDevice(c->device), "hipSetDevice");
        set_params(c, sc, mat_params, n_mat, light_emission, n_light);
    });
}

extern "C" int bn_backward_pt(bn_ctx* ctx, int32_t scene, const bn_pt_params* params, const float* d_film,
                              float* d_mat, float* d_light) {
    return guard([&] {
        Ctx* c = reinterpret_cast<Ctx*>(ctx);
        if (!params) throw Error(BN_E_INVALID, "null params");
        SceneRec& sc = get_scene(c, scene);
        hip_check(hipSetDevice(c->device), "hipSetDevice");
        backward_pt(c, sc, *params, d_film, d_mat, d_light);
    });
}
