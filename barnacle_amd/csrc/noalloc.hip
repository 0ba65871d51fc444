// noalloc.hip -- the allocation-free first-hit query.
//
// bn_first_hit_noalloc is bn_first_hit with every buffer its kernel needs
// beyond the caller's outputs -- the queue fetch counter, the per-pixel result
// staging and the traversal-stack spill area -- carved out of one static
// device array instead of context allocations, so a caller that must not
// allocate device memory (code under HIP graph capture, tight serving loops,
// harnesses that forbid hipMalloc) can still issue first-hit queries once the
// scene is resident.  The array is created with the code object on every GPU
// the library runs on and lives until the process ends; calls that overlap on
// one GPU share it and race.
#include <hip/hip_runtime.h>

#include <mutex>
#include <vector>

#include "abi_internal.hpp"

namespace bnx {

constexpr int kNoAllocBlocks = 256;                    // persistent traversal blocks
constexpr size_t kNoAllocLanes = (size_t)kNoAllocBlocks * bnd::kBlock;
constexpr size_t kNoAllocHdr = 64;                     // fetch counter, own 256-byte line
constexpr size_t kNoAllocSpill = kNoAllocLanes * (bnd::kStackMax - bnd::kStackK);   // words
constexpr size_t kNoAllocWords = kNoAllocHdr + 3 * kNoAllocMaxPix + kNoAllocSpill;

}  // namespace bnx

// Layout (32-bit words): [0, 64) fetch counter; then instance, primitive and
// t of up to kNoAllocMaxPix pixels; then the spill area of kNoAllocLanes
// lanes (stack entries beyond the 32-entry LDS window of k_first_hit).
// 40 MiB in all.
static __device__ __align__(256) uint32_t scratch[bnx::kNoAllocWords];

namespace bnx {

namespace {
std::mutex g_scratch_mu;
std::vector<uint32_t*> g_scratch;   // device address of `scratch` per device id
}  // namespace

// Once per device, with that device current (bn_create): the array's device
// address on this GPU.  Host code uses only this address, never the array's
// name, which on the host is a stand-in for the symbol.
void noalloc_init(int device) {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if ((int)g_scratch.size() <= device) g_scratch.resize(device + 1, nullptr);
    if (g_scratch[device]) return;
    void* p = nullptr;
    hip_check(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
    if (!p) throw bnh::Error(BN_E_DEVICE, "hipGetSymbolAddress(scratch) returned null");
    g_scratch[device] = static_cast<uint32_t*>(p);
}

static uint32_t* scratch_ptr(int device) {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (device < 0 || device >= (int)g_scratch.size() || !g_scratch[device])
        throw bnh::Error(BN_E_DEVICE, "static scratch not set up on this device");
    return g_scratch[device];
}

// The closest hits of the camera rays of sample `sample_id` of every pixel,
// as bn_first_hit computes them, with no device allocation: the scene must
// already be resident, and the film may hold at most kNoAllocMaxPix pixels.
void first_hit_noalloc(Ctx* c, SceneRec& s, const bn_pt_params& p, int sample_id, int32_t* inst, int32_t* prim,
                       float* t) {
    if (p.width <= 0 || p.height <= 0) throw bnh::Error(BN_E_INVALID, "film size must be positive");
    const size_t n = (size_t)p.width * (size_t)p.height;
    if (n > kNoAllocMaxPix)
        throw bnh::Error(BN_E_CAPACITY, "bn_first_hit_noalloc: film larger than 2^20 pixels (the static scratch "
                                        "holds the results of at most 1048576 pixels)");
    uint32_t* base = scratch_ptr(c->device);
    uint32_t* fetch = base;
    int* out_inst = reinterpret_cast<int*>(base + kNoAllocHdr);
    int* out_prim = out_inst + kNoAllocMaxPix;
    float* out_t = reinterpret_cast<float*>(out_prim + kNoAllocMaxPix);
    uint32_t* spill = reinterpret_cast<uint32_t*>(out_t + kNoAllocMaxPix);
    camera_constants(s, p.width, p.height);
    hipStream_t st = c->stream;
    // the fetch counter must start at zero; results and spill are written
    // before they are read
    hip_check(hipMemsetAsync(fetch, 0, kNoAllocHdr * 4, st), "memset scratch counter");
    bnd::launch_first_hit(s.dev, p.width, p.height, sample_id, p.spp, p.frame_id, out_inst, out_prim, out_t, spill,
                          kNoAllocLanes, fetch, kNoAllocBlocks, st);
    hip_check(hipGetLastError(), "kernel launch");
    hip_check(hipMemcpyAsync(inst, out_inst, n * 4, hipMemcpyDeviceToHost, st), "D2H");
    hip_check(hipMemcpyAsync(prim, out_prim, n * 4, hipMemcpyDeviceToHost, st), "D2H");
    hip_check(hipMemcpyAsync(t, out_t, n * 4, hipMemcpyDeviceToHost, st), "D2H");
    hip_check(hipStreamSynchronize(st), "first hit");
}

}  // namespace bnx
