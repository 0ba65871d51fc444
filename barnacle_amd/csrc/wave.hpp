// wave.hpp -- wave and workgroup helpers shared by the kernels (gfx950,
// 64-lane waves).
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace bnd {

__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }
// Lanes below this one whose bit is set in m.
__device__ __forceinline__ uint32_t mbcnt(uint64_t m) {
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}
// Octant of a ray direction (sign bits of x, y, z).
__device__ __forceinline__ int octant(float4 d) {
    return (d.x < 0.0f ? 1 : 0) | (d.y < 0.0f ? 2 : 0) | (d.z < 0.0f ? 4 : 0);
}

// Block-aggregated positions in two queues, each block's entries grouped by
// a 3-bit key (thread order within a key).  For queue o: rank[o] is the
// thread's rank among the block's entries when p[o] holds, base[o] the
// position of the block's first entry and n[o] the block's entry count.
// Four ballots per queue and wave give the per-key masks; one wave scans the
// 2 x 8 keys x 4 waves counts.  Every thread of the 256-thread block must call
// it; the caller synchronises the block before the next call.
__device__ __forceinline__ void block_rank_oct(const bool (&p)[2], const int (&key)[2], uint32_t* const (&ctr)[2],
                                               uint32_t (&rank)[2], uint32_t (&base)[2], uint32_t (&n)[2]) {
    constexpr int NW = 4;
    __shared__ uint32_t cnt[2][8][NW];
    __shared__ uint32_t tot[2][2];
    const int w = (int)(threadIdx.x >> 6), l = (int)lane_id();
    uint64_t mk[2];
    for (int o = 0; o < 2; ++o) {
        const uint64_t mv = __ballot(p[o]);
        const uint64_t m0 = __ballot(key[o] & 1), m1 = __ballot(key[o] & 2), m2 = __ballot(key[o] & 4);
        auto of_key = [&](int b) {
            return mv & (b & 1 ? m0 : ~m0) & (b & 2 ? m1 : ~m1) & (b & 4 ? m2 : ~m2);
        };
        mk[o] = of_key(key[o]);
        if (l < 8) cnt[o][l][w] = (uint32_t)__popcll(of_key(l));
    }
    __syncthreads();
    if (w == 0) {
        // lanes [0, 32): queue 0, [32, 64): queue 1; entry j = key * 4 + wave
        const int o = l >> 5, j = l & 31;
        const uint32_t c = cnt[o][j >> 2][j & 3];
        uint32_t incl = c;
        for (int off = 1; off < 32; off <<= 1) {
            const uint32_t u = __shfl_up(incl, off);
            if (j >= off) incl += u;
        }
        const uint32_t total = __shfl(incl, o * 32 + 31);
        uint32_t b0 = 0;
        if (j == 0 && total) b0 = atomicAdd(ctr[o], total);
        b0 = __shfl(b0, o * 32);
        cnt[o][j >> 2][j & 3] = incl - c;
        if (j == 0) {
            tot[o][0] = total;
            tot[o][1] = b0;
        }
    }
    __syncthreads();
    for (int o = 0; o < 2; ++o) {
        rank[o] = cnt[o][key[o]][w] + mbcnt(mk[o]);
        n[o] = tot[o][0];
        base[o] = tot[o][1];
    }
}

}  // namespace bnd
