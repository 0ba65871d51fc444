// bvh_gpu.hpp -- binned-SAH BVH construction on the GPU (bvh_gpu.hip).
#pragma once
#include <vector>

#include "../../include/barnacle_hip.h"

namespace bnx {

struct Ctx;

// Builds one BVH per item set in a single GPU pass over all sets: sets[s]
// holds counts[s] items of 6 floats (pmin xyz, pmax xyz), all finite.  The
// nodes and permutation of each set are those of bnh::build_bvh (the
// reference's BVHNode.Build), bit for bit.  Runs on the context's stream and
// returns after the results are on the host.
void bvh_build_gpu(Ctx* c, const std::vector<const float*>& sets, const std::vector<int>& counts,
                   std::vector<std::vector<bn_bvh_node>>& nodes_out, std::vector<std::vector<int32_t>>& perm_out);

}  // namespace bnx
