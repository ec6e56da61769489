// Device BVH build (bvh_gpu.hip): linear BVH + refit + 4-wide collapse.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>

#include "gi_internal.h"

namespace gi {

struct LbvhResult {
  int64_t nnodes = 0;   // wide nodes written
  int32_t depth = 0;    // wide levels below the root
};

// Builds the 4-wide BVH of n triangles whose vertices a, b, c are float4
// 0, 1, 2 of records `stride` float4 apart (the ShadeTri table), over the
// centroid bounds clo..chi.  Writes nodes[0..nnodes) (capacity >= n) and
// the 3n float4 of the leaf-ordered intersection triangles.  Synchronises
// the stream; scratch memory is allocated and freed inside.
bool lbvh_build(hipStream_t s, const float4 *tri, uint32_t stride, uint32_t n, const float clo[3],
                const float chi[3], int max_leaf, WideNode *nodes, float4 *itris, LbvhResult &res,
                std::string &err);

}  // namespace gi
