// Device BVH build (bvh_gpu.hip): a binary tree (linear BVH + refit, PLOC
// clustering or binned SAH) collapsed into 4-wide nodes.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>

#include "gi_internal.h"

namespace gi {

struct LbvhResult {
  int64_t nnodes = 0;   // wide nodes written
  int32_t depth = 0;    // wide levels below the root
  int32_t rounds = 0;   // PLOC clustering rounds (0 for the radix tree)
};

enum class DeviceBvhMethod { kLbvh = 0, kPloc = 1, kSah = 2 };

// Builds the 4-wide BVH of n triangles whose vertices a, b, c are float4
// 0, 1, 2 of records `stride` float4 apart (the ShadeTri table), over the
// centroid bounds clo..chi.  Writes nodes[0..nnodes) (capacity >= n) and
// the 3n float4 of the leaf-ordered intersection triangles.  kLbvh: the
// radix tree over the Morton keys; kPloc: clustering with the given search
// radius (1..64); kSah: the host builder's binned SAH (clo, chi and radius
// unused).  width: child slots filled per wide node (2..4).  Synchronises
// the stream; scratch memory is allocated and freed inside.
// The 3n float4 of the leaf-ordered intersection triangles (a with the
// original index in w, b - a, c - a) from the ShadeTri-strided vertex table
// and a leaf-order permutation already on the device.  Asynchronous.
hipError_t launch_leaf_tris(hipStream_t s, const float4 *tri, uint32_t stride, uint32_t n, const uint32_t *order,
                            float4 *itris);

bool device_bvh_build(hipStream_t s, const float4 *tri, uint32_t stride, uint32_t n, const float clo[3],
                      const float chi[3], int max_leaf, DeviceBvhMethod method, int radius, int width,
                      WideNode *nodes, float4 *itris, LbvhResult &res, std::string &err);

}  // namespace gi
