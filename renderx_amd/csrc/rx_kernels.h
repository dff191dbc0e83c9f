// Host-callable launchers of the HIP kernels (defined in the .hip files).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "rx_dev.h"

namespace rx {

constexpr int kTraceBlock = 128;   // threads per traversal block (2 waves)
#ifndef RX_LDS_DEPTH
#define RX_LDS_DEPTH 10
#endif
constexpr int kLdsDepth = RX_LDS_DEPTH;  // LDS stack entries per lane
// the smallest LDS stack depth of any extend-kernel launch shape: the
// per-thread spill buffer is sized for it
constexpr int kMinLdsDepth = kLdsDepth < 6 ? kLdsDepth : 6;

struct IsectArgs {
  uint32_t n;
  const float* org;
  const float* dir;
  const int32_t* inside;
  const int32_t* contact;
  const int32_t* skip;  // per-ray triangle id never hit (IntersectionGPU::Ray::last_tid), or NULL
  float* dist;
  int32_t* tid;
  float* bary;
  uint32_t* counts;
  uint2* spill;        // grid_threads * spill_depth entries
  int spill_depth;     // tree depth beyond the LDS stack
  int32_t* overflow;   // set to 1 on stack overflow
  uint32_t* work;      // shared work counter, zero at launch
};

// Nearest hit over a batch in grid_blocks contiguous ranges (grid_blocks
// also sizes the spill buffer).
hipError_t launch_intersect(const DevTree& T, const IsectArgs& a, uint32_t flags, int grid_blocks,
                            hipStream_t s);
// Resident blocks per CU of the batch traversal kernel (LDS / registers).
int intersect_blocks_per_cu();

}  // namespace rx
