// Host-callable launchers of the HIP kernels (defined in the .hip files).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "rx_dev.h"

namespace rx {

constexpr int kTraceBlock = 128;   // threads per traversal block (2 waves)
constexpr int kLdsDepth = 16;      // LDS stack entries per lane

struct IsectArgs {
  uint32_t n;
  const float* org;
  const float* dir;
  const int32_t* inside;
  const int32_t* contact;
  float* dist;
  int32_t* tid;
  float* bary;
  uint32_t* counts;
  uint2* spill;        // grid_threads * spill_depth entries
  int spill_depth;     // tree depth beyond the LDS stack
  int32_t* overflow;   // set to 1 on stack overflow
};

// Grid-stride nearest-hit over a batch; grid_blocks bounds the spill buffer.
hipError_t launch_intersect(const DevTree& T, const IsectArgs& a, uint32_t flags, int grid_blocks,
                            hipStream_t s);

}  // namespace rx
