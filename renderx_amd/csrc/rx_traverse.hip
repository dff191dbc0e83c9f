// Batch nearest-hit kernel: the HIP replacement of RenderX's GLSL traversal
// (Shader/Intersection.frag:25-84, one fragment per ray) with the CPU
// semantics of KDTree::intersect (src/KDTree.cpp:295-339), plus the shader's
// per-ray last_tid skip (Shader/Intersection.frag:63-64).
//
// Each wave takes chunks of the batch from shared counters, one ray per
// lane, and lanes take the chunk's next ray as theirs ends (trace_range); the
// traversal stack is a per-lane LDS column (conflict-free: lane-contiguous)
// with a per-thread global spill for deep trees.
#include "rx_kernels.h"

namespace rx {

template <bool EXH, bool LEG, bool STATS>
__global__ void __launch_bounds__(kTraceBlock) k_intersect(DevTree T, IsectArgs a) {
  __shared__ uint2 stack[kLdsDepth * kTraceBlock];
  const uint32_t gtid = blockIdx.x * blockDim.x + threadIdx.x;
  bool ovf = false;
  trace_range<EXH, LEG, STATS, kLdsDepth, true>(
      T, a.n, a.work, stack + threadIdx.x, a.spill + (size_t)gtid * a.spill_depth, a.spill_depth, &ovf,
      [&](uint32_t i, Vec3& o, Vec3& d, int32_t& ins, int32_t& con, TreeRoot&, int32_t& skip) {
        o = v3(a.org[3 * (size_t)i], a.org[3 * (size_t)i + 1], a.org[3 * (size_t)i + 2]);
        d = v3(a.dir[3 * (size_t)i], a.dir[3 * (size_t)i + 1], a.dir[3 * (size_t)i + 2]);
        if (LEG) {
          ins = a.inside[i];
          con = a.contact[i];
        }
        skip = a.skip ? a.skip[i] : -1;
      },
      [&](uint32_t i, Hit h, const TraceStats& st) {
        a.dist[i] = h.dist;
        a.tid[i] = h.pos >= 0 ? __float_as_int(T.tv0[h.pos].w) : -1;
        if (a.bary) {
          float u = -1.0f, w = -1.0f;
          if (h.pos >= 0) {  // FP64 solve of the ray/plane point (Moller-Trumbore in double)
            const Vec3 o = v3(a.org[3 * (size_t)i], a.org[3 * (size_t)i + 1], a.org[3 * (size_t)i + 2]);
            const Vec3 d = v3(a.dir[3 * (size_t)i], a.dir[3 * (size_t)i + 1], a.dir[3 * (size_t)i + 2]);
            float4 A = T.tv0[h.pos], B = T.tv1[h.pos], C = T.tv2[h.pos];
            double e1[3] = {(double)B.x - A.x, (double)B.y - A.y, (double)B.z - A.z};
            double e2[3] = {(double)C.x - A.x, (double)C.y - A.y, (double)C.z - A.z};
            double s[3] = {(double)o.x - A.x, (double)o.y - A.y, (double)o.z - A.z};
            double dd[3] = {d.x, d.y, d.z};
            double p[3] = {dd[1] * e2[2] - dd[2] * e2[1], dd[2] * e2[0] - dd[0] * e2[2], dd[0] * e2[1] - dd[1] * e2[0]};
            double q[3] = {s[1] * e1[2] - s[2] * e1[1], s[2] * e1[0] - s[0] * e1[2], s[0] * e1[1] - s[1] * e1[0]};
            double det = p[0] * e1[0] + p[1] * e1[1] + p[2] * e1[2];
            u = (float)((p[0] * s[0] + p[1] * s[1] + p[2] * s[2]) / det);
            w = (float)((q[0] * dd[0] + q[1] * dd[1] + q[2] * dd[2]) / det);
          }
          a.bary[2 * (size_t)i] = u;
          a.bary[2 * (size_t)i + 1] = w;
        }
        if (STATS && a.counts) {
          a.counts[2 * (size_t)i] = st.boxes;
          a.counts[2 * (size_t)i + 1] = st.tris;
        }
      });
  if (ovf) *a.overflow = 1;
}

int intersect_blocks_per_cu() {
  int nb = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_intersect<false, true, false>, kTraceBlock, 0) !=
          hipSuccess ||
      nb < 1)
    nb = 4;
  return nb;
}

hipError_t launch_intersect(const DevTree& T, const IsectArgs& a, uint32_t flags, int grid_blocks,
                            hipStream_t s) {
  const bool exh = flags & 1u, leg = flags & 2u, stats = a.counts != nullptr;
  dim3 g(grid_blocks), b(kTraceBlock);
#define RX_LAUNCH(E, L, S) hipLaunchKernelGGL((k_intersect<E, L, S>), g, b, 0, s, T, a)
  if (exh) {
    if (leg) { if (stats) RX_LAUNCH(true, true, true); else RX_LAUNCH(true, true, false); }
    else { if (stats) RX_LAUNCH(true, false, true); else RX_LAUNCH(true, false, false); }
  } else {
    if (leg) { if (stats) RX_LAUNCH(false, true, true); else RX_LAUNCH(false, true, false); }
    else { if (stats) RX_LAUNCH(false, false, true); else RX_LAUNCH(false, false, false); }
  }
#undef RX_LAUNCH
  return hipGetLastError();
}

}  // namespace rx
