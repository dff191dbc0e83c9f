// Host-side state of the renderers (one per context): the loaded scene's
// device buffers and the wavefront path state, shared by the path tracer
// (rx_render.hip) and the bidirectional tracer (rx_bpt.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <vector>

#include "rx_pathtrace.h"

namespace rx {

struct SceneBufs {
  TreeBufs global;  // the context's tree is used; this holds volume trees
  std::vector<TreeBufs*> vols;
  DBuf vol_trees, vol_obj, objs, etri, enrm, ecdf;
  DevScene dev{};
  uint32_t max_depth = 0;
  std::vector<DevObject> host_objs;
  ~SceneBufs() {
    for (auto* v : vols) delete v;
  }
};

struct RenderState {
  SceneBufs scene;
  DBuf o, d, beta, lastp, L, s, hit, q0, q1, sq, cnt, so, sd, sc, ss, sum, sumsq, mean, stats;
  DBuf rec, bcount, total;                // guided-mode training records
  DBuf gquery, ghead, grow;               // guided mode: network query and decoded output per row, row per entry
  DBuf bverts, bnv, bvis, bpairs, bsplat, bown;  // bidirectional tracer: subpath vertices, counts, visibility, splats, own sums
  uint32_t cap = 0;
  float box_lo[3], box_hi[3];
};

// The bidirectional tracer's render (rx_render with RX_MODE_BPT).
int render_bpt(rx_ctx* c, const rx_camera_desc* cam, const rx_render_params* prm, float* rgb_mean, double* sum,
               double* sumsq, rx_render_stats* stats);

}  // namespace rx
