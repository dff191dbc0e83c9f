// The rx_ctx behind the C ABI: device buffers, streams and loaded scene state.
#pragma once
#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "../../include/rxhip.h"
#include "rx_dev.h"
#include "rx_tree.h"

namespace rx {

// rx_debug_fail_device_alloc: the countdown to a device allocation that is
// made to fail (negative: off).
extern int g_fail_alloc_countdown;

struct DBuf {
  void* p = nullptr;
  size_t bytes = 0;
  hipError_t ensure(size_t n) {
    if (n <= bytes && p) return hipSuccess;
    release();
    if (n == 0) n = 16;
    if (g_fail_alloc_countdown >= 0 && g_fail_alloc_countdown-- == 0) return hipErrorOutOfMemory;
    hipError_t e = hipMalloc(&p, n);
    if (e != hipSuccess) {
      p = nullptr;
      return e;
    }
    bytes = n;
    return hipSuccess;
  }
  hipError_t upload(const void* src, size_t n, hipStream_t s) {
    hipError_t e = ensure(n);
    if (e != hipSuccess) return e;
    if (n == 0) return hipSuccess;
    e = hipMemcpyAsync(p, src, n, hipMemcpyHostToDevice, s);
    if (e != hipSuccess) return e;
    return hipStreamSynchronize(s);
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
  }
  template <class T>
  T* as() const { return reinterpret_cast<T*>(p); }
  ~DBuf() { release(); }
  DBuf() = default;
  DBuf(const DBuf&) = delete;
  DBuf& operator=(const DBuf&) = delete;
};

// A tree resident in HBM.
struct TreeBufs {
  DBuf nodes, leaf_start, tv0, tv1, tv2, tn;
  DevTree dev{};
  uint32_t max_depth = 0;
  uint32_t n_nodes = 0;
  hipError_t upload(const DeviceTreeHost& h, uint32_t depth, hipStream_t s);
};

struct RenderState;  // wavefront buffers (rx_render.hip)
struct GuideState;   // guiding network (rx_guide.hip)

}  // namespace rx

struct rx_ctx {
  int device = 0;
  int n_cu = 256;
  int isect_blocks_per_cu = 0, render_blocks_per_cu = 0;  // resident traversal blocks (0: not queried yet)
  rx_render_tuning tuning{};                               // rx_set_render_tuning (zero = defaults)
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  std::string err;
  // tree loaded by rx_load_tree / rx_load_scene
  rx::HostTree host_tree;
  std::vector<float> leaf_verts;
  std::vector<int32_t> tri_obj;     // by triangle id (empty: unknown)
  std::vector<float> tri_normal;    // by triangle id, optional caller normals
  rx::TreeBufs tree;
  bool tree_loaded = false;
  // traversal scratch
  rx::DBuf spill, overflow, io_org, io_dir, io_in, io_con, io_skip, io_dist, io_tid, io_bary, io_counts;
  bool isect_queued = false;    // an rx_intersect_dev whose overflow flag rx_synchronize has not read yet
  bool noalloc_queued = false;  // the same for rx_intersect_dev_noalloc (flag at the start of `scratch`)
  void* scratch = nullptr;      // this device's address of rx_intersect_dev_noalloc's static scratch
  float last_ms = 0.0f;
  // scene
  bool scene_loaded = false;
  rx::RenderState* render = nullptr;
  rx::GuideState* guide = nullptr;
};

namespace rx {
int set_err(rx_ctx* c, int code, const std::string& msg);
int hip_err(rx_ctx* c, hipError_t e, const char* where);
int ensure_spill(rx_ctx* c, int grid_blocks, int block, int depth_needed, int* spill_depth);
int upload_tree(rx_ctx* c);
}  // namespace rx
