// C ABI of librxhip.so: context, tree and batch intersection entry points.
// Scene loading and rendering live in rx_render.hip.
#include <algorithm>
#include <cstring>
#include <mutex>
#include <new>

#include "rx_context.h"
#include "rx_kernels.h"

struct rx_tree {
  rx::HostTree t;
  std::vector<float> verts;
  std::vector<uint32_t> tris;
};

namespace rx {

int g_fail_alloc_countdown = -1;

// Traversal scratch of rx_intersect_dev_noalloc: its work counters and
// overflow flag, then its traversal-stack spill, in a file-scope device array
// instead of per-call allocations.  Each GPU the library runs on holds its
// own copy until the process ends.  Host code never names the array in a
// kernel argument or copy: it uses the per-device address from
// hipGetSymbolAddress (scratch_addr).
constexpr size_t kScratchBytes = size_t(32) << 20;
constexpr size_t kScratchHeader = 4096;  // overflow flag + work counters, then the spill
static_assert(4 * (1 + kWorkWords) <= kScratchHeader, "work counters fit the scratch header");
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static void* g_scratch_addr[kMaxDevices];

// The device address of `scratch` on `device` (current), looked up once per device.
static int scratch_addr(int device, void** out) {
  if (device < 0 || device >= kMaxDevices) return RX_ERR_INVALID;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_addr[device]) {
    void* p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return RX_ERR_HIP;
    g_scratch_addr[device] = p;
  }
  *out = g_scratch_addr[device];
  return RX_OK;
}

__global__ void k_zero_words(uint32_t* p, uint32_t n) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = 0u;
}

int set_err(rx_ctx* c, int code, const std::string& msg) {
  if (c) c->err = msg;
  return code;
}

int hip_err(rx_ctx* c, hipError_t e, const char* where) {
  if (e == hipSuccess) return RX_OK;
  std::string m = std::string(where) + ": " + hipGetErrorString(e);
  return set_err(c, e == hipErrorOutOfMemory ? RX_ERR_NO_MEMORY : RX_ERR_HIP, m);
}

hipError_t TreeBufs::upload(const DeviceTreeHost& h, uint32_t depth, hipStream_t s) {
  hipError_t e;
  if ((e = nodes.upload(h.nodes.data(), h.nodes.size() * 4, s)) != hipSuccess) return e;
  if ((e = leaf_start.upload(h.leaf_start.data(), h.leaf_start.size() * 4, s)) != hipSuccess) return e;
  if ((e = tv0.upload(h.tv0.data(), h.tv0.size() * 4, s)) != hipSuccess) return e;
  if ((e = tv1.upload(h.tv1.data(), h.tv1.size() * 4, s)) != hipSuccess) return e;
  if ((e = tv2.upload(h.tv2.data(), h.tv2.size() * 4, s)) != hipSuccess) return e;
  if ((e = tn.upload(h.tn.data(), h.tn.size() * 4, s)) != hipSuccess) return e;
  dev.nodes = nodes.as<float4>();
  dev.leaf_start = leaf_start.as<int32_t>();
  dev.tv0 = tv0.as<float4>();
  dev.tv1 = tv1.as<float4>();
  dev.tv2 = tv2.as<float4>();
  dev.tn = tn.as<float4>();
  memcpy(dev.root_lo, h.root_lo, 12);
  memcpy(dev.root_hi, h.root_hi, 12);
  dev.root_ref = h.root_ref;
  dev.n_leaf_tris = (uint32_t)(h.tv0.size() / 4);
  max_depth = depth;
  n_nodes = h.n_interior + h.n_leaves;
  return hipSuccess;
}

int ensure_spill(rx_ctx* c, int grid_blocks, int block, int depth_needed, int* spill_depth) {
  int sd = std::max(1, depth_needed + 2 - kMinLdsDepth);
  size_t bytes = (size_t)grid_blocks * block * sd * sizeof(uint2);
  if (int r = hip_err(c, c->spill.ensure(bytes), "spill alloc")) return r;
  if (int r = hip_err(c, c->overflow.ensure(4 * (1 + kWorkWords)), "overflow alloc")) return r;
  *spill_depth = sd;
  return RX_OK;
}

int upload_tree(rx_ctx* c) {
  DeviceTreeHost dh;
  std::vector<uint32_t> local;  // object-local triangle index by triangle id
  const uint32_t ntri = (uint32_t)c->host_tree.leaf_tris.size();
  uint32_t maxtid = 0;
  for (uint32_t t : c->host_tree.leaf_tris) maxtid = std::max(maxtid, t + 1);
  const int32_t* obj = nullptr;
  if (!c->tri_obj.empty()) {
    if (c->tri_obj.size() < maxtid) return set_err(c, RX_ERR_INVALID, "legality arrays shorter than triangle ids");
    obj = c->tri_obj.data();
    local.resize(c->tri_obj.size());
    int32_t prev = -2;
    uint32_t base = 0;
    for (uint32_t t = 0; t < (uint32_t)c->tri_obj.size(); t++) {
      if (c->tri_obj[t] != prev) {
        prev = c->tri_obj[t];
        base = t;
      }
      local[t] = t - base;
    }
  }
  make_device_tree(c->host_tree, c->leaf_verts, obj, local.empty() ? nullptr : local.data(), dh);
  if (!c->tri_normal.empty()) {  // caller-provided legality normals by triangle id
    for (uint32_t k = 0; k < ntri; k++) {
      uint32_t tid = c->host_tree.leaf_tris[k];
      memcpy(&dh.tn[4 * (size_t)k], &c->tri_normal[3 * (size_t)tid], 12);
    }
  }
  c->tree_loaded = false;  // the device tree is incomplete until every upload has succeeded
  if (int r = hip_err(c, c->tree.upload(dh, c->host_tree.max_depth, c->stream), "tree upload")) return r;
  c->tree_loaded = true;
  return RX_OK;
}

// All blocks of a traversal launch are resident at once: each block runs a
// fixed range of the batch, so a block that started late would finish late.
static int max_blocks(rx_ctx* c, int block) {
  (void)block;
  if (!c->isect_blocks_per_cu) c->isect_blocks_per_cu = intersect_blocks_per_cu();
  return c->n_cu * c->isect_blocks_per_cu;
}

}  // namespace rx

using namespace rx;

extern "C" {

int rx_tree_build(uint32_t n_verts, const float* verts, uint32_t n_tris, const uint32_t* tris, uint32_t flags,
                  rx_tree** out) {
  if (!out || (n_tris && (!verts || !tris))) return RX_ERR_INVALID;
  for (size_t i = 0; i < 3 * (size_t)n_tris; i++)
    if (tris[i] >= n_verts) return RX_ERR_INVALID;
  rx_tree* t = new (std::nothrow) rx_tree;
  if (!t) return RX_ERR_NO_MEMORY;
  try {
    t->verts.assign(verts, verts + 3 * (size_t)n_verts);
    t->tris.assign(tris, tris + 3 * (size_t)n_tris);
    build_tree(t->verts.data(), n_verts, t->tris.data(), n_tris, (flags & RX_TREE_ABS_INT) != 0, t->t);
  } catch (const std::bad_alloc&) {
    delete t;
    return RX_ERR_NO_MEMORY;
  }
  *out = t;
  return RX_OK;
}

void rx_tree_free(rx_tree* tree) { delete tree; }

int rx_tree_info(const rx_tree* tree, uint32_t* n_nodes, uint32_t* n_leaf_tris, uint32_t* max_depth) {
  if (!tree) return RX_ERR_INVALID;
  if (n_nodes) *n_nodes = (uint32_t)tree->t.nodes.size();
  if (n_leaf_tris) *n_leaf_tris = (uint32_t)tree->t.leaf_tris.size();
  if (max_depth) *max_depth = tree->t.max_depth;
  return RX_OK;
}

int rx_tree_serialize(const rx_tree* tree, float* nodes4, float* min4, float* max4, float* v1, float* v2,
                      float* v3) {
  if (!tree || !nodes4 || !min4 || !max4 || !v1 || !v2 || !v3) return RX_ERR_INVALID;
  const HostTree& t = tree->t;
  for (size_t i = 0; i < t.nodes.size(); i++) {
    const PNode& p = t.nodes[i];
    float* n = nodes4 + 4 * i;
    n[0] = (float)p.parent;
    if (p.right >= 0) {
      n[1] = (float)p.right;
      n[2] = -1;
      n[3] = -1;
    } else {
      n[1] = -1;
      n[2] = (float)p.leaf_begin;
      n[3] = (float)p.leaf_end;
    }
    for (int a = 0; a < 3; a++) {
      min4[4 * i + a] = p.lo[a];
      max4[4 * i + a] = p.hi[a];
    }
    min4[4 * i + 3] = 0;
    max4[4 * i + 3] = 0;
  }
  for (size_t k = 0; k < t.leaf_tris.size(); k++) {
    uint32_t tr = t.leaf_tris[k];
    float* o[3] = {v1 + 4 * k, v2 + 4 * k, v3 + 4 * k};
    for (int j = 0; j < 3; j++) {
      memcpy(o[j], &tree->verts[3 * (size_t)tree->tris[3 * (size_t)tr + j]], 12);
      o[j][3] = 0;
    }
    v1[4 * k + 3] = (float)tr;
  }
  return RX_OK;
}

int rx_create(int device, rx_ctx** out) {
  if (!out) return RX_ERR_INVALID;
  *out = nullptr;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return RX_ERR_NO_DEVICE;
  if (device < 0 || device >= n) return RX_ERR_INVALID;
  rx_ctx* c = new (std::nothrow) rx_ctx;
  if (!c) return RX_ERR_NO_MEMORY;
  c->device = device;
  hipDeviceProp_t prop;
  if (hipSetDevice(device) != hipSuccess || hipGetDeviceProperties(&prop, device) != hipSuccess ||
      hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev1) != hipSuccess) {
    rx_destroy(c);
    return RX_ERR_HIP;
  }
  if (strncmp(prop.gcnArchName, "gfx950", 6) != 0) {
    rx_destroy(c);
    return RX_ERR_NO_DEVICE;
  }
  c->n_cu = prop.multiProcessorCount;
  if (scratch_addr(device, &c->scratch) != RX_OK) {  // the no-allocation entry point's scratch on this device
    rx_destroy(c);
    return RX_ERR_HIP;
  }
  *out = c;
  return RX_OK;
}

void rx_render_state_free(rx_ctx* c);
void rx_guide_state_free(rx_ctx* c);

void rx_destroy(rx_ctx* c) {
  if (!c) return;
  (void)hipSetDevice(c->device);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  rx_render_state_free(c);
  rx_guide_state_free(c);
  if (c->ev0) (void)hipEventDestroy(c->ev0);
  if (c->ev1) (void)hipEventDestroy(c->ev1);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
}

const char* rx_last_error(const rx_ctx* c) { return c ? c->err.c_str() : "null context"; }

int rx_debug_fail_device_alloc(int32_t nth) {
  g_fail_alloc_countdown = nth < 0 ? -1 : nth;
  return RX_OK;
}

int rx_synchronize(rx_ctx* c) {
  if (!c) return RX_ERR_INVALID;
  if (int r = hip_err(c, hipStreamSynchronize(c->stream), "synchronize")) return r;
  // the traversal stack flags of the queued rx_intersect_dev / rx_intersect_dev_noalloc calls
  const void* flags[2] = {c->isect_queued ? c->overflow.p : nullptr, c->noalloc_queued ? c->scratch : nullptr};
  c->isect_queued = c->noalloc_queued = false;
  for (const void* f : flags) {
    if (!f) continue;
    int32_t ovf = 0;
    if (int r = hip_err(c, hipMemcpy(&ovf, f, 4, hipMemcpyDeviceToHost), "overflow flag")) return r;
    if (ovf) return set_err(c, RX_ERR_OVERFLOW, "rx_intersect_dev: traversal stack overflow");
  }
  return RX_OK;
}

int rx_load_tree(rx_ctx* c, uint32_t n_nodes, const float* nodes4, const float* min4, const float* max4,
                 uint32_t n_leaf_tris, const float* v1, const float* v2, const float* v3) {
  if (!c || (n_nodes && (!nodes4 || !min4 || !max4)) || (n_leaf_tris && (!v1 || !v2 || !v3)))
    return set_err(c, RX_ERR_INVALID, "rx_load_tree: null array");
  (void)hipSetDevice(c->device);
  c->scene_loaded = false;  // the previous scene and tree are gone from here on
  c->tree_loaded = false;
  try {
    if (!parse_serialized(n_nodes, nodes4, min4, max4, n_leaf_tris, v1, v2, v3, c->host_tree, c->leaf_verts))
      return set_err(c, RX_ERR_INVALID, "rx_load_tree: malformed serialized tree");
    c->tri_obj.clear();  // legality data belongs to the previous tree
    c->tri_normal.clear();
    return upload_tree(c);
  } catch (const std::bad_alloc&) {
    return set_err(c, RX_ERR_NO_MEMORY, "rx_load_tree: host allocation failed");
  }
}

int rx_load_legality(rx_ctx* c, uint32_t n_tris, const int32_t* tri_obj, const float* tri_normal) {
  if (!c || !tri_obj) return set_err(c, RX_ERR_INVALID, "rx_load_legality: null array");
  try {
    c->tri_obj.assign(tri_obj, tri_obj + n_tris);
    if (tri_normal)
      c->tri_normal.assign(tri_normal, tri_normal + 3 * (size_t)n_tris);
    else
      c->tri_normal.clear();
    if (c->tree_loaded) {
      c->scene_loaded = false;  // the render scene's device tree is replaced
      return upload_tree(c);
    }
  } catch (const std::bad_alloc&) {
    c->tree_loaded = c->scene_loaded = false;
    return set_err(c, RX_ERR_NO_MEMORY, "rx_load_legality: host allocation failed");
  }
  return RX_OK;
}

int rx_intersect_dev(rx_ctx* c, uint32_t n, const float* org, const float* dir, const int32_t* inside,
                     const int32_t* contact, const int32_t* skip_tid, uint32_t flags, float* dist, int32_t* tid,
                     float* bary, uint32_t* counts) {
  if (!c) return RX_ERR_INVALID;
  if (!c->tree_loaded) return set_err(c, RX_ERR_STATE, "rx_intersect: no tree loaded");
  if (n == 0) return RX_OK;
  if (!org || !dir || !dist || !tid) return set_err(c, RX_ERR_INVALID, "rx_intersect: null array");
  if ((flags & RX_ISECT_LEGALITY) && (!inside || !contact))
    return set_err(c, RX_ERR_INVALID, "rx_intersect: legality needs inside/contact");
  (void)hipSetDevice(c->device);
  const int block = kTraceBlock;
  int blocks = (int)std::min<uint64_t>((n + block - 1) / block, (uint64_t)max_blocks(c, block));
  int sd = 0;
  if (int r = ensure_spill(c, blocks, block, (int)c->tree.max_depth, &sd)) return r;
  IsectArgs a;
  a.n = n;
  a.org = org;
  a.dir = dir;
  a.inside = inside;
  a.contact = contact;
  a.skip = skip_tid;
  a.dist = dist;
  a.tid = tid;
  a.bary = bary;
  a.counts = counts;
  a.spill = c->spill.as<uint2>();
  a.spill_depth = sd;
  a.overflow = c->overflow.as<int32_t>();
  a.work = c->overflow.as<uint32_t>() + 1;
  if (int r = hip_err(c, hipMemsetAsync(a.overflow, 0, 4 * (1 + kWorkWords), c->stream), "overflow clear")) return r;
  if (int r = hip_err(c, hipEventRecord(c->ev0, c->stream), "event")) return r;
  if (int r = hip_err(c, launch_intersect(c->tree.dev, a, flags, blocks, c->stream), "intersect launch")) return r;
  if (int r = hip_err(c, hipEventRecord(c->ev1, c->stream), "event")) return r;
  c->isect_queued = true;
  return RX_OK;
}

int rx_intersect_dev_noalloc(rx_ctx* c, uint32_t n, const float* org, const float* dir, const int32_t* inside,
                             const int32_t* contact, const int32_t* skip_tid, uint32_t flags, float* dist,
                             int32_t* tid, float* bary, uint32_t* counts) {
  if (!c) return RX_ERR_INVALID;
  if (!c->tree_loaded) return set_err(c, RX_ERR_STATE, "rx_intersect_dev_noalloc: no tree loaded");
  if (!c->scratch) return set_err(c, RX_ERR_STATE, "rx_intersect_dev_noalloc: no scratch on this device");
  if (n == 0) return RX_OK;
  if (!org || !dir || !dist || !tid) return set_err(c, RX_ERR_INVALID, "rx_intersect_dev_noalloc: null array");
  if ((flags & RX_ISECT_LEGALITY) && (!inside || !contact))
    return set_err(c, RX_ERR_INVALID, "rx_intersect_dev_noalloc: legality needs inside/contact");
  (void)hipSetDevice(c->device);
  const int block = kTraceBlock;
  // the spill each thread needs beyond the LDS stack, and how many threads' worth the scratch holds
  const int sd = std::max(1, (int)c->tree.max_depth + 2 - kLdsDepth);
  const size_t per_block = (size_t)block * sd * sizeof(uint2);
  const size_t fit = (kScratchBytes - kScratchHeader) / per_block;
  if (fit < 1)
    return set_err(c, RX_ERR_INVALID,
                   "rx_intersect_dev_noalloc: the tree is too deep for the fixed traversal scratch (" +
                       std::to_string(per_block) + " bytes per block needed, " +
                       std::to_string(kScratchBytes - kScratchHeader) + " available)");
  const int blocks = (int)std::min<uint64_t>(std::min<uint64_t>((n + block - 1) / block, (uint64_t)max_blocks(c, block)),
                                             (uint64_t)fit);
  unsigned char* base = static_cast<unsigned char*>(c->scratch);
  IsectArgs a;
  a.n = n;
  a.org = org;
  a.dir = dir;
  a.inside = inside;
  a.contact = contact;
  a.skip = skip_tid;
  a.dist = dist;
  a.tid = tid;
  a.bary = bary;
  a.counts = counts;
  a.overflow = reinterpret_cast<int32_t*>(base);
  a.work = reinterpret_cast<uint32_t*>(base) + 1;
  a.spill = reinterpret_cast<uint2*>(base + kScratchHeader);
  a.spill_depth = sd;
  // the previous call's counters are still there: clear them with a kernel (no memset node under capture)
  hipLaunchKernelGGL(k_zero_words, dim3(1), dim3(256), 0, c->stream, reinterpret_cast<uint32_t*>(base),
                     (uint32_t)(1 + kWorkWords));
  if (int r = hip_err(c, hipGetLastError(), "scratch clear")) return r;
  if (int r = hip_err(c, launch_intersect(c->tree.dev, a, flags, blocks, c->stream), "intersect launch")) return r;
  c->noalloc_queued = true;
  return RX_OK;
}

int rx_last_kernel_ms(rx_ctx* c, float* ms) {
  if (!c || !ms) return RX_ERR_INVALID;
  if (int r = hip_err(c, hipEventSynchronize(c->ev1), "event sync")) return r;
  return hip_err(c, hipEventElapsedTime(ms, c->ev0, c->ev1), "elapsed");
}

int rx_intersect(rx_ctx* c, uint32_t n, const float* org, const float* dir, const int32_t* inside,
                 const int32_t* contact, const int32_t* skip_tid, uint32_t flags, float* dist, int32_t* tid,
                 float* bary, uint32_t* counts) {
  if (!c) return RX_ERR_INVALID;
  if (!c->tree_loaded) return set_err(c, RX_ERR_STATE, "rx_intersect: no tree loaded");
  if (n == 0) return RX_OK;
  if (!org || !dir || !dist || !tid) return set_err(c, RX_ERR_INVALID, "rx_intersect: null array");
  const bool leg = (flags & RX_ISECT_LEGALITY) != 0;
  if (leg && (!inside || !contact)) return set_err(c, RX_ERR_INVALID, "rx_intersect: legality needs inside/contact");
  (void)hipSetDevice(c->device);
  hipStream_t s = c->stream;
#define RXCK(x, w) \
  if (int r_ = hip_err(c, (x), w)) return r_;
  RXCK(c->io_org.upload(org, 12 * (size_t)n, s), "upload org");
  RXCK(c->io_dir.upload(dir, 12 * (size_t)n, s), "upload dir");
  if (leg) {
    RXCK(c->io_in.upload(inside, 4 * (size_t)n, s), "upload inside");
    RXCK(c->io_con.upload(contact, 4 * (size_t)n, s), "upload contact");
  }
  if (skip_tid) RXCK(c->io_skip.upload(skip_tid, 4 * (size_t)n, s), "upload skip");
  RXCK(c->io_dist.ensure(4 * (size_t)n), "alloc dist");
  RXCK(c->io_tid.ensure(4 * (size_t)n), "alloc tid");
  if (bary) RXCK(c->io_bary.ensure(8 * (size_t)n), "alloc bary");
  if (counts) RXCK(c->io_counts.ensure(8 * (size_t)n), "alloc counts");
  int r = rx_intersect_dev(c, n, c->io_org.as<float>(), c->io_dir.as<float>(), leg ? c->io_in.as<int32_t>() : nullptr,
                           leg ? c->io_con.as<int32_t>() : nullptr, skip_tid ? c->io_skip.as<int32_t>() : nullptr,
                           flags, c->io_dist.as<float>(),
                           c->io_tid.as<int32_t>(), bary ? c->io_bary.as<float>() : nullptr,
                           counts ? c->io_counts.as<uint32_t>() : nullptr);
  if (r) return r;
  int32_t ovf = 0;
  RXCK(hipMemcpyAsync(dist, c->io_dist.p, 4 * (size_t)n, hipMemcpyDeviceToHost, s), "download dist");
  RXCK(hipMemcpyAsync(tid, c->io_tid.p, 4 * (size_t)n, hipMemcpyDeviceToHost, s), "download tid");
  if (bary) RXCK(hipMemcpyAsync(bary, c->io_bary.p, 8 * (size_t)n, hipMemcpyDeviceToHost, s), "download bary");
  if (counts) RXCK(hipMemcpyAsync(counts, c->io_counts.p, 8 * (size_t)n, hipMemcpyDeviceToHost, s), "download counts");
  RXCK(hipMemcpyAsync(&ovf, c->overflow.p, 4, hipMemcpyDeviceToHost, s), "download overflow");
  RXCK(hipStreamSynchronize(s), "synchronize");
  c->isect_queued = false;  // its flag was read here
#undef RXCK
  if (ovf) return set_err(c, RX_ERR_OVERFLOW, "rx_intersect: traversal stack overflow");
  return RX_OK;
}

}  // extern "C"

namespace rx {
// Streaming copy, 16 B per lane per iteration: reads and writes each byte once.
__global__ void k_copy16(const uint4* __restrict__ src, uint4* __restrict__ dst, size_t n16) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n16; i += (size_t)gridDim.x * blockDim.x)
    dst[i] = src[i];
}
}  // namespace rx

extern "C" int rx_copy_kernel(rx_ctx* c, void* dst, const void* src, uint64_t bytes) {
  if (!c || !dst || !src) return set_err(c, RX_ERR_INVALID, "rx_copy_kernel: null pointer");
  if ((bytes & 15) || ((uintptr_t)dst & 15) || ((uintptr_t)src & 15))
    return set_err(c, RX_ERR_INVALID, "rx_copy_kernel: pointers and size must be multiples of 16 bytes");
  if (bytes == 0) return RX_OK;
  (void)hipSetDevice(c->device);
  const size_t n16 = bytes / 16;
  const int blocks = (int)std::min<size_t>((n16 + 255) / 256, (size_t)c->n_cu * 32);
  hipLaunchKernelGGL(k_copy16, dim3(blocks), dim3(256), 0, c->stream, (const uint4*)src, (uint4*)dst, n16);
  if (int r = hip_err(c, hipGetLastError(), "copy launch")) return r;
  return hip_err(c, hipStreamSynchronize(c->stream), "copy synchronize");
}
