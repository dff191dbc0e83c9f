// Multi-scene batch renderer: many scenes resident together, rendered as one
// wavefront job.  No counterpart in the reference (which renders one scene per
// process, src/Renderer.cpp:50); each scene's image is bit-identical to
// rx_render of that scene alone.
//
// The scenes' device trees, objects and emitters are concatenated with their
// node / leaf / triangle / object indices rebased, so every kernel sees one
// DevScene; a path carries its scene index, and the scene's root, emitter
// range, camera, seed and depth limits come from a per-scene table.  A host
// schedule packs (scene, pixel range, sample) segments into passes of at most
// `path_slots` paths with at most one sample of any pixel per pass, in sample
// order, so the FP64 film adds every pixel's samples in the same order as a
// single-scene render.  The schedule lives on the device and a device counter
// selects the pass, so one pass is captured into a hipGraph and replayed.
#include <hip/hip_cooperative_groups.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <new>
#include <vector>

#include "rx_pathtrace.h"

namespace rx {

struct BatchItem {
  rx_camera_desc cam;
  rx_render_params prm;
  uint32_t npix;
  SceneInfo info;
};

// The whole batch as one launch: the grid stays resident (cooperative
// launch) and runs every pass's stages in order, generate, then per bounce
// extend, shade and (NEE) shadow, then the film, with a grid-wide barrier
// where the stream launches have a kernel boundary.  The stages are the
// stream kernels' own bodies; one LDS buffer serves them in turn.
template <int MODE>
union PersistLds {
  uint2 stack[kLdsDepth * kRBlock];
  ShadeLds<MODE> shade;
};
constexpr int kPersistMinWaves = 4;  // resident waves per SIMD the registers must allow
template <int MODE>
__global__ void __launch_bounds__(kRBlock, kPersistMinWaves)
    k_batch_persistent(DevScene S, PathBufs B, PassArgs P, BatchArgs BA, uint32_t n_passes, uint32_t Dmax) {
  __shared__ PersistLds<MODE> sh;
  cooperative_groups::grid_group grid = cooperative_groups::this_grid();
  const GuideArgs GA{};
  const DevCam C0{};
  for (uint32_t k = 0; k < n_passes; k++) {  // every block runs the same barriers
    generate_body<true, false>(S, C0, B, P, BA, sh.stack);
    grid.sync();
    for (uint32_t bb = 0; bb <= Dmax; bb++) {
      if (bb == 0) extend_body<false, false, true, kLdsDepth>(S.tree, B, bb, BA, sh.stack);
      else extend_body<true, false, true, kLdsDepth>(S.tree, B, bb, BA, sh.stack);
      grid.sync();
      shade_body<MODE, true, false>(S, B, P, bb, GA, BA, sh.shade);
      grid.sync();
      if (MODE == RX_MODE_NEE) {
        shadow_body<true, false>(S, B, P, bb, GA, BA, sh.stack);
        grid.sync();
      }
    }
    accumulate_batch_body(B, BA);
    grid.sync();
    if (blockIdx.x == 0 && threadIdx.x == 0) *BA.pass_id += 1;
    grid.sync();
  }
}

}  // namespace rx

struct rx_batch {
  rx_ctx* c = nullptr;
  std::vector<rx::BatchItem> items;
  int mode = -1;
  // concatenated scene data (host copies until uploaded)
  std::vector<float> nodes, tv0, tv1, tv2, tn, etri, enrm, ecdf;
  std::vector<int32_t> leaf_start;
  std::vector<rx::DevObject> objs;
  uint32_t max_tree_depth = 0;
  uint64_t total_pix = 0;
  bool dirty = true;
  rx::DBuf d_nodes, d_leaf, d_tv0, d_tv1, d_tv2, d_tn, d_objs, d_etri, d_enrm, d_ecdf, d_info;
  rx::DBuf segs, passes, pass_id, scn, pix, smp;
  rx::DBuf o, d, beta, lastp, L, s, hit, q0, q1, sq, cnt, so, sd, sc, ss, sum, sumsq, stats;
  uint32_t cap = 0;
  rx::DevScene dev{};
};

using namespace rx;

extern "C" {

int rx_batch_create(rx_ctx* c, rx_batch** out) {
  if (!c || !out) return RX_ERR_INVALID;
  rx_batch* b = new (std::nothrow) rx_batch;
  if (!b) return RX_ERR_NO_MEMORY;
  b->c = c;
  *out = b;
  return RX_OK;
}

void rx_batch_destroy(rx_batch* b) {
  if (!b) return;
  if (b->c) {
    (void)hipSetDevice(b->c->device);
    (void)hipStreamSynchronize(b->c->stream);
  }
  delete b;
}

int rx_batch_add_scene(rx_batch* b, uint32_t n_verts, const float* verts, uint32_t n_tris, const uint32_t* tris,
                       uint32_t n_objects, const rx_object_desc* objects, uint32_t tree_flags,
                       const rx_camera_desc* cam, const rx_render_params* prm, uint32_t* index) {
  if (!b) return RX_ERR_INVALID;
  rx_ctx* c = b->c;
  if (!verts || !tris || !objects || !cam || !prm || n_tris == 0 || n_objects == 0)
    return set_err(c, RX_ERR_INVALID, "rx_batch_add_scene: empty scene or null argument");
  if (cam->width <= 0 || cam->height <= 0 || (uint64_t)cam->width * cam->height > (1ull << 26))
    return set_err(c, RX_ERR_INVALID, "rx_batch_add_scene: bad resolution");
  if (prm->max_depth >= 64) return set_err(c, RX_ERR_INVALID, "rx_batch_add_scene: max_depth must be < 64");
  if (prm->mode > RX_MODE_NEE) return set_err(c, RX_ERR_INVALID, "rx_batch_add_scene: batches render PT or NEE");
  if (b->mode >= 0 && (int)prm->mode != b->mode)
    return set_err(c, RX_ERR_INVALID, "rx_batch_add_scene: every scene of a batch uses the same mode");
  for (size_t i = 0; i < 3 * (size_t)n_tris; i++)
    if (tris[i] >= n_verts) return set_err(c, RX_ERR_INVALID, "rx_batch_add_scene: vertex index out of range");
  const int32_t obj_base = (int32_t)b->objs.size();
  std::vector<int32_t> tri_obj(n_tris, -1);
  std::vector<uint32_t> local(n_tris, 0);
  for (uint32_t o = 0; o < n_objects; o++) {
    const rx_object_desc& od = objects[o];
    if (od.type < 0 || od.type > RX_OBJ_VPM) return set_err(c, RX_ERR_INVALID, "rx_batch_add_scene: bad object type");
    if (od.type == RX_OBJ_VPM)
      return set_err(c, RX_ERR_INVALID, "rx_batch_add_scene: media need rx_load_scene (per-object trees)");
    if ((uint64_t)od.tri_begin + od.tri_count > n_tris)
      return set_err(c, RX_ERR_INVALID, "rx_batch_add_scene: object range");
    for (uint32_t t = 0; t < od.tri_count; t++) {
      if (tri_obj[od.tri_begin + t] != -1) return set_err(c, RX_ERR_INVALID, "rx_batch_add_scene: overlapping objects");
      tri_obj[od.tri_begin + t] = obj_base + (int32_t)o;
      local[od.tri_begin + t] = t;
    }
  }
  for (uint32_t t = 0; t < n_tris; t++)
    if (tri_obj[t] < 0) return set_err(c, RX_ERR_INVALID, "rx_batch_add_scene: triangle without object");
  std::vector<float> etri, enrm, ecdf;
  const double area = collect_emitters(verts, tris, n_objects, objects, obj_base, etri, enrm, ecdf);
  if (prm->mode == RX_MODE_NEE && ecdf.empty())
    return set_err(c, RX_ERR_INVALID, "rx_batch_add_scene: next-event estimation needs an emitter");
  HostTree ht;
  DeviceTreeHost dh;
  try {
    build_tree(verts, n_verts, tris, n_tris, (tree_flags & RX_TREE_ABS_INT) != 0, ht);
    std::vector<float> lv(9 * ht.leaf_tris.size());
    for (size_t k = 0; k < ht.leaf_tris.size(); k++)
      for (int j = 0; j < 3; j++)
        memcpy(&lv[9 * k + 3 * j], &verts[3 * (size_t)tris[3 * (size_t)ht.leaf_tris[k] + j]], 12);
    make_device_tree(ht, lv, tri_obj.data(), local.data(), dh);
  } catch (const std::bad_alloc&) {
    return set_err(c, RX_ERR_NO_MEMORY, "rx_batch_add_scene: host memory");
  }
  // append the tree with node, leaf and triangle indices rebased
  const int32_t node_base = (int32_t)(b->nodes.size() / 16);
  const int32_t leaf_base = (int32_t)b->leaf_start.size();
  const int32_t tri_base = (int32_t)(b->tv0.size() / 4);
  auto rebase = [&](int32_t ref) { return ref >= 0 ? ref + node_base : ~(~ref + leaf_base); };
  for (size_t i = 0; i < dh.nodes.size() / 16; i++) {
    float f[16];
    memcpy(f, &dh.nodes[16 * i], 64);
    int32_t* r = reinterpret_cast<int32_t*>(f + 12);
    r[0] = rebase(r[0]);
    r[1] = rebase(r[1]);
    b->nodes.insert(b->nodes.end(), f, f + 16);
  }
  for (uint32_t l = 0; l < dh.n_leaves; l++) b->leaf_start.push_back(dh.leaf_start[l] + tri_base);
  b->tv0.insert(b->tv0.end(), dh.tv0.begin(), dh.tv0.end());
  b->tv1.insert(b->tv1.end(), dh.tv1.begin(), dh.tv1.end());
  b->tv2.insert(b->tv2.end(), dh.tv2.begin(), dh.tv2.end());
  b->tn.insert(b->tn.end(), dh.tn.begin(), dh.tn.end());
  for (uint32_t o = 0; o < n_objects; o++) b->objs.push_back(make_devobject(objects[o]));
  BatchItem it;
  it.cam = *cam;
  it.prm = *prm;
  it.npix = (uint32_t)(cam->width * cam->height);
  SceneInfo& I = it.info;
  memset(&I, 0, sizeof(I));
  I.root_ref = rebase(dh.root_ref);
  memcpy(I.root_lo, dh.root_lo, 12);
  memcpy(I.root_hi, dh.root_hi, 12);
  I.e_base = (int32_t)(b->ecdf.size());
  I.n_etri = (int32_t)ecdf.size();
  I.total_area = (float)area;
  I.key1 = seed_key(prm->seed);
  I.max_depth = prm->max_depth;
  I.rr_start = prm->rr_start_depth;
  I.film_base = (uint32_t)b->total_pix;
  I.cam = make_devcam(cam);
  b->etri.insert(b->etri.end(), etri.begin(), etri.end());
  b->enrm.insert(b->enrm.end(), enrm.begin(), enrm.end());
  b->ecdf.insert(b->ecdf.end(), ecdf.begin(), ecdf.end());
  b->max_tree_depth = std::max(b->max_tree_depth, ht.max_depth);
  b->total_pix += it.npix;
  if (b->total_pix > (1ull << 31)) return set_err(c, RX_ERR_INVALID, "rx_batch_add_scene: too many pixels");
  b->mode = (int)prm->mode;
  if (index) *index = (uint32_t)b->items.size();
  b->items.push_back(it);
  b->dirty = true;
  return RX_OK;
}

namespace {
// rx_batch_render; max_passes > 0 renders only the schedule's first passes
// (the autotuner's timing runs).
int batch_render(rx_batch* b, rx_render_stats* stats, uint32_t max_passes) {
  rx_ctx* c = b->c;
  if (b->items.empty()) return set_err(c, RX_ERR_STATE, "rx_batch_render: no scenes");
  (void)hipSetDevice(c->device);
  hipStream_t s = c->stream;
#define RXCK(x, w) \
  if (int r_ = hip_err(c, (x), w)) return r_;
  if (b->dirty) {  // upload the concatenated scenes
    std::vector<int32_t> ls = b->leaf_start;
    ls.push_back((int32_t)(b->tv0.size() / 4));
    std::vector<SceneInfo> info(b->items.size());
    for (size_t k = 0; k < b->items.size(); k++) info[k] = b->items[k].info;
    RXCK(b->d_nodes.upload(b->nodes.data(), b->nodes.size() * 4, s), "upload nodes");
    RXCK(b->d_leaf.upload(ls.data(), ls.size() * 4, s), "upload leaves");
    RXCK(b->d_tv0.upload(b->tv0.data(), b->tv0.size() * 4, s), "upload triangles");
    RXCK(b->d_tv1.upload(b->tv1.data(), b->tv1.size() * 4, s), "upload triangles");
    RXCK(b->d_tv2.upload(b->tv2.data(), b->tv2.size() * 4, s), "upload triangles");
    RXCK(b->d_tn.upload(b->tn.data(), b->tn.size() * 4, s), "upload normals");
    RXCK(b->d_objs.upload(b->objs.data(), b->objs.size() * sizeof(DevObject), s), "upload objects");
    RXCK(b->d_etri.upload(b->etri.data(), b->etri.size() * 4, s), "upload emitters");
    RXCK(b->d_enrm.upload(b->enrm.data(), b->enrm.size() * 4, s), "upload emitters");
    RXCK(b->d_ecdf.upload(b->ecdf.data(), b->ecdf.size() * 4, s), "upload emitters");
    RXCK(b->d_info.upload(info.data(), info.size() * sizeof(SceneInfo), s), "upload scene table");
    DevScene& D = b->dev;
    memset(&D, 0, sizeof(D));
    D.tree.nodes = b->d_nodes.as<float4>();
    D.tree.leaf_start = b->d_leaf.as<int32_t>();
    D.tree.tv0 = b->d_tv0.as<float4>();
    D.tree.tv1 = b->d_tv1.as<float4>();
    D.tree.tv2 = b->d_tv2.as<float4>();
    D.tree.tn = b->d_tn.as<float4>();
    D.tree.n_leaf_tris = (uint32_t)(b->tv0.size() / 4);
    D.objs = b->d_objs.as<DevObject>();
    D.n_objs = (int32_t)b->objs.size();
    D.etri = b->d_etri.as<float4>();
    D.enrm = b->d_enrm.as<float4>();
    D.ecdf = b->d_ecdf.as<float>();
    b->dirty = false;
  }
  // schedule: sample rounds over the scenes, packed into passes of <= N slots
  // holding at most one segment of each scene
  const uint32_t slots = c->tuning.path_slots ? c->tuning.path_slots : kDefaultPathSlots;
  const uint32_t N = (uint32_t)std::min<uint64_t>(slots, b->total_pix);
  uint32_t max_spp = 0, Dmax = 0;
  for (auto& it : b->items) {
    max_spp = std::max(max_spp, it.prm.spp);
    Dmax = std::max(Dmax, it.prm.max_depth);
  }
  std::vector<uint4> segs, passes;
  {
    std::vector<uint32_t> in_pass(b->items.size(), UINT32_MAX);
    uint32_t used = 0, seg0 = 0;
    auto close = [&]() {
      if (used) passes.push_back(make_uint4(seg0, (uint32_t)segs.size() - seg0, used, 0));
      seg0 = (uint32_t)segs.size();
      used = 0;
    };
    for (uint32_t smp = 0; smp < max_spp; smp++) {
      for (uint32_t k = 0; k < (uint32_t)b->items.size(); k++) {
        const BatchItem& it = b->items[k];
        if (smp >= it.prm.spp) continue;
        uint32_t first = 0;
        while (first < it.npix) {
          if (used == N || in_pass[k] == (uint32_t)passes.size()) close();
          const uint32_t take = std::min(it.npix - first, N - used);
          segs.push_back(make_uint4(k, first, it.prm.sample_offset + smp, used));
          in_pass[k] = (uint32_t)passes.size();
          used += take;
          first += take;
        }
      }
    }
    close();
  }
  RXCK(b->segs.upload(segs.data(), segs.size() * sizeof(uint4), s), "upload schedule");
  RXCK(b->passes.upload(passes.data(), passes.size() * sizeof(uint4), s), "upload schedule");
  RXCK(b->pass_id.ensure(64), "alloc");
  if (b->cap < N) {
    for (DBuf* x : {&b->o, &b->d, &b->beta, &b->lastp, &b->L, &b->s, &b->so, &b->sd, &b->sc})
      RXCK(x->ensure(16 * (size_t)N), "alloc paths");
    for (DBuf* x : {&b->hit, &b->ss}) RXCK(x->ensure(8 * (size_t)N), "alloc paths");
    for (DBuf* x : {&b->q0, &b->q1, &b->sq, &b->scn, &b->pix, &b->smp}) RXCK(x->ensure(4 * (size_t)N), "alloc paths");
    b->cap = N;
  }
  RXCK(b->sum.ensure(24 * (size_t)b->total_pix), "alloc film");
  RXCK(b->sumsq.ensure(24 * (size_t)b->total_pix), "alloc film");
  RXCK(b->cnt.ensure(4 * kCntWords), "alloc");
  RXCK(b->stats.ensure(64), "alloc");
  if (!c->render_blocks_per_cu) {
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_extend<true, false, true>, kRBlock, 0) != hipSuccess ||
        nb < 1)
      nb = 4;
    c->render_blocks_per_cu = nb;
  }
  const int per_cu = c->tuning.blocks_per_cu ? std::min<int>((int)c->tuning.blocks_per_cu, c->render_blocks_per_cu)
                                             : c->render_blocks_per_cu;
  const int blocks = (int)std::min<uint64_t>((N + kRBlock - 1) / kRBlock, (uint64_t)c->n_cu * per_cu);
  // shade and shadow take any grid; theirs is what they hold resident (see rx_render)
  auto resident = [&](const void* k) {
    int nr = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nr, k, kRBlock, 0) != hipSuccess || nr < 1) nr = per_cu;
    return (int)std::min<uint64_t>((N + kRBlock - 1) / kRBlock, (uint64_t)c->n_cu * std::min(nr, per_cu));
  };
  const int sblocks = resident(b->mode == RX_MODE_NEE ? (const void*)k_shade<1, true, false>
                                                     : (const void*)k_shade<0, true, false>);
  const int shblocks = resident((const void*)k_shadow<true, false>);
  int sd = 0;
  if (int r = ensure_spill(c, blocks, kRBlock, (int)b->max_tree_depth, &sd)) return r;
  PathBufs B;
  memset(&B, 0, sizeof(B));
  B.o = b->o.as<float4>();
  B.d = b->d.as<float4>();
  B.beta = b->beta.as<float4>();
  B.lastp = b->lastp.as<float4>();
  B.L = b->L.as<float4>();
  B.s = b->s.as<int4>();
  B.hit = b->hit.as<int2>();
  B.q0 = b->q0.as<uint32_t>();
  B.q1 = b->q1.as<uint32_t>();
  B.sq = b->sq.as<uint32_t>();
  B.cnt = b->cnt.as<uint32_t>();
  B.so = b->so.as<float4>();
  B.sd = b->sd.as<float4>();
  B.sc = b->sc.as<float4>();
  B.ss = b->ss.as<int2>();
  B.sum = b->sum.as<double>();
  B.sumsq = b->sumsq.as<double>();
  B.stats = b->stats.as<unsigned long long>();
  B.spill = c->spill.as<uint2>();
  B.spill_depth = sd;
  B.overflow = c->overflow.as<int32_t>();
  BatchArgs BA;
  BA.info = b->d_info.as<SceneInfo>();
  BA.segs = b->segs.as<uint4>();
  BA.passes = b->passes.as<uint4>();
  BA.pass_id = b->pass_id.as<uint32_t>();
  BA.scn = b->scn.as<uint32_t>();
  BA.pix = b->pix.as<uint32_t>();
  BA.smp = b->smp.as<uint32_t>();
  PassArgs P;
  memset(&P, 0, sizeof(P));
  P.mode = (uint32_t)b->mode;
  GuideArgs GA{};
  const DevScene S = b->dev;
  const DevCam C0{};
  hipLaunchKernelGGL(k_fill, dim3(1024), dim3(256), 0, s, B.sum, 3 * (size_t)b->total_pix);
  hipLaunchKernelGGL(k_fill, dim3(1024), dim3(256), 0, s, B.sumsq, 3 * (size_t)b->total_pix);
  hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(64), 0, s, B.stats, 8, B.overflow);
  hipLaunchKernelGGL(k_pass_set, dim3(1), dim3(64), 0, s, BA.pass_id, 0u);
  RXCK(hipGetLastError(), "launch fill");
  const uint32_t sdep = c->tuning.stack_depth, sblk = c->tuning.block_size;
  // camera rays one per lane unless the tuning asks for packets: a batch pass holds one sample per
  // pixel, so a wave's 64 rays are 64 neighbouring pixels, whose walks share too little
  const bool packets = c->tuning.camera_rays == 2 && b->max_tree_depth < (uint32_t)kPacketDepth;
  auto launch_pass = [&]() {
    hipLaunchKernelGGL((k_generate<true, false>), dim3(blocks), dim3(kRBlock), 0, s, S, C0, B, P, BA);
    for (uint32_t bb = 0; bb <= Dmax; bb++) {
      if (bb == 0 && packets)
        hipLaunchKernelGGL(k_extend_packet<true>, dim3(blocks), dim3(kRBlock), 0, s, S.tree, B, bb, BA);
      else if (bb == 0) launch_extend<false, false, true>(sdep, sblk, blocks, s, S.tree, B, bb, BA);
      else launch_extend<true, false, true>(sdep, sblk, blocks, s, S.tree, B, bb, BA);
      if (b->mode == RX_MODE_NEE) {
        hipLaunchKernelGGL((k_shade<1, true, false>), dim3(sblocks), dim3(kRBlock), 0, s, S, B, P, bb, GA, BA);
        hipLaunchKernelGGL((k_shadow<true, false>), dim3(shblocks), dim3(kRBlock), 0, s, S, B, P, bb, GA, BA);
      } else {
        hipLaunchKernelGGL((k_shade<0, true, false>), dim3(sblocks), dim3(kRBlock), 0, s, S, B, P, bb, GA, BA);
      }
    }
    hipLaunchKernelGGL(k_accumulate_batch, dim3(blocks), dim3(kRBlock), 0, s, B, BA);
    hipLaunchKernelGGL(k_pass_next, dim3(1), dim3(64), 0, s, BA.pass_id);
  };
  const size_t n_passes = max_passes ? std::min<size_t>(max_passes, passes.size()) : passes.size();
  RXCK(hipEventRecord(c->ev0, s), "event");
  if (c->tuning.use_graph == 2) {  // one persistent cooperative launch for every pass
    void* fn = b->mode == RX_MODE_NEE ? (void*)k_batch_persistent<RX_MODE_NEE> : (void*)k_batch_persistent<RX_MODE_PT>;
    int nb = 0;
    RXCK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, fn, kRBlock, 0), "persistent occupancy");
    if (nb < 1) return set_err(c, RX_ERR_INVALID, "rx_batch_render: the persistent kernel cannot be resident");
    // the grid must be resident at once; it also must not exceed the spill buffer's threads
    const int pgrid = std::min(blocks, c->n_cu * std::min(nb, per_cu));
    uint32_t np = (uint32_t)n_passes, dm = Dmax;
    void* args[] = {(void*)&S, (void*)&B, (void*)&P, (void*)&BA, (void*)&np, (void*)&dm};
    RXCK(hipLaunchCooperativeKernel(fn, dim3(pgrid), dim3(kRBlock), args, 0, s), "persistent launch");
  } else if (c->tuning.use_graph) {  // record one pass, replay it for every pass
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    RXCK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal), "begin capture");
    launch_pass();
    hipError_t e = hipStreamEndCapture(s, &graph);
    if (e == hipSuccess) e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    for (size_t k = 0; e == hipSuccess && k < n_passes; k++) e = hipGraphLaunch(exec, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    RXCK(e, "pass graph");
  } else {
    for (size_t k = 0; k < n_passes; k++) launch_pass();
    RXCK(hipGetLastError(), "launch pass");
  }
  RXCK(hipEventRecord(c->ev1, s), "event");
  unsigned long long st[8] = {0};
  int32_t ovf = 0;
  RXCK(hipMemcpyAsync(st, B.stats, 64, hipMemcpyDeviceToHost, s), "download stats");
  RXCK(hipMemcpyAsync(&ovf, B.overflow, 4, hipMemcpyDeviceToHost, s), "download overflow");
  RXCK(hipStreamSynchronize(s), "batch synchronize");
  float ms = 0;
  RXCK(hipEventElapsedTime(&ms, c->ev0, c->ev1), "elapsed");
  if (stats) {
    memset(stats, 0, sizeof(*stats));
    stats->rays = st[0];
    stats->extend_rays = st[1];
    stats->shadow_rays = st[2];
    stats->segments = st[3];
    uint64_t samples = 0;  // camera paths of the passes rendered (all passes: every scene's pixels x spp)
    for (size_t k = 0; k < n_passes; k++) samples += passes[k].z;
    stats->samples = samples;
    stats->ms_total = ms;
    stats->extend_launches = (uint64_t)n_passes * (Dmax + 1);
  }
#undef RXCK
  if (ovf) return set_err(c, RX_ERR_OVERFLOW, "rx_batch_render: traversal stack overflow");
  return RX_OK;
}
}  // namespace

int rx_batch_render(rx_batch* b, rx_render_stats* stats) {
  if (!b) return RX_ERR_INVALID;
  return batch_render(b, stats, 0);
}

int rx_batch_autotune(rx_batch* b, uint32_t passes, rx_render_tuning* best, float* best_ms) {
  if (!b) return RX_ERR_INVALID;
  rx_ctx* c = b->c;
  if (b->items.empty()) return set_err(c, RX_ERR_STATE, "rx_batch_autotune: no scenes");
  const rx_render_tuning keep = c->tuning;
  rx_render_tuning win = keep;
  float win_ms = 0.0f;
  bool have = false;
  auto trial = [&](const rx_render_tuning& t) -> int {
    c->tuning = t;
    rx_render_stats st;
    int r = batch_render(b, &st, passes);
    if (r == RX_OK) r = batch_render(b, &st, passes);  // the second run is timed (first-use costs excluded)
    if (r) {
      c->tuning = keep;
      return r;
    }
    // per pass: the candidates differ in path slots, so in their pass counts
    const float per = (float)(st.ms_total / std::max<double>(1.0, (double)st.samples));
    if (!have || per < win_ms) {
      win = t;
      win_ms = per;
      have = true;
    }
    return RX_OK;
  };
  if (!c->render_blocks_per_cu) {  // occupancy limit, set by a first render
    rx_render_tuning t = keep;
    c->tuning = t;
    rx_render_stats st;
    if (int r = batch_render(b, &st, 1)) {
      c->tuning = keep;
      return r;
    }
  }
  const int occ = c->render_blocks_per_cu;
  // coordinate search: wavefront size (path slots, also the queue capacity) x
  // blocks per CU with graph replay; then the extend kernel's (stack depth,
  // block size); then the launch form (stream launches, one persistent launch)
  const uint32_t slot_opts[] = {1u << 23, 1u << 24, 1u << 25, 1u << 26};
  const int blk_opts[] = {occ, std::max(2, (3 * occ) / 4)};
  for (uint32_t sl : slot_opts)
    for (int bk : blk_opts) {
      rx_render_tuning t{};
      t.path_slots = sl;
      t.blocks_per_cu = (uint32_t)bk;
      t.use_graph = 1;
      if (int r = trial(t)) return r;
    }
  rx_render_tuning base = win;
  for (const ExtendShape& e : kExtendShapes) {
    if (e.depth == kLdsDepth && e.block == kRBlock) continue;
    rx_render_tuning t = base;
    t.stack_depth = (uint32_t)e.depth;
    t.block_size = (uint32_t)e.block;
    if (int r = trial(t)) return r;
  }
  base = win;  // camera rays in packets instead of one per lane
  {
    rx_render_tuning t = base;
    t.camera_rays = 2;
    if (int r = trial(t)) return r;
  }
  base = win;
  for (uint32_t g : {0u, 2u}) {
    rx_render_tuning t = base;
    t.use_graph = g;
    if (g == 2) t.stack_depth = t.block_size = t.camera_rays = 0;  // the persistent kernel has the default shape
    if (int r = trial(t)) return r;
  }
  c->tuning = win;
  if (best) *best = win;
  if (best_ms) *best_ms = win_ms;  // per sample
  return RX_OK;
}

int rx_batch_get_image(rx_batch* b, uint32_t index, float* rgb_mean, double* sum, double* sumsq) {
  if (!b) return RX_ERR_INVALID;
  rx_ctx* c = b->c;
  if (index >= b->items.size()) return set_err(c, RX_ERR_INVALID, "rx_batch_get_image: no such scene");
  if (b->sum.bytes < 24 * b->total_pix) return set_err(c, RX_ERR_STATE, "rx_batch_get_image: render first");
  const BatchItem& it = b->items[index];
  const size_t off = 3 * (size_t)it.info.film_base, n = 3 * (size_t)it.npix;
  std::vector<double> tmp;
  double* s_out = sum;
  if (!s_out && rgb_mean) {
    tmp.resize(n);
    s_out = tmp.data();
  }
  hipStream_t s = c->stream;
  if (s_out && hip_err(c, hipMemcpyAsync(s_out, b->sum.as<double>() + off, 8 * n, hipMemcpyDeviceToHost, s), "image"))
    return RX_ERR_HIP;
  if (sumsq && hip_err(c, hipMemcpyAsync(sumsq, b->sumsq.as<double>() + off, 8 * n, hipMemcpyDeviceToHost, s), "image"))
    return RX_ERR_HIP;
  if (int r = hip_err(c, hipStreamSynchronize(s), "image")) return r;
  if (rgb_mean && it.prm.spp) {  // the same double product as rx_render's k_mean
    const double inv = 1.0 / (double)it.prm.spp;
    for (size_t i = 0; i < n; i++) rgb_mean[i] = (float)(s_out[i] * inv);
  }
  return RX_OK;
}

}  // extern "C"
