// Scene loading and the wavefront path tracer.
//
// One pass renders one sample per pixel as a sequence of kernels over
// structure-of-arrays path state resident in HBM:
//   generate -> { extend(b) -> shade(b) -> shadow(b) } for b = 0..max_depth -> accumulate
// extend is the nearest-hit traversal of the reference tree with the
// NoSelfIntersectionCondition legality test (no condition for camera rays,
// src/Camera.cpp:23); shade restates the reference's scatter functions
// (src/DiffuseMaterial.cpp, src/GlossyMaterial.cpp, src/RefractiveMaterial.cpp,
// src/SceneEmissiveObject.cpp, src/SceneVPMObject.cpp) and the PathTracer
// product (src/PathTracer.cpp:45-67) incrementally; shadow answers the
// next-event queries.  Queues are compacted with a wave ballot + one atomic
// per wave.  Random numbers are Philox counters keyed by (pixel, seed) with
// counter (sample, bounce, stream): each path's result is independent of
// queue order, and the film adds each pixel's samples in sample order in
// FP64, so renders are bit-reproducible.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>
#include <deque>
#include <new>
#include <vector>

#include "rx_render_state.h"

namespace rx {

static void leaf_verts_of(const HostTree& t, const float* verts, const uint32_t* tris, std::vector<float>& lv) {
  lv.resize(9 * t.leaf_tris.size());
  for (size_t k = 0; k < t.leaf_tris.size(); k++)
    for (int j = 0; j < 3; j++) memcpy(&lv[9 * k + 3 * j], &verts[3 * (size_t)tris[3 * (size_t)t.leaf_tris[k] + j]], 12);
}

}  // namespace rx

namespace rx {
int load_scene_body(rx_ctx* c, uint32_t n_verts, const float* verts, uint32_t n_tris, const uint32_t* tris,
                    uint32_t n_objects, const rx_object_desc* objects, uint32_t tree_flags,
                    const std::vector<int32_t>& tri_obj, const std::vector<uint32_t>& local);
}

using namespace rx;

extern "C" {

void rx_render_state_free(rx_ctx* c) {
  if (c && c->render) {
    delete c->render;
    c->render = nullptr;
  }
}

int rx_load_scene(rx_ctx* c, uint32_t n_verts, const float* verts, uint32_t n_tris, const uint32_t* tris,
                  uint32_t n_objects, const rx_object_desc* objects, uint32_t tree_flags) {
  if (!c) return RX_ERR_INVALID;
  if (!verts || !tris || !objects || n_tris == 0 || n_objects == 0)
    return set_err(c, RX_ERR_INVALID, "rx_load_scene: empty scene");
  for (size_t i = 0; i < 3 * (size_t)n_tris; i++)
    if (tris[i] >= n_verts) return set_err(c, RX_ERR_INVALID, "rx_load_scene: vertex index out of range");
  try {
    std::vector<int32_t> tri_obj(n_tris, -1);
    std::vector<uint32_t> local(n_tris, 0);
    for (uint32_t o = 0; o < n_objects; o++) {
      const rx_object_desc& od = objects[o];
      if (od.type < 0 || od.type > RX_OBJ_VPM) return set_err(c, RX_ERR_INVALID, "rx_load_scene: bad object type");
      if ((uint64_t)od.tri_begin + od.tri_count > n_tris) return set_err(c, RX_ERR_INVALID, "rx_load_scene: object range");
      for (uint32_t t = 0; t < od.tri_count; t++) {
        if (tri_obj[od.tri_begin + t] != -1) return set_err(c, RX_ERR_INVALID, "rx_load_scene: overlapping objects");
        tri_obj[od.tri_begin + t] = (int32_t)o;
        local[od.tri_begin + t] = t;
      }
    }
    for (uint32_t t = 0; t < n_tris; t++)
      if (tri_obj[t] < 0) return set_err(c, RX_ERR_INVALID, "rx_load_scene: triangle without object");
    (void)hipSetDevice(c->device);
    // from the first destructive step on, the context holds no scene and no
    // tree until every upload below has succeeded
    c->scene_loaded = false;
    c->tree_loaded = false;
    rx_render_state_free(c);
    return load_scene_body(c, n_verts, verts, n_tris, tris, n_objects, objects, tree_flags, tri_obj, local);
  } catch (const std::bad_alloc&) {
    c->scene_loaded = c->tree_loaded = false;
    return set_err(c, RX_ERR_NO_MEMORY, "rx_load_scene: host allocation failed");
  }
}

}  // extern "C"

namespace rx {

int load_scene_body(rx_ctx* c, uint32_t n_verts, const float* verts, uint32_t n_tris, const uint32_t* tris,
                    uint32_t n_objects, const rx_object_desc* objects, uint32_t tree_flags,
                    const std::vector<int32_t>& tri_obj, const std::vector<uint32_t>& local) {
  RenderState* R = new (std::nothrow) RenderState;
  if (!R) return RX_ERR_NO_MEMORY;
  c->render = R;
  const bool abs_int = (tree_flags & RX_TREE_ABS_INT) != 0;
  // global tree (Scene::buildKDTree)
  build_tree(verts, n_verts, tris, n_tris, abs_int, c->host_tree);
  leaf_verts_of(c->host_tree, verts, tris, c->leaf_verts);
  c->tri_obj = tri_obj;
  c->tri_normal.clear();
  {
    DeviceTreeHost dh;
    make_device_tree(c->host_tree, c->leaf_verts, tri_obj.data(), local.data(), dh);
    if (int r = hip_err(c, c->tree.upload(dh, c->host_tree.max_depth, c->stream), "tree upload")) return r;
    c->tree_loaded = true;
  }
  SceneBufs& SB = R->scene;
  SB.max_depth = c->host_tree.max_depth;
  memcpy(R->box_lo, c->host_tree.nodes[0].lo, 12);
  memcpy(R->box_hi, c->host_tree.nodes[0].hi, 12);
  // objects
  std::vector<DevTree> vt;
  std::vector<int32_t> vobj;
  SB.host_objs.resize(n_objects);
  for (uint32_t o = 0; o < n_objects; o++) {
    const rx_object_desc& od = objects[o];
    DevObject& D = SB.host_objs[o];
    D = make_devobject(od);
    if (od.type == RX_OBJ_VPM) {  // per-object tree (Scene::buildObjKDTrees)
      HostTree ht;
      build_tree(verts, n_verts, tris + 3 * (size_t)od.tri_begin, od.tri_count, abs_int, ht);
      std::vector<float> lv;
      leaf_verts_of(ht, verts, tris + 3 * (size_t)od.tri_begin, lv);
      DeviceTreeHost dh;
      make_device_tree(ht, lv, nullptr, nullptr, dh);
      TreeBufs* tb = new TreeBufs;
      SB.vols.push_back(tb);
      if (int r = hip_err(c, tb->upload(dh, ht.max_depth, c->stream), "volume tree upload")) return r;
      SB.max_depth = std::max(SB.max_depth, ht.max_depth);
      D.vol = (int32_t)vt.size();
      vt.push_back(tb->dev);
      vobj.push_back((int32_t)o);
    }
  }
  std::vector<float> etri, enrm, ecdf;
  const double total = collect_emitters(verts, tris, n_objects, objects, 0, etri, enrm, ecdf);
  hipStream_t s = c->stream;
  if (int r = hip_err(c, SB.objs.upload(SB.host_objs.data(), SB.host_objs.size() * sizeof(DevObject), s), "objs")) return r;
  if (int r = hip_err(c, SB.vol_trees.upload(vt.data(), vt.size() * sizeof(DevTree), s), "vol trees")) return r;
  if (int r = hip_err(c, SB.vol_obj.upload(vobj.data(), vobj.size() * 4, s), "vol obj")) return r;
  if (int r = hip_err(c, SB.etri.upload(etri.data(), etri.size() * 4, s), "etri")) return r;
  if (int r = hip_err(c, SB.enrm.upload(enrm.data(), enrm.size() * 4, s), "enrm")) return r;
  if (int r = hip_err(c, SB.ecdf.upload(ecdf.data(), ecdf.size() * 4, s), "ecdf")) return r;
  DevScene& D = SB.dev;
  D.tree = c->tree.dev;
  D.objs = SB.objs.as<DevObject>();
  D.n_objs = (int32_t)n_objects;
  D.vol = SB.vol_trees.as<DevTree>();
  D.vol_obj = SB.vol_obj.as<int32_t>();
  D.n_vol = (int32_t)vt.size();
  D.etri = SB.etri.as<float4>();
  D.enrm = SB.enrm.as<float4>();
  D.ecdf = SB.ecdf.as<float>();
  D.n_etri = (int32_t)ecdf.size();
  D.total_area = (float)total;
  c->scene_loaded = true;
  return RX_OK;
}

}  // namespace rx

extern "C" {

int rx_render(rx_ctx* c, const rx_camera_desc* cam, const rx_render_params* prm, float* rgb_mean, double* sum,
              double* sumsq, rx_render_stats* stats) {
  if (!c || !cam || !prm) return set_err(c, RX_ERR_INVALID, "rx_render: null argument");
  if (!c->scene_loaded || !c->render) return set_err(c, RX_ERR_STATE, "rx_render: no scene loaded");
  if (cam->width <= 0 || cam->height <= 0 || (uint64_t)cam->width * cam->height > (1ull << 28))
    return set_err(c, RX_ERR_INVALID, "rx_render: bad resolution");
  if (prm->max_depth >= (uint32_t)kMaxBounce) return set_err(c, RX_ERR_INVALID, "rx_render: max_depth must be < 64");
  if (prm->mode > RX_MODE_BPT) return set_err(c, RX_ERR_INVALID, "rx_render: bad mode");
  if (prm->mode >= RX_MODE_NEE && c->render->scene.dev.n_etri == 0)
    return set_err(c, RX_ERR_INVALID, "rx_render: next-event estimation needs an emitter");
  if (prm->mode == RX_MODE_BPT) return render_bpt(c, cam, prm, rgb_mean, sum, sumsq, stats);
  (void)hipSetDevice(c->device);
  RenderState& R = *c->render;
  hipStream_t s = c->stream;
  const uint32_t npix = (uint32_t)(cam->width * cam->height);
  // samples per pass: as many as the path-slot budget holds, so that the later
  // bounces (after Russian roulette) still fill the GPU
  const uint32_t slots = c->tuning.path_slots ? c->tuning.path_slots : kDefaultPathSlots;
  const uint32_t S_max = std::max<uint32_t>(1u, std::min<uint32_t>(std::max(prm->spp, 1u), slots / npix));
  const uint32_t N = npix * S_max;  // path slots
#define RXCK(x, w) \
  if (int r_ = hip_err(c, (x), w)) return r_;
  if (R.cap < N) {
    RXCK(R.o.ensure(16 * (size_t)N), "alloc");
    RXCK(R.d.ensure(16 * (size_t)N), "alloc");
    RXCK(R.beta.ensure(16 * (size_t)N), "alloc");
    RXCK(R.lastp.ensure(16 * (size_t)N), "alloc");
    RXCK(R.L.ensure(16 * (size_t)N), "alloc");
    RXCK(R.s.ensure(16 * (size_t)N), "alloc");
    RXCK(R.hit.ensure(8 * (size_t)N), "alloc");
    RXCK(R.q0.ensure(4 * (size_t)N), "alloc");
    RXCK(R.q1.ensure(4 * (size_t)N), "alloc");
    RXCK(R.sq.ensure(4 * (size_t)N), "alloc");
    RXCK(R.so.ensure(16 * (size_t)N), "alloc");
    RXCK(R.sd.ensure(16 * (size_t)N), "alloc");
    RXCK(R.sc.ensure(16 * (size_t)N), "alloc");
    RXCK(R.ss.ensure(8 * (size_t)N), "alloc");
    R.cap = N;
  }
  RXCK(R.sum.ensure(24 * (size_t)npix), "alloc");
  RXCK(R.sumsq.ensure(24 * (size_t)npix), "alloc");
  RXCK(R.mean.ensure(12 * (size_t)npix), "alloc");
  RXCK(R.cnt.ensure(4 * kCntWords), "alloc");
  RXCK(R.stats.ensure(8 * 8), "alloc");
  if (!c->render_blocks_per_cu) {  // launches are sized to be resident at once (see k_extend)
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_extend<true, false, false>, kRBlock, 0) != hipSuccess || nb < 1)
      nb = 4;
    c->render_blocks_per_cu = nb;
  }
  const int per_cu = c->tuning.blocks_per_cu ? std::min<int>((int)c->tuning.blocks_per_cu, c->render_blocks_per_cu)
                                             : c->render_blocks_per_cu;
  const int blocks = (int)std::min<uint64_t>((N + kRBlock - 1) / kRBlock, (uint64_t)c->n_cu * per_cu);
  // the shade and shadow kernels take any grid (chunks of their queue), so
  // theirs is what they hold resident (they need more registers than the
  // extend kernel): a larger grid would run a second, partial round of blocks
  auto resident = [&](const void* k) {
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k, kRBlock, 0) != hipSuccess || nb < 1) nb = per_cu;
    return (int)std::min<uint64_t>((N + kRBlock - 1) / kRBlock, (uint64_t)c->n_cu * std::min(nb, per_cu));
  };
  int sd = 0;
  if (int r = ensure_spill(c, blocks, kRBlock, (int)R.scene.max_depth, &sd)) return r;
  PathBufs B;
  B.o = R.o.as<float4>();
  B.d = R.d.as<float4>();
  B.beta = R.beta.as<float4>();
  B.lastp = R.lastp.as<float4>();
  B.L = R.L.as<float4>();
  B.s = R.s.as<int4>();
  B.hit = R.hit.as<int2>();
  B.q0 = R.q0.as<uint32_t>();
  B.q1 = R.q1.as<uint32_t>();
  B.sq = R.sq.as<uint32_t>();
  B.cnt = R.cnt.as<uint32_t>();
  B.so = R.so.as<float4>();
  B.sd = R.sd.as<float4>();
  B.sc = R.sc.as<float4>();
  B.ss = R.ss.as<int2>();
  B.sum = R.sum.as<double>();
  B.sumsq = R.sumsq.as<double>();
  B.stats = R.stats.as<unsigned long long>();
  B.spill = c->spill.as<uint2>();
  B.spill_depth = sd;
  B.overflow = c->overflow.as<int32_t>();
  const DevCam C = make_devcam(cam);
  PassArgs P;
  P.n_paths = N;
  P.npix = npix;
  P.spass = 1;
  P.key1 = seed_key(prm->seed);
  P.max_depth = prm->max_depth;
  P.rr_start = prm->rr_start_depth;
  P.mode = prm->mode;
  const DevScene S = R.scene.dev;
  const bool vol = S.n_vol > 0;  // kernels without the medium code when the scene has no volumes
  const bool timing = (prm->flags & RX_RENDER_TIMING) != 0;
  const bool count = (prm->flags & RX_RENDER_COUNT_WORK) != 0;
  const int mode = (int)prm->mode;
  GuideArgs GA{};
  const BatchArgs BA0{};
  uint32_t train_passes = 0, batch = 0, steps = 0, store_cap = 0, store_total = 0, s_train = 1, train_end = 0;
  bool keep = false;
  if (mode == RX_MODE_GUIDED) {
    if (!c->guide) return set_err(c, RX_ERR_STATE, "rx_render: guided mode needs rx_guide_init");
    GuideState& G = *c->guide;
    for (int a = 0; a < 3; a++) {  // encode positions relative to the scene box (slightly padded)
      float pad = 1e-3f * (R.box_hi[a] - R.box_lo[a]) + 1e-6f;
      G.lo[a] = R.box_lo[a] - pad;
      G.hi[a] = R.box_hi[a] + pad;
    }
    // The online warm-up: train_passes passes of s_train samples per pixel
    // each, the leading train_end samples of every pixel.  A pass has enough
    // paths for its sample store (samples_per_pass rows, by default batch *
    // steps, an evenly strided subset of the pass's records) and at least
    // kTrainPassPaths paths: a
    // pass with fewer waits on its longest paths at every bounce.  A render
    // resumed with sample_offset >= train_end only uses the network.
    constexpr uint32_t kTrainPassPaths = 1u << 20;
    train_passes = G.prm.train_passes;
    batch = std::max(256u, G.prm.batch_size);
    steps = std::max(1u, G.prm.steps_per_pass);
    const uint64_t want = G.prm.samples_per_pass ? G.prm.samples_per_pass : (uint64_t)batch * steps;
    store_cap = (uint32_t)std::min<uint64_t>(std::max<uint64_t>(want, batch), 1u << 26);
    // with kept samples the store holds every warm-up pass's samples, pass k at rows k * store_cap
    keep = (G.prm.flags & RX_GUIDE_KEEP_SAMPLES) != 0;
    if (keep && (uint64_t)store_cap * std::max(1u, G.prm.train_passes) > (1u << 27))
      return set_err(c, RX_ERR_INVALID, "rx_render: kept guide samples (samples_per_pass * train_passes) exceed 2^27");
    store_total = keep ? store_cap * std::max(1u, G.prm.train_passes) : store_cap;
    s_train = std::max((store_cap + npix - 1) / npix, (kTrainPassPaths + npix - 1) / npix);
    s_train = std::max(1u, std::min(S_max, s_train));
    train_end = train_passes * s_train;
    G.decay_steps = train_passes * steps;
    const uint32_t Rs = prm->max_depth + 1;
    // a training pass keeps one record per path slot and bounce, ranked with 32-bit indices
    if ((uint64_t)npix * s_train * Rs >= (1ull << 31))
      return set_err(c, RX_ERR_INVALID, "rx_render: guided mode needs pixels * (max_depth + 1) < 2^31");
    if (int r = guide_reserve(c, G, batch)) return r;  // the training step's rows
    if (prm->sample_offset < train_end && prm->spp) {
      if (int r = guide_reserve_store(c, G, (store_total + 1) & ~1u)) return r;
      // the warm-up starts (or the network was re-initialised): no kept samples yet (zero weights add
      // nothing to the loss)
      if (keep && (prm->sample_offset == 0 || G.step == 0))
        hipLaunchKernelGGL(k_fill, dim3(256), dim3(256), 0, s, G.swt.as<double>(), (size_t)(store_total + 1) / 2);
      const size_t K = (size_t)npix * s_train * Rs;
      RXCK(R.rec.ensure(K * kRecFloats * 4), "alloc records");
      RXCK(R.bcount.ensure((K / kRecChunk + 2) * 4), "alloc records");
      RXCK(R.total.ensure(64), "alloc records");
      // tags restart at 1 every render: clear stale records once
      hipLaunchKernelGGL(k_fill, dim3(1024), dim3(256), 0, s, R.rec.as<double>(), K * kRecFloats / 2);
    }
    GA.net = G.net();
    RXCK(R.ghead.ensure(kGHead * (size_t)N), "alloc guide heads");      // 208 B per path slot
    RXCK(R.gquery.ensure(32 * (size_t)N), "alloc guide queries");       // 32 B per path slot
    RXCK(R.grow.ensure(4 * (size_t)N), "alloc guide rows");
    GA.head = R.ghead.as<unsigned char>();
    GA.query = R.gquery.as<float4>();
    GA.row = R.grow.as<uint32_t>();
    GA.rec = R.rec.as<float>();
    GA.R = Rs;
  }
  // camera rays as packets unless the tuning asks for lanes or the tree is deeper than the packet stack
  // (the later bounces' rays leave a pixel's hit points in scattered directions: packets of them walk
  // the union of their paths through the tree and are slower than one ray per lane)
  const bool packets = c->tuning.camera_rays != 1 && R.scene.max_depth < (uint32_t)kPacketDepth;
  std::vector<hipEvent_t> evs;
  auto mark = [&](void) -> hipEvent_t {
    hipEvent_t e;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    evs.push_back(e);
    (void)hipEventRecord(e, s);
    return e;
  };
  hipLaunchKernelGGL(k_fill, dim3(1024), dim3(256), 0, s, B.sum, 3 * (size_t)npix);
  hipLaunchKernelGGL(k_fill, dim3(1024), dim3(256), 0, s, B.sumsq, 3 * (size_t)npix);
  hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(64), 0, s, B.stats, 8, B.overflow);
  RXCK(hipGetLastError(), "launch fill");
  RXCK(hipEventRecord(c->ev0, s), "event");
  std::vector<std::pair<hipEvent_t, hipEvent_t>> ext, shd, sdw;
  std::vector<uint32_t> ext_b;  // bounce of each timed extend launch
  // counting render: the work counters after each counted extend launch (differences give the bounce's work)
  // (a deque: an element keeps its address while its copy is in flight)
  std::deque<std::array<unsigned long long, 8>> snaps;
  std::vector<uint32_t> snap_b;
  uint64_t launches = 0, train_rows = 0;
  uint32_t pass = 0;
  // RX_MODE_PT without media: the bounces from tail_b on run as one k_tail launch
  // (the counting render keeps one counted extend launch per bounce)
  const void* shade_fn = mode == RX_MODE_GUIDED ? (vol ? (const void*)k_shade<2, false, true> : (const void*)k_shade<2, false, false>)
                         : mode == RX_MODE_NEE   ? (vol ? (const void*)k_shade<1, false, true> : (const void*)k_shade<1, false, false>)
                                                 : (vol ? (const void*)k_shade<0, false, true> : (const void*)k_shade<0, false, false>);
  const int sblocks = resident(shade_fn);
  // the shadow kernel's waves take their first chunk by wave index: a wave
  // that is not resident yet would hold its chunk back until a second round
  const int shblocks = resident(vol ? (const void*)k_shadow<false, true> : (const void*)k_shadow<false, false>);
  const uint32_t tail_b = (mode == RX_MODE_PT && !vol && !count)
                              ? std::max(1u, c->tuning.tail_bounce ? c->tuning.tail_bounce : kDefaultTailBounce)
                              : ~0u;
  for (uint32_t done = 0; done < prm->spp; pass++) {
    P.sample = prm->sample_offset + done;
    // a training pass renders up to s_train samples per pixel (to the next pass boundary); otherwise up to
    // S_max samples per pass
    const bool train = mode == RX_MODE_GUIDED && P.sample < train_end;
    const uint32_t spass = train ? std::min(s_train - P.sample % s_train, prm->spp - done)
                                 : std::min(S_max, prm->spp - done);
    P.n_paths = npix * spass;
    P.spass = spass;
    done += spass;
    GA.tag = train ? (float)(pass + 1) : 0.0f;
    if (vol) hipLaunchKernelGGL((k_generate<false, true>), dim3(blocks), dim3(kRBlock), 0, s, S, C, B, P, BA0);
    else hipLaunchKernelGGL((k_generate<false, false>), dim3(blocks), dim3(kRBlock), 0, s, S, C, B, P, BA0);
    for (uint32_t b = 0; b <= prm->max_depth; b++) {
      if (b == tail_b) {  // bounces b.. to the paths' ends in one launch, timed as an extend launch
        hipEvent_t e0 = timing ? mark() : nullptr;
        hipLaunchKernelGGL(k_tail, dim3(blocks), dim3(kRBlock), 0, s, S, B, P, b);
        launches++;
        if (timing) {
          ext.push_back({e0, mark()});
          ext_b.push_back(b);
        }
        break;
      }
      hipEvent_t e0 = timing ? mark() : nullptr;
      const uint32_t sdep = c->tuning.stack_depth, sblk = c->tuning.block_size;
      if (b == 0) {
        if (count) launch_extend<false, true, false>(sdep, sblk, blocks, s, S.tree, B, b, BA0);
        else if (packets) hipLaunchKernelGGL(k_extend_packet<false>, dim3(blocks), dim3(kRBlock), 0, s, S.tree, B, b, BA0);
        else launch_extend<false, false, false>(sdep, sblk, blocks, s, S.tree, B, b, BA0);
      } else {
        if (count) launch_extend<true, true, false>(sdep, sblk, blocks, s, S.tree, B, b, BA0);
        else launch_extend<true, false, false>(sdep, sblk, blocks, s, S.tree, B, b, BA0);
      }
      launches++;
      if (count) {  // pageable destination: the copy is ordered on the stream and read after the final sync
        snaps.emplace_back();
        snap_b.push_back(b);
        RXCK(hipMemcpyAsync(snaps.back().data(), B.stats, 64, hipMemcpyDeviceToHost, s), "download counters");
      }
      hipEvent_t e1 = timing ? mark() : nullptr;
      if (mode == RX_MODE_GUIDED) {  // network inference for this bounce's vertices
        if (vol) hipLaunchKernelGGL(k_guide_query<true>, dim3(blocks), dim3(kRBlock), 0, s, S, B, P, GA, b);
        else hipLaunchKernelGGL(k_guide_query<false>, dim3(blocks), dim3(kRBlock), 0, s, S, B, P, GA, b);
        RXCK(guide_infer(GA.net, GA.query, GA.query + 1, 2, B.cnt + kGuideRowBase + b, P.n_paths, nullptr, GA.head,
                         c->n_cu, s),
             "guide inference");
        if (vol) hipLaunchKernelGGL((k_shade<2, false, true>), dim3(sblocks), dim3(kRBlock), 0, s, S, B, P, b, GA, BA0);
        else hipLaunchKernelGGL((k_shade<2, false, false>), dim3(sblocks), dim3(kRBlock), 0, s, S, B, P, b, GA, BA0);
      } else if (mode == RX_MODE_NEE) {
        if (vol) hipLaunchKernelGGL((k_shade<1, false, true>), dim3(sblocks), dim3(kRBlock), 0, s, S, B, P, b, GA, BA0);
        else hipLaunchKernelGGL((k_shade<1, false, false>), dim3(sblocks), dim3(kRBlock), 0, s, S, B, P, b, GA, BA0);
      } else {
        if (vol) hipLaunchKernelGGL((k_shade<0, false, true>), dim3(sblocks), dim3(kRBlock), 0, s, S, B, P, b, GA, BA0);
        else hipLaunchKernelGGL((k_shade<0, false, false>), dim3(sblocks), dim3(kRBlock), 0, s, S, B, P, b, GA, BA0);
      }
      hipEvent_t e2 = timing ? mark() : nullptr;
      hipEvent_t e3 = nullptr;
      if (mode) {
        if (vol) hipLaunchKernelGGL((k_shadow<false, true>), dim3(shblocks), dim3(kRBlock), 0, s, S, B, P, b, GA, BA0);
        else hipLaunchKernelGGL((k_shadow<false, false>), dim3(shblocks), dim3(kRBlock), 0, s, S, B, P, b, GA, BA0);
        e3 = timing ? mark() : nullptr;
      }
      if (timing) {
        ext.push_back({e0, e1});
        ext_b.push_back(b);
        shd.push_back({e1, e2});
        if (mode) sdw.push_back({e2, e3});
      }
    }
    hipLaunchKernelGGL(k_accumulate, dim3(blocks), dim3(kAccPix), 0, s, B, P);
    RXCK(hipGetLastError(), "launch pass");
    if (GA.tag > 0.0f) {  // online training on this pass's radiance samples
      GuideState& G = *c->guide;
      const uint32_t K = P.n_paths * GA.R;
      const uint32_t nb = (K + kRecChunk - 1) / kRecChunk;
      hipLaunchKernelGGL(k_rec_count, dim3(nb), dim3(256), 0, s, GA.rec, K, GA.tag, R.bcount.as<uint32_t>());
      hipLaunchKernelGGL(k_rec_scan, dim3(1), dim3(kScanThreads), 0, s, R.bcount.as<uint32_t>(), nb,
                         R.total.as<uint32_t>());
      // this warm-up pass's rows of the store (with kept samples, after the earlier passes' rows)
      const uint32_t tp = std::min(P.sample / s_train, std::max(1u, train_passes) - 1);
      const size_t off = keep ? (size_t)tp * store_cap : 0;
      const uint32_t kept = keep ? (tp + 1) * store_cap : 0;
      const TrainOut ST{G.spos.as<float4>() + off, G.sdir.as<float4>() + off, G.swdir.as<float4>() + off,
                        G.spb.as<float>() + off, G.swt.as<float>() + off, G.srows.as<uint32_t>()};
      const TrainOut SK{G.spos.as<float4>(), G.sdir.as<float4>(), G.swdir.as<float4>(), G.spb.as<float>(),
                        G.swt.as<float>(), G.srows.as<uint32_t>()};
      const TrainOut TB{G.pos.as<float4>(), G.dir.as<float4>(), G.wdir.as<float4>(), G.pb.as<float>(),
                        G.wt.as<float>(), G.rows.as<uint32_t>()};
      hipLaunchKernelGGL(k_rec_write, dim3(nb), dim3(256), 0, s, GA.rec, K, GA.tag, R.bcount.as<uint32_t>(),
                         R.total.as<uint32_t>(), store_cap, B.L, GA.R, ST);
      RXCK(hipGetLastError(), "launch record write");
      for (uint32_t k = 0; k < steps; k++) {
        hipLaunchKernelGGL(k_guide_gather, dim3(std::min<uint32_t>((batch + 255) / 256, 1024u)), dim3(256), 0, s, SK,
                           R.total.as<uint32_t>(), store_cap, kept, k, batch, TB);
        RXCK(hipGetLastError(), "launch gather");
        if (int r = guide_train_rows(c, batch)) return r;
        train_rows += batch;
      }
    }
  }
  RXCK(hipEventRecord(c->ev1, s), "event");
  if (rgb_mean && prm->spp) {
    hipLaunchKernelGGL(k_mean, dim3(1024), dim3(256), 0, s, B.sum, R.mean.as<float>(), 3 * (size_t)npix,
                       1.0 / (double)prm->spp);
    RXCK(hipMemcpyAsync(rgb_mean, R.mean.p, 12 * (size_t)npix, hipMemcpyDeviceToHost, s), "download mean");
  }
  if (sum) RXCK(hipMemcpyAsync(sum, B.sum, 24 * (size_t)npix, hipMemcpyDeviceToHost, s), "download sum");
  if (sumsq) RXCK(hipMemcpyAsync(sumsq, B.sumsq, 24 * (size_t)npix, hipMemcpyDeviceToHost, s), "download sumsq");
  unsigned long long st[8] = {0};
  int32_t ovf = 0;
  RXCK(hipMemcpyAsync(st, B.stats, 64, hipMemcpyDeviceToHost, s), "download stats");
  RXCK(hipMemcpyAsync(&ovf, B.overflow, 4, hipMemcpyDeviceToHost, s), "download overflow");
  RXCK(hipStreamSynchronize(s), "render synchronize");
  float ms = 0;
  RXCK(hipEventElapsedTime(&ms, c->ev0, c->ev1), "elapsed");
  if (stats) {
    memset(stats, 0, sizeof(*stats));
    stats->rays = st[0];
    stats->extend_rays = st[1];
    stats->shadow_rays = st[2];
    stats->segments = st[3];
    stats->box_tests = st[4];
    stats->tri_tests = st[5];
    stats->samples = (uint64_t)npix * prm->spp;
    stats->ms_total = ms;
    stats->extend_launches = launches;
    stats->guide_rows = st[6];
    stats->guide_train_rows = train_rows;
    auto tsum = [](const std::vector<std::pair<hipEvent_t, hipEvent_t>>& v) {
      double t = 0;
      for (auto& e : v) {
        float m = 0;
        if (e.first && e.second && hipEventElapsedTime(&m, e.first, e.second) == hipSuccess) t += m;
      }
      return t;
    };
    stats->ms_extend = tsum(ext);
    const uint32_t lastb = RX_STATS_BOUNCES - 1;
    for (size_t i = 0; i < ext.size(); i++) {
      float m = 0;
      if (ext[i].first && ext[i].second && hipEventElapsedTime(&m, ext[i].first, ext[i].second) == hipSuccess)
        stats->ms_extend_bounce[std::min(ext_b[i], lastb)] += m;
    }
    std::array<unsigned long long, 8> prev{};
    for (size_t i = 0; i < snaps.size(); i++) {
      const uint32_t b = std::min(snap_b[i], lastb);
      stats->extend_rays_bounce[b] += snaps[i][1] - prev[1];
      stats->box_tests_bounce[b] += snaps[i][4] - prev[4];
      stats->tri_tests_bounce[b] += snaps[i][5] - prev[5];
      prev = snaps[i];
    }
    stats->ms_shade = tsum(shd);
    stats->ms_shadow = tsum(sdw);
  }
  for (auto e : evs) (void)hipEventDestroy(e);
#undef RXCK
  if (ovf) return set_err(c, RX_ERR_OVERFLOW, "rx_render: traversal stack overflow");
  return RX_OK;
}

int rx_set_render_tuning(rx_ctx* c, const rx_render_tuning* t) {
  if (!c || !t) return RX_ERR_INVALID;
  if (t->path_slots && (t->path_slots < 4096 || t->path_slots > (1u << 27)))
    return set_err(c, RX_ERR_INVALID, "rx_set_render_tuning: path_slots must be 0 or in [4096, 2^27]");
  if (t->blocks_per_cu > 64) return set_err(c, RX_ERR_INVALID, "rx_set_render_tuning: blocks_per_cu > 64");
  if (t->use_graph > 2) return set_err(c, RX_ERR_INVALID, "rx_set_render_tuning: use_graph is 0, 1 or 2");
  if (t->camera_rays > 2) return set_err(c, RX_ERR_INVALID, "rx_set_render_tuning: camera_rays is 0, 1 or 2");
  if (!extend_shape_ok(t->stack_depth, t->block_size))
    return set_err(c, RX_ERR_INVALID,
                   "rx_set_render_tuning: (stack_depth, block_size) must be (0|10, 0|64|128|256) or (6|8|12|16, 0|128)");
  c->tuning = *t;
  return RX_OK;
}

int rx_get_render_tuning(rx_ctx* c, rx_render_tuning* t) {
  if (!c || !t) return RX_ERR_INVALID;
  *t = c->tuning;
  if (!t->path_slots) t->path_slots = kDefaultPathSlots;
  return RX_OK;
}

int rx_autotune_render(rx_ctx* c, const rx_camera_desc* cam, const rx_render_params* prm, rx_render_tuning* best,
                       float* best_ms) {
  if (!c || !cam || !prm) return RX_ERR_INVALID;
  if (!c->render) return set_err(c, RX_ERR_STATE, "rx_autotune_render: no scene loaded");
  rx_render_params p = *prm;
  if (p.mode == RX_MODE_GUIDED) p.mode = RX_MODE_NEE;  // timing must not train the network
  p.flags = 0;
  const rx_render_tuning keep = c->tuning;
  if (!c->render_blocks_per_cu) {  // occupancy limit, queried by a first render
    rx_render_params q = p;
    q.spp = 1;
    if (int r = rx_render(c, cam, &q, nullptr, nullptr, nullptr, nullptr)) return r;
  }
  const int occ = c->render_blocks_per_cu;
  rx_render_tuning win = keep;
  float win_ms = 0.0f;
  bool have = false;
  auto trial = [&](const rx_render_tuning& t) -> int {
    c->tuning = t;
    rx_render_stats st;
    if (int r = rx_render(c, cam, &p, nullptr, nullptr, nullptr, &st)) {
      c->tuning = keep;
      return r;
    }
    if (!have || st.ms_total < win_ms) {
      win = t;
      win_ms = (float)st.ms_total;
      have = true;
    }
    return RX_OK;
  };
  // coordinate search: path slots x blocks per CU at the default extend
  // shape, then the extend kernel's (stack depth, block size) at the winner,
  // the camera-ray form and, for RX_MODE_PT, the tail launch's first bounce
  const uint32_t slot_opts[] = {1u << 22, 1u << 23, 1u << 24, 1u << 25, 1u << 26};
  const int blk_opts[] = {occ, std::max(2, (3 * occ) / 4)};
  for (uint32_t sl : slot_opts) {
    for (int bk : blk_opts) {
      rx_render_tuning t{};
      t.path_slots = sl;
      t.blocks_per_cu = (uint32_t)bk;
      if (int r = trial(t)) return r;
    }
  }
  rx_render_tuning base = win;
  for (const ExtendShape& e : kExtendShapes) {
    if (e.depth == kLdsDepth && e.block == kRBlock) continue;  // timed above
    rx_render_tuning t = base;
    t.stack_depth = (uint32_t)e.depth;
    t.block_size = (uint32_t)e.block;
    if (int r = trial(t)) return r;
  }
  base = win;  // then camera rays one per lane instead of in packets
  {
    rx_render_tuning t = base;
    t.camera_rays = 1;
    if (int r = trial(t)) return r;
  }
  if (p.mode == RX_MODE_PT) {  // then the first bounce of the one-launch tail (64: none)
    base = win;
    for (uint32_t tb : {2u, 4u, 64u}) {
      rx_render_tuning t = base;
      t.tail_bounce = tb;
      if (int r = trial(t)) return r;
    }
  }
  c->tuning = win;
  if (best) *best = win;
  if (best_ms) *best_ms = win_ms;
  return RX_OK;
}

}  // extern "C"
