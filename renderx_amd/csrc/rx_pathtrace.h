// Device side of the wavefront path tracer: path-state layout, the
// restated scatter functions and the generate / extend / shade / shadow /
// accumulate kernels, shared by the single-scene renderer (rx_render.hip)
// and the multi-scene batch renderer (rx_batch.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <stdint.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "rx_context.h"
#include "rx_guide.h"
#include "rx_kernels.h"

namespace rx {

enum { OBJ_DIFFUSE = 0, OBJ_GLOSSY = 1, OBJ_PHONG = 2, OBJ_REFRACTIVE = 3, OBJ_EMISSIVE = 4, OBJ_VPM = 5 };
constexpr int NONE = -1, CAMERA = -2;
static_assert(kNoCondition != NONE && kNoCondition != CAMERA, "distinct contact sentinels");
constexpr int kRBlock = kTraceBlock;  // render kernels share the traversal block shape
#ifndef RX_EXTEND_MIN_WAVES
#define RX_EXTEND_MIN_WAVES 7
#endif
// minimum resident waves per SIMD the extend kernel's registers must allow:
// 7 (72 VGPRs, a 44-byte spill at the ray set-up; 14 blocks of 10 KB of LDS
// stack per CU) traces the bench render's bounces in 27.2 ms against 27.9 at
// 6 (80 VGPRs), 27.9 at 8 and 29.3 at 5
constexpr int kExtendMinWaves = RX_EXTEND_MIN_WAVES;
#ifndef RX_SHADOW_MIN_WAVES
#define RX_SHADOW_MIN_WAVES 4
#endif
constexpr int kShadowMinWaves = RX_SHADOW_MIN_WAVES;  // the same for the shadow kernel
constexpr int kMaxBounce = 64;
constexpr int kShadowBase = 96;       // cnt[kShadowBase + b]: shadow queue length of bounce b
constexpr int kSpareWords = 192;      // cnt[192, 256): single counters of other passes
constexpr int kGuideRowBase = 324;    // cnt[kGuideRowBase + b]: network rows of bounce b (guided mode)
// work counter blocks (kWorkWords each): extend of bounce b at kWorkBase + b * kWorkWords,
// shadow of bounce b at kShadowWorkBase + b * kWorkWords, one spare block after them
constexpr int kWorkBase = 512;
constexpr int kShadowWorkBase = kWorkBase + kMaxBounce * (int)kWorkWords;
constexpr int kSpareWork = kShadowWorkBase + kMaxBounce * (int)kWorkWords;
constexpr int kCntWords = kSpareWork + (int)kWorkWords;
static_assert(kShadowBase + kMaxBounce <= kSpareWords && kSpareWords + 64 <= kGuideRowBase &&
                  kGuideRowBase + kMaxBounce <= kWorkBase,
              "counter regions do not overlap");
constexpr uint32_t kDefaultPathSlots = 1u << 26;  // rx_render_tuning.path_slots default (12.6 GB of path state)
constexpr int kMaxCross = 4;          // index-matched boundary crossings per shadow ray
constexpr uint32_t FLAG_MIS = 1u;     // last scatter was RANDOM (only transparent crossings since)
// Guided mode: bits 8.. of the flags count the path's guided vertices so far.
// Only the first kMaxGuided scattering vertices of a path are guided: a
// vertex's weight f / p_mix can exceed the plain strategy's by up to
// 1 / (1 - alpha) where the mixture misses, and over a long random walk in a
// medium such factors compound into heavy-tailed throughputs.
constexpr uint32_t kGuidedShift = 8;
#ifndef RX_GUIDE_MAX_VERTICES
#define RX_GUIDE_MAX_VERTICES 3
#endif
constexpr uint32_t kMaxGuided = RX_GUIDE_MAX_VERTICES;

struct DevObject {
  int32_t type, vol;
  float c1[3], c2[3];
  float expo, ior, g;
  float st[3];     // sigma_t = sigma_s + sigma_a
  float lum_st;    // y(sigma_t)
  float albedo;    // y(sigma_s) / y(sigma_t)
  float max1, max2;
};

struct DevScene {
  DevTree tree;
  const DevObject* objs;
  int32_t n_objs;
  const DevTree* vol;
  const int32_t* vol_obj;
  int32_t n_vol;
  const float4* etri;  // 3 per emitter triangle; etri[3k].w = object id bits
  const float4* enrm;
  const float* ecdf;   // normalised cumulative area
  int32_t n_etri;
  float total_area;
};

struct DevCam {
  float pos[3], front[3], right[3], up[3];
  float sight;
  int32_t W, H;
};

struct PathBufs {
  float4 *o, *d, *beta, *lastp, *L;
  int4* s;     // inside, contact, contact position, flags
  int2* hit;   // dist bits, leaf position
  uint32_t *q0, *q1, *sq, *cnt;
  float4 *so, *sd, *sc;
  int2* ss;
  double *sum, *sumsq;
  unsigned long long* stats;  // 0 rays, 1 extend rays, 2 shadow rays, 3 segments, 4 boxes, 5 tris, 6 guide rows
  uint2* spill;
  int32_t spill_depth;
  int32_t* overflow;
};

struct PassArgs {
  uint32_t n_paths, sample, key1, max_depth, rr_start, mode;
  uint32_t npix;  // pixels
  uint32_t spass; // samples of every pixel in this pass
};

#ifndef RX_PIXEL_MAJOR
#define RX_PIXEL_MAJOR 1
#endif
// Path slot of (pixel q, j-th sample of the pass).  Pixel-major: the pass's
// samples of one pixel occupy adjacent slots, so a wave's camera rays (and
// the first bounces after them) start together and share most of their
// traversal.  Slot order changes no result: the random streams are keyed by
// (pixel, sample) and the film adds a pixel's samples in sample order.
__host__ __device__ __forceinline__ uint32_t pass_slot(const PassArgs& P, uint32_t q, uint32_t j) {
  return RX_PIXEL_MAJOR ? q * P.spass + j : j * P.npix + q;
}
__device__ __forceinline__ void slot_pixel(const PassArgs& P, uint32_t p, uint32_t& pix, uint32_t& smp) {
  if (RX_PIXEL_MAJOR) {
    pix = p / P.spass;
    smp = P.sample + (p - pix * P.spass);
  } else {
    pix = p % P.npix;
    smp = P.sample + p / P.npix;
  }
}

// Multi-scene batch: the scenes' trees, objects and emitters are concatenated
// (node, leaf, triangle and object indices rebased), so one DevScene holds
// them all; each scene keeps its root, emitter range, camera and parameters.
struct SceneInfo {
  int32_t root_ref, e_base, n_etri;
  float root_lo[3], root_hi[3], total_area;
  uint32_t key1, max_depth, rr_start, film_base;
  DevCam cam;
};
// The pass schedule of a batch lives on the device: pass `*pass_id` renders
// segments passes[pass].x .. +passes[pass].y (scene, first pixel, sample,
// first path slot), passes[pass].z paths; one sample of a pixel per pass.
struct BatchArgs {
  const SceneInfo* info;
  const uint4* segs;
  const uint4* passes;
  uint32_t* pass_id;
  uint32_t *scn, *pix, *smp;  // per path slot
};

// Per-path constants: pixel, sample and the scene's key and depth limits; in
// a batch the scene's emitters and root replace the shared ones in S.
struct PathCtx {
  uint32_t pix, smp, key1, max_depth, rr_start;
};
template <bool BATCH>
__device__ __forceinline__ PathCtx path_ctx(const PassArgs& P, const BatchArgs& BA, uint32_t p, DevScene& S) {
  if (BATCH) {
    const SceneInfo& I = BA.info[BA.scn[p]];
    S.ecdf += I.e_base;
    S.etri += 3 * I.e_base;
    S.enrm += I.e_base;
    S.n_etri = I.n_etri;
    S.total_area = I.total_area;
    S.tree.root_ref = I.root_ref;
    for (int a = 0; a < 3; a++) {
      S.tree.root_lo[a] = I.root_lo[a];
      S.tree.root_hi[a] = I.root_hi[a];
    }
    return PathCtx{BA.pix[p], BA.smp[p], I.key1, I.max_depth, I.rr_start};
  }
  PathCtx c{0u, 0u, P.key1, P.max_depth, P.rr_start};
  slot_pixel(P, p, c.pix, c.smp);
  return c;
}

// Guided mode: the decoded network output of each queue entry of the current
// bounce (kGHead bytes per network row) and the training
// records of the current pass (16 floats per (path, bounce) slot: x, -d, w,
// p_mix(w), p_bsdf(w), lum L before the vertex, lum throughput arriving at
// the vertex, RR probability, lum of the vertex's own NEE, tag).
struct GuideArgs {
  GuideNet net;
  float4* query;  // per network row: vertex position, direction to the viewer
  unsigned char* head;  // per network row: decoded output (kGHead bytes)
  uint32_t* row;  // network row of each queue entry of the bounce (kNoRow: not a guided vertex)
  float* rec;
  uint32_t R;
  float tag;  // > 0: record this pass with that tag
};
constexpr uint32_t kNoRow = 0xffffffffu;
constexpr int kRecFloats = 16;

// ------------------------------------------------------------------ helpers
__device__ __forceinline__ float lum(Vec3 v) { return 0.212671f * v.x + 0.715160f * v.y + 0.072169f * v.z; }
__device__ __forceinline__ Vec3 ld3(const float* c) { return v3(c[0], c[1], c[2]); }
__device__ __forceinline__ float maxc(Vec3 c) { return fmaxf(c.x, fmaxf(c.y, c.z)); }
__device__ __forceinline__ Vec3 vexp(Vec3 a) { return v3(__expf(a.x), __expf(a.y), __expf(a.z)); }

struct Frame {
  Vec3 s, t, n;
};
// LocalFrame::buildFromNormal (include/LocalFrame.h:13-24); n is the local y axis
__device__ __forceinline__ Frame frame_from(Vec3 n) {
  Frame f;
  f.n = n;
  Vec3 tmp = fabsf(n.z) > 0.99f ? v3(1, 0, 0) : v3(0, 0, 1);
  f.s = normalize(cross(n, tmp));
  f.t = normalize(cross(f.s, n));
  return f;
}
__device__ __forceinline__ Vec3 to_world(const Frame& f, Vec3 a) { return f.s * a.x + f.n * a.y + f.t * a.z; }

// genHemiCosDirection (src/RandGenerator.cpp:152-166): cos(theta) = u1^(1/(n+1))
__device__ __forceinline__ Vec3 sample_lobe(const Frame& f, float u0, float u1, float expo) {
  float phi = 6.28318530717958647692f * u0;
  float c = expo == 1.0f ? sqrtf(u1) : powf(u1, 1.0f / (expo + 1.0f));
  float s = sqrtf(fmaxf(0.0f, 1.0f - c * c));
  float sp, cp;
  __sincosf(phi, &sp, &cp);
  return normalize(to_world(f, v3(cp * s, c, sp * s)));
}
// CosineSphericalSampler::getProbDensity (include/CosineSphericalSampler.h:29-34)
__device__ __forceinline__ float lobe_pdf(Vec3 axis, Vec3 w, float expo) {
  float c = fmaxf(dot(axis, w), 0.0f);
  float pc = expo == 1.0f ? c : powf(c, expo);
  return pc * (expo + 1.0f) * 0.15915494309189533577f;
}
__device__ __forceinline__ Vec3 mirror(Vec3 n, Vec3 d) { return (-dot(n, d)) * n * 2.0f + d; }

// HGPhaseFunc::evaluate / HGPhaseSampler::getProbDensity (include/HGPhaseFunc.h:11-17)
__device__ __forceinline__ float hg(float g, float cosTheta) {
  float temp = 1.0f + g * g - 2.0f * g * cosTheta;
  return 0.07957747154594767f * (1.0f - g * g) / (temp * sqrtf(temp));
}

__device__ __forceinline__ float refr_coeff(const DevScene& S, int o) {
  if (o < 0) return 1.0f;
  int t = S.objs[o].type;
  return (t == OBJ_REFRACTIVE || t == OBJ_VPM) ? S.objs[o].ior : 1.0f;
}

// Ray::getRadianceDecay over a segment inside object `ins`
__device__ __forceinline__ Vec3 decay(const DevScene& S, int ins, float dist) {
  if (ins < 0) return v3(1, 1, 1);
  const DevObject& O = S.objs[ins];
  if (O.type == OBJ_VPM) return vexp(v3(-O.st[0] * dist, -O.st[1] * dist, -O.st[2] * dist));
  if (O.type == OBJ_REFRACTIVE) return v3(powf(O.c2[0], dist), powf(O.c2[1], dist), powf(O.c2[2], dist));
  return v3(1, 1, 1);
}
// SceneVPMObject::P_surface / p_medium (include/SceneVPMObject.h:34-47)
__device__ __forceinline__ float P_surface(const DevObject& O, float d) {
  return lum(vexp(v3(-O.st[0] * d, -O.st[1] * d, -O.st[2] * d)));
}
__device__ __forceinline__ float p_medium(const DevObject& O, float s) {
  Vec3 e = vexp(v3(-O.st[0] * s, -O.st[1] * s, -O.st[2] * s));
  return lum(v3(O.st[0] * e.x, O.st[1] * e.y, O.st[2] * e.z));
}

// Dielectric boundary (the Snell/Fresnel step shared by RefractiveMaterial and
// SceneVPMObject, src/RefractiveMaterial.cpp:189-232): returns false on total
// internal reflection; t = refracted direction, er = unpolarised reflectance.
__device__ __forceinline__ bool refract(Vec3 d, Vec3 n, float n_i, float n_t, Vec3& t, float& er) {
  if (n_i == n_t) {  // index-matched: no bend, no reflection
    t = d;
    er = 0.0f;
    return true;
  }
  float cd = dot(d, n);
  Vec3 nf = cd < 0 ? n : -n;
  float ci = fminf(fabsf(cd), 1.0f);
  float eta = n_i / n_t;
  float s2 = eta * eta * (1.0f - ci * ci);
  if (s2 >= 1.0f) return false;
  float ct = sqrtf(1.0f - s2);
  t = normalize(d * eta + nf * (eta * ci - ct));
  float rs = (n_i * ci - n_t * ct) / (n_i * ci + n_t * ct);
  float rp = (n_t * ci - n_i * ct) / (n_t * ci + n_i * ct);
  er = 0.5f * (rs * rs + rp * rp);
  return true;
}

// Queue appends staged in LDS: the entries a block produces over a chunk of
// its input go to an LDS list (one LDS atomic per wave), and the block then
// claims their slots in the global queue with one atomic for the whole chunk.
// Appends from every wave of the grid otherwise meet at one counter address,
// where they queue one after another.  The order inside a chunk is the LDS
// atomics' order, so neighbouring entries stay neighbours.
constexpr uint32_t kStageIters = 8;                  // input entries per thread per chunk
constexpr uint32_t kStageCap = kRBlock * kStageIters;  // chunk size and LDS list capacity
struct StagedQueue {
  uint32_t* list;   // LDS, kStageCap entries
  uint32_t* count;  // LDS: [0] entries staged, [1] global base
};
__device__ __forceinline__ void staged_reset(StagedQueue q) {
  if (threadIdx.x == 0) q.count[0] = 0;
}
__device__ __forceinline__ void staged_push(StagedQueue q, bool pred, uint32_t v) {
  const unsigned long long m = __ballot(pred);
  const int lane = threadIdx.x & 63;
  const int leader = m ? (__ffsll((long long)m) - 1) : 0;
  uint32_t base = 0;
  if (m && lane == leader) base = atomicAdd(q.count, (uint32_t)__popcll(m));
  base = __shfl(base, leader);
  if (pred) q.list[base + (uint32_t)__popcll(m & ((1ull << lane) - 1ull))] = v;
}
// After a __syncthreads: one thread claims the global slots.
__device__ __forceinline__ void staged_claim(StagedQueue q, uint32_t* counter) {
  if (threadIdx.x == 0) q.count[1] = q.count[0] ? atomicAdd(counter, q.count[0]) : 0u;
}
// After another __syncthreads: copy the list out.
__device__ __forceinline__ void staged_flush(StagedQueue q, uint32_t* out) {
  const uint32_t nq = q.count[0], base = q.count[1];
  for (uint32_t k = threadIdx.x; k < nq; k += blockDim.x) out[base + k] = q.list[k];
}
__device__ __forceinline__ void wave_add(unsigned long long* dst, uint32_t v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
  if ((threadIdx.x & 63) == 0 && v) atomicAdd(dst, (unsigned long long)v);
}

// Scene::findInsideObject (src/Scene.cpp:354-390): the volumetric object whose
// own tree is hit both backwards and forwards through outward-facing faces.
__device__ int find_inside(const DevScene& S, Vec3 o, Vec3 d, int current, uint2* lds, const PathBufs& B,
                           uint32_t& rays) {
  uint2* spill = B.spill + (size_t)(blockIdx.x * blockDim.x + threadIdx.x) * B.spill_depth;
  bool ovf = false;
  TraceStats st;
  for (int v = S.n_vol - 1; v >= 0; v--) {
    const int obj = S.vol_obj[v];
    if (obj == current) continue;
    const DevTree& T = S.vol[v];
    Hit hb = trace<false, false, false, kLdsDepth>(T, o, -d, NONE, NONE, lds, spill, B.spill_depth, &st, &ovf);
    rays++;
    if (!(hb.dist >= 1e-6f)) continue;
    if (dot(xyz(T.tn[hb.pos]), -d) < 1e-6f) continue;
    Hit hf = trace<false, false, false, kLdsDepth>(T, o, d, NONE, NONE, lds, spill, B.spill_depth, &st, &ovf);
    rays++;
    if (!(hf.dist >= 1e-6f)) continue;
    if (dot(xyz(T.tn[hf.pos]), d) < 1e-6f) continue;
    return obj;
  }
  if (ovf) *B.overflow = 1;
  return NONE;
}

__device__ __forceinline__ float4 f4(Vec3 v, float w) { return make_float4(v.x, v.y, v.z, w); }

// ------------------------------------------------------------------ generate
// Camera::generateRay / generateRayBase (src/Camera.cpp:38-47,4-36); the
// camera cosine and the vignetting division of src/PathTracer.cpp:55,67
// (Camera::eliminateVignetting, src/Camera.cpp:78-96) start the throughput.
template <bool BATCH, bool VOL>
__device__ __forceinline__ void generate_body(DevScene S, DevCam C0, PathBufs B, PassArgs P, BatchArgs BA,
                                              uint2* stack) {
  const uint32_t stride = gridDim.x * blockDim.x;
  uint32_t rays = 0;
  uint32_t n_paths = P.n_paths;
  uint4 pass = make_uint4(0, 0, 0, 0);
  if (BATCH) {
    pass = BA.passes[*BA.pass_id];
    n_paths = pass.z;
  }
  if (blockIdx.x == 0)
    for (int k = threadIdx.x; k < kCntWords; k += blockDim.x) B.cnt[k] = (k == 0) ? n_paths : 0u;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n_paths; i += stride) {
    uint32_t pix, smp, key1;
    DevCam C;
    if (BATCH) {  // the pass segment holding slot i: last one starting at or before it
      uint32_t lo = pass.x, hi = pass.x + pass.y - 1;
      while (lo < hi) {
        const uint32_t mid = (lo + hi + 1) >> 1;
        if (BA.segs[mid].w <= i) lo = mid;
        else hi = mid - 1;
      }
      const uint4 sg = BA.segs[lo];
      const SceneInfo& I = BA.info[sg.x];
      pix = sg.y + (i - sg.w);
      smp = sg.z;
      key1 = I.key1;
      C = I.cam;
      BA.scn[i] = sg.x;
      BA.pix[i] = pix;
      BA.smp[i] = smp;
    } else {
      slot_pixel(P, i, pix, smp);
      key1 = P.key1;
      C = C0;
    }
    const Vec3 front = ld3(C.front), right = ld3(C.right), up = ld3(C.up), pos = ld3(C.pos);
    const int x = (int)(pix % (uint32_t)C.W), y = (int)(pix / (uint32_t)C.W);
    U4 r = philox(U4{smp, 0xFFFFu, 0u, 0u}, pix, key1);
    float xp = ((float)x + u01(r.x)) - (float)C.W * 0.5f;
    float yp = (float)C.H * 0.5f - ((float)y + u01(r.y));
    Vec3 d = normalize(front * C.sight + right * xp + up * yp);
    Vec3 dc = normalize(front * C.sight + right * ((float)x + 0.5f - (float)C.W * 0.5f) +
                        up * ((float)C.H * 0.5f - 0.5f - (float)y));
    float beta = fabsf(dot(front, d)) / fabsf(dot(dc, front));
    int inside = (VOL && !BATCH && S.n_vol) ? find_inside(S, pos, d, NONE, stack + threadIdx.x, B, rays) : NONE;
    B.o[i] = f4(pos, 0.0f);
    B.d[i] = f4(d, 0.0f);
    B.s[i] = make_int4(inside, CAMERA, -1, 0);
    B.beta[i] = make_float4(beta, beta, beta, 0.0f);
    B.L[i] = make_float4(0, 0, 0, 0);
    B.q0[i] = i;
  }
  wave_add(&B.stats[0], rays);
}

template <bool BATCH, bool VOL>
__global__ void __launch_bounds__(kRBlock) k_generate(DevScene S, DevCam C0, PathBufs B, PassArgs P, BatchArgs BA) {
  __shared__ uint2 stack[kLdsDepth * kRBlock];
  generate_body<BATCH, VOL>(S, C0, B, P, BA, stack);
}

// ------------------------------------------------------------------ extend
// Nearest hit of every queued path: the waves take chunks of the queue from
// the bounce's work counter (trace_range); every ray's result is independent
// of the lane that ran it.
// DEPTH: LDS stack entries per lane; BLOCK: threads per block (the launch
// shapes the autotuner chooses between, see launch_extend).
template <bool LEG, bool STATS, bool BATCH, int DEPTH>
__device__ __forceinline__ void extend_body(DevTree T, PathBufs B, uint32_t b, BatchArgs BA, uint2* stack) {
  const uint32_t n = B.cnt[b];
  const uint32_t* q = (b & 1) ? B.q1 : B.q0;
  const uint32_t gtid = blockIdx.x * blockDim.x + threadIdx.x;
  uint32_t boxes = 0, tris = 0;
  bool ovf = false;
  trace_range<false, LEG, STATS, DEPTH>(
      T, n, B.cnt + kWorkBase + b * kWorkWords, stack + threadIdx.x, B.spill + (size_t)gtid * B.spill_depth, B.spill_depth, &ovf,
      [&](uint32_t i, Vec3& o, Vec3& d, int32_t& ins, int32_t& con, TreeRoot& root) {
        const uint32_t p = q[i];
        o = xyz(B.o[p]);
        d = xyz(B.d[p]);
        if (LEG) {
          const int4 s = B.s[p];
          ins = s.x;
          con = s.y;
        }
        if (BATCH) {
          const SceneInfo& I = BA.info[BA.scn[p]];
          root.ref = I.root_ref;
          for (int a = 0; a < 3; a++) {
            root.lo[a] = I.root_lo[a];
            root.hi[a] = I.root_hi[a];
          }
        }
      },
      [&](uint32_t i, Hit h, const TraceStats& st) {
        B.hit[q[i]] = make_int2(__float_as_int(h.dist), h.pos);
        if (STATS) {
          boxes += st.boxes;
          tris += st.tris;
        }
      });
  if (ovf) *B.overflow = 1;
  if (STATS) {
    wave_add(&B.stats[4], boxes);
    wave_add(&B.stats[5], tris);
  }
  if (gtid == 0) {
    atomicAdd(&B.stats[0], (unsigned long long)n);
    atomicAdd(&B.stats[1], (unsigned long long)n);
  }
}

template <bool LEG, bool STATS, bool BATCH, int DEPTH = kLdsDepth, int BLOCK = kRBlock>
__global__ void __launch_bounds__(BLOCK, kExtendMinWaves) k_extend(DevTree T, PathBufs B, uint32_t b, BatchArgs BA) {
  __shared__ uint2 stack[DEPTH * BLOCK];
  extend_body<LEG, STATS, BATCH, DEPTH>(T, B, b, BA, stack);
}

// Camera rays (bounce 0, no condition) as packets of 64: a wave traces the
// 64 neighbouring queue entries of its chunk (with pixel-major slots, the
// samples of the same pixels; in a batch, neighbouring pixels of a scene) as
// one packet (trace_packet), its first chunk by its wave index, the others
// from the bounce's work counter.  BATCH: the lanes of a chunk that belong
// to different scenes (at a segment boundary) form one packet per scene,
// each from its scene's root.  The hits are the per-lane kernel's.
constexpr int kPacketDepth = 64;  // wave stack entries: trees up to 64 levels
template <bool BATCH>
__global__ void __launch_bounds__(kRBlock, kExtendMinWaves)
    k_extend_packet(DevTree T, PathBufs B, uint32_t b, BatchArgs BA) {
  __shared__ uint4 wstack[kRBlock / 64][kPacketDepth];
  const uint32_t n = B.cnt[b];
  const uint32_t* q = (b & 1) ? B.q1 : B.q0;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const uint32_t waves = gridDim.x * (blockDim.x >> 6);
  uint32_t* ctr = B.cnt + kWorkBase + b * kWorkWords;
  uint32_t base = (blockIdx.x * (blockDim.x >> 6) + (uint32_t)wv) * 64u;
  bool ovf = false;
  while (base < n) {  // wave-uniform
    const uint32_t i = base + (uint32_t)lane;
    const bool valid = i < n;
    const uint32_t p = valid ? q[i] : 0u;
    Vec3 o = v3(0, 0, 0), d = v3(0, 0, 1);
    uint32_t scn = 0;
    if (valid) {
      o = xyz(B.o[p]);
      d = xyz(B.d[p]);
      if (BATCH) scn = BA.scn[p];
    }
    bool pending = valid;
    for (unsigned long long left = __ballot(pending); left; left = __ballot(pending)) {  // a packet per scene
      const uint32_t cur = BATCH ? (uint32_t)__shfl((int)scn, __ffsll((long long)left) - 1) : 0u;
      const bool mine = pending && (!BATCH || scn == cur);
      TreeRoot root = tree_root(T);
      if (BATCH) {
        const SceneInfo& I = BA.info[cur];
        root.ref = I.root_ref;
        for (int a = 0; a < 3; a++) {
          root.lo[a] = I.root_lo[a];
          root.hi[a] = I.root_hi[a];
        }
      }
      TraceState ts;
      trace_packet<false, false>(T, root, mine, o, d, -1, kNoCondition, wstack[wv], kPacketDepth, ts, nullptr,
                                 &ovf);
      if (mine) {
        B.hit[p] = ts.has ? make_int2(__float_as_int(ts.best), ts.bestPos) : make_int2(__float_as_int(-1.0f), -1);
        pending = false;
      }
    }
    uint32_t next = 0;
    if (lane == 0) next = waves * 64u + atomicAdd(ctr, 64u);
    base = __builtin_amdgcn_readfirstlane(next);
  }
  if (ovf) *B.overflow = 1;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    atomicAdd(&B.stats[0], (unsigned long long)n);
    atomicAdd(&B.stats[1], (unsigned long long)n);
  }
}

// Launch shapes of the extend kernel: (LDS stack depth, block size) pairs
// with an instantiation.  Every shape gives the same hits; the grid keeps
// the thread count of the default shape (blocks * kRBlock threads), which
// the spill buffer is sized for.
struct ExtendShape {
  int depth, block;
};
constexpr ExtendShape kExtendShapes[] = {{kLdsDepth, kRBlock}, {6, 128}, {8, 128}, {12, 128}, {16, 128},
                                         {kLdsDepth, 64}, {kLdsDepth, 256}};
inline bool extend_shape_ok(uint32_t depth, uint32_t block) {
  for (const ExtendShape& e : kExtendShapes)
    if ((uint32_t)e.depth == (depth ? depth : (uint32_t)kLdsDepth) && (uint32_t)e.block == (block ? block : (uint32_t)kRBlock))
      return true;
  return false;
}
template <bool LEG, bool STATS, bool BATCH>
inline void launch_extend(uint32_t depth, uint32_t block, int blocks, hipStream_t s, const DevTree& T,
                          const PathBufs& B, uint32_t b, const BatchArgs& BA) {
  const int d = depth ? (int)depth : kLdsDepth, k = block ? (int)block : kRBlock;
  const int threads = blocks * kRBlock;
#define RX_EXTEND_SHAPE(D, K)                                                                                  \
  if (d == D && k == K) {                                                                                      \
    hipLaunchKernelGGL((k_extend<LEG, STATS, BATCH, D, K>), dim3(std::max(1, threads / K)), dim3(K), 0, s, T, B, \
                       b, BA);                                                                                 \
    return;                                                                                                    \
  }
  if constexpr (!STATS) {  // the counting renders are untimed: default shape only
    RX_EXTEND_SHAPE(6, 128)
    RX_EXTEND_SHAPE(8, 128)
    RX_EXTEND_SHAPE(12, 128)
    RX_EXTEND_SHAPE(16, 128)
    RX_EXTEND_SHAPE(kLdsDepth, 64)
    RX_EXTEND_SHAPE(kLdsDepth, 256)
  }
#undef RX_EXTEND_SHAPE
  hipLaunchKernelGGL((k_extend<LEG, STATS, BATCH>), dim3(blocks), dim3(kRBlock), 0, s, T, B, b, BA);
}

// ------------------------------------------------------------------ guide queries
// Encode (hit point, view direction) of every queued path of bounce b; row i
// of the network batch belongs to queue entry i.
// Whether the shade kernel's scatter at this hit will be a guided scatter (a
// diffuse, glossy or Phong surface, or a scattering event inside a medium):
// the object resolution of its level loop (media, delegation from a
// refractive inside object) without the scatter itself.  tv receives the
// vertex's distance along the ray (the hit, or the medium's sampled distance).
template <bool VOL>
__device__ __forceinline__ bool guided_vertex(const DevScene& S, int inside, int hitObj, int pos, float t, Vec3 d,
                                              Vec3 hn, float u5, uint32_t flags, float& tv) {
  int cur = inside != NONE ? inside : hitObj;
  tv = t;
  if (cur == NONE || (flags >> kGuidedShift) >= kMaxGuided) return false;
  for (int level = 0; level < 3; level++) {
    const DevObject& O = S.objs[cur];
    if (VOL && O.type == OBJ_VPM) {
      const float sampDist = -__logf(1.0f - u5) / O.lum_st;
      const bool be_in = (inside == cur);
      const bool out_of_vol = !(pos >= 0) || sampDist >= t;
      if (be_in && out_of_vol && pos >= 0 && hitObj != cur) {  // a surface inside the medium
        cur = hitObj;
        continue;
      }
      if (be_in && !out_of_vol) {  // a scattering event inside the medium
        tv = sampDist;
        return true;
      }
      return false;  // boundary crossing or nothing
    }
    if (O.type == OBJ_REFRACTIVE && hitObj != cur && inside == cur) {
      if (pos < 0) return false;
      cur = hitObj;
      continue;
    }
    if (pos < 0) return false;
    return dot(d, hn) <= 0 && (O.type == OBJ_DIFFUSE || O.type == OBJ_GLOSSY || O.type == OBJ_PHONG);
  }
  return false;
}

// Guided mode: the network's inputs of path slot p at bounce b, the vertex
// position and the direction it is seen along; true when the vertex is guided.
template <bool VOL>
__device__ __forceinline__ bool guide_input(const DevScene& S, const PathBufs& B, const PassArgs& P, uint32_t b,
                                            uint32_t p, Vec3& x, Vec3& dd) {
  uint32_t pix, smp;
  slot_pixel(P, p, pix, smp);
  const float4 o = B.o[p], d = B.d[p];
  const int2 h = B.hit[p];
  const int pos = h.y;
  const float t = __int_as_float(h.x);
  int hitObj = NONE;
  Vec3 hn = v3(0, 0, 0);
  if (pos >= 0) {
    hitObj = __float_as_int(S.tree.tv1[pos].w);
    hn = xyz(S.tree.tn[pos]);
  }
  dd = xyz(d);
  const float u5 = VOL ? u01(philox(U4{smp, b, 1u, 0u}, pix, P.key1).y) : 0.0f;
  float tv;
  const int4 st = B.s[p];
  const bool g = guided_vertex<VOL>(S, st.x, hitObj, pos, t, dd, hn, u5, (uint32_t)st.w, tv);
  x = xyz(o) + dd * tv;
  return g;
}
// The network's queries of bounce b: the guided vertices of the queue get
// rows [0, cnt[kGuideRowBase + b]) (two float4 each: position, direction to
// the viewer), appended through an LDS list per block; row[i] maps queue
// entry i to its row or kNoRow.
template <bool VOL>
__global__ void __launch_bounds__(kRBlock) k_guide_query(DevScene S, PathBufs B, PassArgs P, GuideArgs GA, uint32_t b) {
  __shared__ uint32_t list[kStageCap], counts[2];
  const StagedQueue qg{list, counts};
  const uint32_t n = B.cnt[b];
  const uint32_t* q = (b & 1) ? B.q1 : B.q0;
  const uint32_t chunk = blockDim.x * min(kStageIters, max(1u, n / (gridDim.x * blockDim.x)));
  for (uint32_t c0 = blockIdx.x * chunk; c0 < n; c0 += gridDim.x * chunk) {
    staged_reset(qg);
    __syncthreads();
    for (uint32_t i0 = c0; i0 < c0 + chunk; i0 += blockDim.x) {
      const uint32_t i = i0 + threadIdx.x;
      bool want = false;
      if (i < n) {
        Vec3 x, dd;
        want = guide_input<VOL>(S, B, P, b, q[i], x, dd);
        if (!want) GA.row[i] = kNoRow;
      }
      staged_push(qg, want, i);
    }
    __syncthreads();
    staged_claim(qg, &B.cnt[kGuideRowBase + b]);
    __syncthreads();
    if (threadIdx.x == 0 && counts[0]) atomicAdd(&B.stats[6], (unsigned long long)counts[0]);
    // the staged entries get rows base + k; their vertices are computed again (cached inputs)
    for (uint32_t k = threadIdx.x; k < counts[0]; k += blockDim.x) {
      const uint32_t i = list[k], r = counts[1] + k;
      Vec3 x, dd;
      guide_input<VOL>(S, B, P, b, q[i], x, dd);
      GA.row[i] = r;
      GA.query[2 * (size_t)r] = f4(x, 0.0f);
      GA.query[2 * (size_t)r + 1] = f4(-dd, 0.0f);
    }
    __syncthreads();
  }
}

// ------------------------------------------------------------------ shade
struct Scat {
  Vec3 x, dir, color;
  float dirProb, originProb;
  int inside, contact, contactPos;
  bool random, emitter, transparent;
  int neeKind;   // 0 none, 1 surface (neeObj), 2 medium (neeObj)
  int neeObj;
  float pdfB;    // density of the continuation direction (RR survival included)
  bool guided;   // sampled by the guide/BSDF one-sample MIS mixture
  float pmix, pb, qrr;
};

// Effective BSDF f (Phong: half the lobe sum, the expectation of the uniform
// lobe pick) and the BSDF-sampling density without Russian roulette.
__device__ __forceinline__ void bsdf_nrr(const DevObject& O, Vec3 n, Vec3 d, Vec3 w, Vec3& f, float& pdf) {
  const bool front = dot(n, d) <= 0 && dot(n, w) >= 0;
  Vec3 fd = v3(0, 0, 0), fg = v3(0, 0, 0);
  float pd = 0.0f, pg = 0.0f;
  if (O.type == OBJ_DIFFUSE || O.type == OBJ_PHONG) {
    Vec3 rho = O.type == OBJ_DIFFUSE ? ld3(O.c1) : ld3(O.c2);
    if (front) fd = rho * 0.31830988618379067154f;
    pd = fmaxf(dot(n, w), 0.0f) * 0.31830988618379067154f;
  }
  if (O.type == OBJ_GLOSSY || O.type == OBJ_PHONG) {
    Vec3 r = normalize(mirror(n, d));
    float rw = dot(r, w);
    if (front && rw > 0) fg = ld3(O.c1) * powf(rw, O.expo);
    pg = lobe_pdf(r, w, O.expo);
  }
  if (O.type == OBJ_PHONG) {
    f = (fd + fg) * 0.5f;
    pdf = 0.5f * (pd + pg);
  } else {
    f = fd + fg;
    pdf = pd + pg;
  }
}
// Russian roulette at a guided vertex: the continuation probability of the
// reference's scatter (src/DiffuseMaterial.cpp:49-70, GlossyMaterial.cpp:31-53);
// for a Phong object its mean over the uniform lobe pick, (max ks + max kd) / 2,
// so guided paths are no longer on average than the plain estimator's (the
// lobe sum max(ks + kd), twice that, made guided paths on the glossy config
// 45 % more expensive to trace for little less variance).
__device__ __forceinline__ float guided_rr(const DevObject& O) {
  if (O.type == OBJ_PHONG) return fminf(1.0f, 0.5f * (O.max1 + O.max2));
  return O.max1;
}

// Guided surface scatter: with probability alpha sample the learned vMF
// mixture, otherwise the reference's BSDF strategy; weight f cos / p_mix with
// p_mix = alpha q + (1 - alpha) p_bsdf (one-sample MIS: unbiased because
// alpha < 1 keeps the BSDF density in the mixture, and alpha <= 0.5 bounds the
// weight by twice the BSDF strategy's).
__device__ void guided_scatter(const DevObject& O, int obj, Vec3 x, Vec3 d, Vec3 n, int hitPos, int insideIn,
                               bool russian, const GuideHead& H, const float* ug, float u_rr, Scat& sc) {
  sc.x = x;
  sc.contact = obj;
  sc.contactPos = hitPos;
  sc.inside = insideIn;
  sc.random = true;
  sc.neeKind = 1;
  sc.neeObj = obj;
  sc.guided = true;
  Vec3 w;
  if (ug[0] < H.alpha) {
    w = guide_sample(H, ug[1], ug[2], ug[3]);
  } else {
    bool glossy = O.type == OBJ_GLOSSY || (O.type == OBJ_PHONG && ug[1] < 0.5f);
    if (glossy) w = sample_lobe(frame_from(normalize(mirror(n, d))), ug[2], ug[3], O.expo);
    else w = sample_lobe(frame_from(n), ug[2], ug[3], 1.0f);
  }
  Vec3 f;
  float pb;
  bsdf_nrr(O, n, d, w, f, pb);
  // The continuation multiplies by the cosine term clamped at 1e-6 (the
  // reference's COS_TERM_MIN); for directions closer to the tangent plane
  // the true cosine is folded into f, so the weight stays f |cos| / p_mix:
  // the BSDF density in p_mix vanishes with the cosine, the clamped term
  // would not, and a guide-proposed grazing direction would get an
  // unbounded weight.
  const float cn = fabsf(dot(n, w));
  if (cn < 1e-6f) f = f * (cn * 1e6f);
  const float q = guide_pdf(H, w);
  const float pmix = H.alpha * q + (1.0f - H.alpha) * pb;
  const float qrr = russian ? guided_rr(O) : 1.0f;
  sc.pmix = pmix;
  sc.pb = pb;
  sc.qrr = qrr;
  if (maxc(f) <= 0.0f || !(pmix > 0.0f) || u_rr >= qrr) {
    sc.dir = v3(0, 0, 0);
    sc.color = v3(0, 0, 0);
    sc.dirProb = 1.0f;
    return;
  }
  sc.dir = w;
  sc.color = f;
  sc.dirProb = pmix * qrr;
  sc.pdfB = pmix * qrr;
}

// Density of the continuation direction w at a surface vertex of object O with
// normal n and incoming direction d (Phong: mixture of its two lobes).
__device__ __forceinline__ void surface_bsdf(const DevObject& O, Vec3 n, Vec3 d, Vec3 w, bool russian, Vec3& f,
                                             float& pdf) {
  const bool front = dot(n, d) <= 0 && dot(n, w) >= 0;
  Vec3 fd = v3(0, 0, 0), fg = v3(0, 0, 0);
  float pd = 0.0f, pg = 0.0f;
  if (O.type == OBJ_DIFFUSE || O.type == OBJ_PHONG) {
    Vec3 rho = O.type == OBJ_DIFFUSE ? ld3(O.c1) : ld3(O.c2);
    float p = russian ? (O.type == OBJ_DIFFUSE ? O.max1 : O.max2) : 1.0f;
    if (front) fd = rho * 0.31830988618379067154f;
    pd = p * fmaxf(dot(n, w), 0.0f) * 0.31830988618379067154f;
  }
  if (O.type == OBJ_GLOSSY || O.type == OBJ_PHONG) {
    Vec3 r = normalize(mirror(n, d));
    float rw = dot(r, w);
    if (front && rw > 0) fg = ld3(O.c1) * powf(rw, O.expo);
    pg = O.max1 * lobe_pdf(r, w, O.expo);
  }
  if (O.type == OBJ_PHONG) {
    f = (fd + fg) * 0.5f;
    pdf = 0.5f * (pd + pg);
  } else {
    f = fd + fg;
    pdf = pd + pg;
  }
}

// One scatter of a surface object (the Material::scatter implementations).
// exit_outside: the object the ray is in once it leaves obj
// (Scene::findInsideObject, computed by the caller when insideIn == obj).
// light: the scatter of a light subpath (fixIsLight), which leaves out the
// refracted radiance's (n_i/n_t)^2 scaling.
__device__ void surface_scatter(const DevScene& S, int obj, Vec3 x, Vec3 d, Vec3 n, int hitPos, int insideIn,
                                bool russian, const float* u, Scat& sc, int exit_outside, bool light = false) {
  const DevObject& O = S.objs[obj];
  sc.x = x;
  sc.contact = obj;
  sc.contactPos = hitPos;
  sc.inside = insideIn;
  sc.dirProb = 1.0f;
  sc.color = v3(1, 1, 1);
  int type = O.type;
  bool diffuse_lobe = type == OBJ_DIFFUSE, glossy_lobe = type == OBJ_GLOSSY;
  if (type == OBJ_PHONG) {  // SceneObject::scatter: uniform material pick (src/SceneObject.cpp:373-380)
    if (u[3] < 0.5f) glossy_lobe = true;
    else diffuse_lobe = true;
  }
  if (diffuse_lobe) {  // DiffuseMaterial::scatter (src/DiffuseMaterial.cpp:4-77)
    Vec3 rho = type == OBJ_DIFFUSE ? ld3(O.c1) : ld3(O.c2);
    if (dot(d, n) > 0) {  // leaving: pass through unchanged
      sc.dir = d;
      sc.random = false;
      return;
    }
    Frame lf = frame_from(n);
    Vec3 w = sample_lobe(lf, u[0], u[1], 1.0f);
    float p = type == OBJ_DIFFUSE ? O.max1 : O.max2;
    sc.random = true;
    sc.neeKind = 1;
    sc.neeObj = obj;
    if (u[2] < p || !russian) {
      sc.dir = w;
      sc.color = (dot(n, d) <= 0 && dot(n, w) >= 0) ? rho * 0.31830988618379067154f : v3(0, 0, 0);
      float pdf = fmaxf(dot(n, w), 0.0f) * 0.31830988618379067154f;
      sc.dirProb = russian ? p * pdf : pdf;
    } else {
      sc.dir = v3(0, 0, 0);
      sc.color = v3(0, 0, 0);
      sc.dirProb = 1 - p;
    }
  } else if (glossy_lobe) {  // GlossyMaterial::scatter (src/GlossyMaterial.cpp:4-57)
    Vec3 r = normalize(mirror(n, d));
    Frame olf = frame_from(r);
    Vec3 w = sample_lobe(olf, u[0], u[1], O.expo);
    Vec3 ks = ld3(O.c1);
    sc.random = true;
    sc.neeKind = 1;
    sc.neeObj = obj;
    float rw = dot(r, w);
    Vec3 col = (dot(n, d) <= 0 && dot(n, w) >= 0 && rw > 0) ? ks * powf(rw, O.expo) : v3(0, 0, 0);
    float pdf = lobe_pdf(r, w, O.expo);
    if (col.x == 0 && col.y == 0 && col.z == 0) {
      sc.dir = v3(0, 0, 0);
      sc.color = col;
      sc.dirProb = 1.0f;
    } else if (u[2] < O.max1) {
      sc.dir = w;
      sc.color = col;
      sc.dirProb = pdf * O.max1;
    } else {
      sc.dir = v3(0, 0, 0);
      sc.color = v3(0, 0, 0);
      sc.dirProb = 1 - O.max1;
    }
  } else if (type == OBJ_REFRACTIVE) {  // RefractiveMaterial::scatter (src/RefractiveMaterial.cpp:132-269)
    sc.random = false;
    int outside = obj;
    if (insideIn == obj) outside = exit_outside;
    float n_i = refr_coeff(S, insideIn), n_t = refr_coeff(S, outside);
    Vec3 t;
    float er;
    if (!refract(d, n, n_i, n_t, t, er)) {  // total internal reflection terminates
      sc.dir = v3(0, 0, 0);
      sc.contact = NONE;
      sc.color = v3(0, 0, 0);
      return;
    }
    float p = fminf(fmaxf(er, 0.0f), 1.0f);
    Vec3 sf = ld3(O.c1);
    if (u[4] < p) {
      Vec3 rd = normalize(mirror(n, d));
      sc.dir = rd;
      sc.color = sf * (er / fmaxf(fabsf(dot(n, rd)), 1e-6f));
      sc.dirProb = p;
    } else {
      sc.dir = t;
      sc.color = (light ? sf : sf * ((n_i * n_i) / (n_t * n_t))) * ((1 - er) / fmaxf(fabsf(dot(n, t)), 1e-6f));
      sc.dirProb = 1 - p;
      sc.inside = outside;
    }
  } else if (type == OBJ_EMISSIVE) {  // SceneEmissiveObject::scatter (src/SceneEmissiveObject.cpp:4-19)
    sc.dir = v3(0, 0, 0);
    sc.inside = NONE;
    sc.emitter = true;
    sc.random = true;
    sc.color = dot(d, n) <= 0 ? ld3(O.c1) : v3(0, 0, 0);
  }
}

// minimum resident waves per SIMD for the plain medium-free shade kernel:
// 5 (96 VGPRs, a 12-byte spill) shades a 64-spp 1024^2 bench render in
// 4.68 ms against 4.79 at 4, 4.82 at 8 and 5.30 at 6 (80 VGPRs, a 68-byte
// spill).  The other variants need far more registers and keep their own
// allocation.
#ifndef RX_SHADE_MIN_WAVES
#define RX_SHADE_MIN_WAVES 5
#endif
// One scatter of samplePathIter (src/MCRenderer.cpp:151-267) after the hit
// of the ray (o, d) is known (t, pos, hitObj, normal hn; inside: the object
// the ray travels in): the object resolution of the level loop (a medium
// the ray is in, a surface inside it, delegation from a refractive object to
// the surface it hits) and that object's scatter.  originScale receives the
// surface-inside-a-medium density the caller multiplies into originProb.
// GUIDED: the guided scatter of guided vertices (row GA.row[i], random
// numbers of (smp, b, pix, key1)), whose mixture head is left in `head`.
// light: the scatter of a light subpath (fixIsLight).
template <bool VOL, bool GUIDED>
__device__ __forceinline__ void scatter_at(const DevScene& S, Vec3 o, Vec3 d, float t, int pos, int hitObj, Vec3 hn,
                                           int inside, bool russian, const float* u, uint2* lds, const PathBufs& B,
                                           uint32_t& rays, bool light, const GuideArgs& GA, uint32_t i, uint32_t smp,
                                           uint32_t b, uint32_t pix, uint32_t key1, Scat& sc, GuideHead& head,
                                           float& originScale) {
  const int obj = inside != NONE ? inside : hitObj;
  sc.dir = v3(0, 0, 0);
  sc.color = v3(1, 1, 1);
  sc.dirProb = 1.0f;
  sc.originProb = 1.0f;
  sc.inside = NONE;
  sc.contact = NONE;
  sc.contactPos = -1;
  sc.random = sc.emitter = sc.transparent = false;
  sc.neeKind = 0;
  sc.neeObj = NONE;
  sc.pdfB = 0.0f;
  sc.guided = false;
  sc.pmix = sc.pb = sc.qrr = 0.0f;
  const Vec3 xh = o + d * t;  // hit point (valid when pos >= 0)
  // the one containment query a scatter can need: leaving the object
  // the path is in through its own boundary (a refractive object, or a
  // medium whose sampled distance passes the boundary)
  int exit_outside = NONE;
  if (VOL && pos >= 0 && inside != NONE && hitObj == inside) {
    const DevObject& I = S.objs[inside];
    if (I.type == OBJ_REFRACTIVE || (I.type == OBJ_VPM && -__logf(1.0f - u[5]) / I.lum_st >= t))
      exit_outside = find_inside(S, xh, d, inside, lds, B, rays);
  }
  int cur = obj;
  originScale = 1.0f;
  bool done = false;
  for (int level = 0; level < 3 && !done; level++) {
    const DevObject& O = S.objs[cur];
    if (VOL && O.type == OBJ_VPM) {  // SceneVPMObject::scatter (src/SceneVPMObject.cpp:4-251)
      const float sampDist = -__logf(1.0f - u[5]) / O.lum_st;
      const bool go_in = (hitObj == cur) && (inside != cur);
      const bool be_in = (inside == cur);
      const bool out_of_vol = !(pos >= 0) || sampDist >= t;
      if (go_in || (be_in && out_of_vol && hitObj == cur)) {
        // cross the own boundary, entering or leaving
        int outside = go_in ? cur : exit_outside;
        float n_i = go_in ? refr_coeff(S, inside) : O.ior, n_t = refr_coeff(S, outside);
        sc.x = xh;
        sc.contact = cur;
        sc.contactPos = pos;
        sc.random = false;
        if (be_in) sc.originProb = P_surface(O, t);
        Vec3 tr;
        float er;
        if (!refract(d, hn, n_i, n_t, tr, er)) {
          sc.dir = v3(0, 0, 0);
          sc.contact = NONE;
          sc.color = v3(0, 0, 0);
        } else {
          float pr = fminf(fmaxf(er, 0.0f), 1.0f);
          if (u[4] < pr) {  // medium boundaries carry no cosine term
            sc.dir = normalize(mirror(hn, d));
            sc.color = v3(er, er, er);
            sc.dirProb = pr;
            sc.inside = inside;
          } else {
            float f = (light ? 1.0f : (n_i * n_i) / (n_t * n_t)) * (1 - er);  // fixIsLight (:84,219)
            sc.dir = tr;
            sc.color = v3(f, f, f);
            sc.dirProb = 1 - pr;
            sc.inside = outside;
            sc.transparent = (n_i == n_t);
          }
        }
        done = true;
      } else if (be_in && !out_of_vol) {  // scatter inside the medium
        sc.x = o + d * sampDist;
        // guided mode: with probability alpha the learned mixture proposes the direction,
        // otherwise the phase function does; one-sample MIS over both densities
        const bool gm = GUIDED && GA.row[i] != kNoRow;
        float ug[4] = {1.0f, 0.0f, 0.0f, 0.0f};
        if (gm) {
          guide_head(GA.head + (size_t)GA.row[i] * kGHead, head);
          U4 r3 = philox(U4{smp, b, 3u, 0u}, pix, key1);
          ug[0] = u01(r3.x);
          ug[1] = u01(r3.y);
          ug[2] = u01(r3.z);
          ug[3] = u01(r3.w);
        }
        const float g = O.g;
        Vec3 w;
        if (gm && ug[0] < head.alpha) {
          w = guide_sample(head, ug[1], ug[2], ug[3]);
        } else {
          Frame lf = frame_from(d);
          float cosT;
          if (fabsf(g) < 1e-5f) cosT = 1 - 2 * u[6];
          else {
            float sq = (1 - g * g) / (1 - g + 2 * g * u[6]);
            cosT = (1 + g * g - sq * sq) / (2 * g);
          }
          float sinT = sqrtf(fmaxf(0.0f, 1 - cosT * cosT));
          float sp, cp;
          __sincosf(6.28318530717958647692f * u[7], &sp, &cp);
          w = normalize(to_world(lf, v3(sinT * cp, cosT, sinT * sp)));
        }
        float ph = hg(g, dot(d, w));
        sc.inside = cur;
        sc.contact = NONE;
        sc.random = true;
        sc.neeKind = 2;
        sc.neeObj = cur;
        sc.originProb = p_medium(O, sampDist);
        float pdir = ph;
        if (gm) {
          pdir = head.alpha * guide_pdf(head, w) + (1.0f - head.alpha) * ph;
          sc.guided = true;
          sc.pmix = pdir;
          sc.pb = ph;
          sc.qrr = russian ? O.albedo : 1.0f;
        }
        if ((u[2] < O.albedo || !russian) && pdir > 0.0f) {
          sc.dir = w;
          sc.color = ld3(O.c1) * ph;
          sc.dirProb = russian ? O.albedo * pdir : pdir;
          sc.pdfB = sc.dirProb;
        } else {
          sc.dir = v3(0, 0, 0);
          sc.color = v3(0, 0, 0);
        }
        done = true;
      } else if (be_in && out_of_vol && pos >= 0) {  // a surface inside the medium
        originScale = P_surface(O, t);
        cur = hitObj;
      } else {  // inside a medium with no hit: nothing to scatter
        sc.dir = v3(0, 0, 0);
        sc.color = v3(0, 0, 0);
        sc.x = o;
        done = true;
      }
    } else if (O.type == OBJ_REFRACTIVE && hitObj != cur && inside == cur) {
      if (pos < 0) {
        sc.x = o;
        done = true;
      } else {
        cur = hitObj;  // RefractiveMaterial delegates to the hit object
      }
    } else {
      if (pos < 0) {
        sc.x = o;
        done = true;
      } else if (GUIDED && dot(d, hn) <= 0 &&
                 (S.objs[cur].type == OBJ_DIFFUSE || S.objs[cur].type == OBJ_GLOSSY ||
                  S.objs[cur].type == OBJ_PHONG) && GA.row[i] != kNoRow) {
        guide_head(GA.head + (size_t)GA.row[i] * kGHead, head);
        U4 r3 = philox(U4{smp, b, 3u, 0u}, pix, key1);
        const float ug[4] = {u01(r3.x), u01(r3.y), u01(r3.z), u01(r3.w)};
        guided_scatter(S.objs[cur], cur, xh, d, hn, pos, inside, russian, head, ug, u[2], sc);
        done = true;
      } else {
        surface_scatter(S, cur, xh, d, hn, pos, inside, russian, u, sc, exit_outside, light);
        done = true;
      }
    }
  }
}

template <int MODE>
struct ShadeLds {
  uint2 stack[kLdsDepth * kRBlock];
  uint32_t cont_list[kStageCap], nee_list[MODE >= 1 ? kStageCap : 1], counts[4];
};
// Bounce b of path p (queue entry i of the bounce): the reference's vertex
// (MCRenderer::samplePathIter :169-201 with PathTracer's accumulation), its
// next-event query (MODE >= 1) and the continuation written back to the path
// state; cont: the path goes on to bounce b + 1, nee: a shadow query was
// written.  shade_body runs it for a bounce's queue, k_tail for the last
// bounces of a path in one lane.
template <int MODE, bool BATCH, bool VOL>
__device__ __forceinline__ void shade_path(const DevScene& S0, const PathBufs& B, const PassArgs& P, uint32_t b,
                                           const GuideArgs& GA, const BatchArgs& BA, uint32_t p, uint32_t i,
                                           uint2* lds, uint32_t& rays, bool& cont, bool& nee) {
  DevScene S = S0;
  const PathCtx pc = path_ctx<BATCH>(P, BA, p, S);
  const uint32_t pix = pc.pix, smp = pc.smp;
  const bool russian = b >= pc.rr_start;
  const float4 o4 = B.o[p], d4 = B.d[p], be = B.beta[p];
  const int4 s = B.s[p];
  const int2 h = B.hit[p];
  const Vec3 o = xyz(o4), d = xyz(d4);
  Vec3 beta = xyz(be);
  const int inside = s.x;
  const float t = __int_as_float(h.x);
  const int pos = h.y;
  U4 r0 = philox(U4{smp, b, 0u, 0u}, pix, pc.key1);
  U4 r1 = philox(U4{smp, b, 1u, 0u}, pix, pc.key1);
  const float u[8] = {u01(r0.x), u01(r0.y), u01(r0.z), u01(r0.w), u01(r1.x), u01(r1.y), u01(r1.z), u01(r1.w)};
  int hitObj = NONE;
  Vec3 hn = v3(0, 0, 0);
  if (pos >= 0) {
    hitObj = __float_as_int(S.tree.tv1[pos].w);
    hn = xyz(S.tree.tn[pos]);
  }
  const int obj = inside != NONE ? inside : hitObj;  // MCRenderer::samplePathIter :169-181
  if (obj != NONE) {
    Scat sc;
    GuideHead head;
    float originScale;
    scatter_at<VOL, MODE == 2>(S, o, d, t, pos, hitObj, hn, inside, russian, u, lds, B, rays, false, GA, i, smp, b,
                               pix, pc.key1, sc, head, originScale);
    sc.originProb *= originScale;
    // segment into the new vertex: decay by the incoming ray's medium, 1/originProb
    const float seg = length(sc.x - o);
    beta = beta * decay(S, inside, seg);
    // vec3 / 0 leaves the vector unchanged in the reference (nvVector.h operator/=)
    if (sc.originProb != 0.0f) beta = beta * (1.0f / sc.originProb);
    if (MODE >= 1 && sc.neeKind && b < pc.max_depth) {
      // next-event estimation to an area light: one light point, MIS with the continuation density.
      // The light point is the path's vertex b + 1 (+ the boundary crossings of the shadow
      // ray), so like a continuation it needs a bounce left.
      U4 r2 = philox(U4{smp, b, 2u, 0u}, pix, pc.key1);
      float us = u01(r2.x), u1 = u01(r2.y), u2 = u01(r2.z);
      int lo = 0, hi = S.n_etri - 1;
      while (lo < hi) {
        int mid = (lo + hi) >> 1;
        if (us < S.ecdf[mid]) hi = mid;
        else lo = mid + 1;
      }
      const float4 e0 = S.etri[3 * lo], e1 = S.etri[3 * lo + 1], e2 = S.etri[3 * lo + 2];
      const Vec3 nL = xyz(S.enrm[lo]);
      float su = sqrtf(u1), bu = 1.0f - su, bv = u2 * su;
      Vec3 y = xyz(e0) * bu + xyz(e1) * bv + xyz(e2) * (1.0f - bu - bv);
      Vec3 wv = y - sc.x;
      float r2l = dot(wv, wv), rl = sqrtf(r2l);
      Vec3 w = wv * (1.0f / rl);
      float cosL = -dot(nL, w);
      if (cosL > 0 && rl > 0) {
        const DevObject& LO = S.objs[__float_as_int(e0.w)];
        float pL = r2l / (cosL * S.total_area);
        Vec3 f;
        float pB, cx = 1.0f;
        if (sc.neeKind == 1 && sc.guided) {
          float pbl;
          bsdf_nrr(S.objs[sc.neeObj], hn, d, w, f, pbl);
          pB = sc.qrr * (head.alpha * guide_pdf(head, w) + (1.0f - head.alpha) * pbl);
          cx = fmaxf(fabsf(dot(hn, w)), 1e-6f);
        } else if (sc.neeKind == 1) {
          surface_bsdf(S.objs[sc.neeObj], hn, d, w, russian, f, pB);
          cx = fmaxf(fabsf(dot(hn, w)), 1e-6f);
        } else {
          const DevObject& M = S.objs[sc.neeObj];
          float ph = hg(M.g, dot(d, w));
          f = ld3(M.c1) * ph;
          pB = (russian ? M.albedo : 1.0f) *
               (sc.guided ? head.alpha * guide_pdf(head, w) + (1.0f - head.alpha) * ph : ph);
        }
        Vec3 c = beta * f * (cx / (pL + pB)) * ld3(LO.c1);
        if (maxc(c) > 0) {
          nee = true;
          B.so[p] = f4(sc.x, rl);
          B.sd[p] = f4(w, 0.0f);
          B.sc[p] = f4(c, 0.0f);
          B.ss[p] = sc.neeKind == 1 ? make_int2(sc.inside, sc.neeObj) : make_int2(sc.neeObj, NONE);
        }
      }
    }
    if (sc.emitter) {
      Vec3 le = sc.dirProb != 0.0f ? sc.color * (1.0f / sc.dirProb) : sc.color;
      float w = 1.0f;
      if (MODE >= 1 && (s.w & FLAG_MIS)) {
        Vec3 lp = xyz(B.lastp[p]);
        Vec3 dv = sc.x - lp;
        float cosL = fabsf(dot(hn, d));
        float pL = dot(dv, dv) / (fmaxf(cosL, 1e-12f) * S.total_area);
        w = be.w / (be.w + pL);
      }
      Vec3 add = beta * le * w;
      float4 L = B.L[p];
      B.L[p] = make_float4(L.x + add.x, L.y + add.y, L.z + add.z, 0.0f);
    } else if (length(sc.dir) >= 0.5f) {
      bool alive = true;
      Vec3 nc = v3(0, 0, 0);
      if (sc.contact != NONE && sc.contactPos >= 0) nc = xyz(S.tree.tn[sc.contactPos]);
      if (sc.inside == NONE && sc.contact != NONE) {  // shading-normal light-leak guard (:183-201)
        float a = dot(nc, -d), c2 = dot(nc, sc.dir);
        if (a * a <= 0 || c2 * c2 <= 0) alive = false;
      }
      if (alive && b + 1 <= pc.max_depth) {
        float cosT = 1.0f;
        if (sc.contact != NONE && S.objs[sc.contact].type != OBJ_VPM)
          cosT = fmaxf(fabsf(dot(nc, sc.dir)), 1e-6f);
        const Vec3 beta_arr = beta;  // throughput arriving at the vertex
        beta = beta * sc.color * (sc.dirProb != 0.0f ? cosT / sc.dirProb : cosT);
        uint32_t flags = (uint32_t)s.w;
        float lastPdf = be.w;
        if (sc.random) {
          flags = FLAG_MIS;
          // the density the direction was drawn with and the vertex, for the MIS
          // weight of an emitter the continuation hits: next-event estimation only
          if (MODE >= 1) {
            if (sc.guided) {
              lastPdf = sc.pdfB;
              if (GA.tag > 0.0f) {  // training record of this vertex
                float* r = GA.rec + ((size_t)p * GA.R + b) * kRecFloats;
                const float4 Lp = B.L[p];
                const float vals[kRecFloats] = {sc.x.x, sc.x.y, sc.x.z, -d.x, -d.y, -d.z, sc.dir.x, sc.dir.y,
                                                sc.dir.z, sc.pmix, sc.pb, lum(v3(Lp.x, Lp.y, Lp.z)), lum(beta_arr),
                                                sc.qrr, 0.0f, GA.tag};
                for (int k = 0; k < kRecFloats; k++) r[k] = vals[k];
              }
            } else if (sc.neeKind == 1) {
              Vec3 fdummy;
              surface_bsdf(S.objs[sc.neeObj], hn, d, sc.dir, russian, fdummy, lastPdf);
            } else {
              lastPdf = sc.dirProb;
            }
            B.lastp[p] = f4(sc.x, 0.0f);
          }
        } else if (!sc.transparent) {
          flags = 0;
        }
        const uint32_t nguided = ((uint32_t)s.w >> kGuidedShift) + (sc.guided ? 1u : 0u);
        flags = (flags & ((1u << kGuidedShift) - 1u)) | (nguided << kGuidedShift);
        B.o[p] = f4(sc.x, 0.0f);
        B.d[p] = f4(sc.dir, 0.0f);
        B.s[p] = make_int4(sc.inside, sc.contact, sc.contactPos, (int)flags);
        B.beta[p] = f4(beta, lastPdf);
        cont = true;
      }
    }
  }
}

template <int MODE, bool BATCH, bool VOL>
__device__ __forceinline__ void shade_body(DevScene S0, PathBufs B, PassArgs P, uint32_t b, GuideArgs GA, BatchArgs BA,
                                           ShadeLds<MODE>& sh) {
  uint2* stack = sh.stack;
  uint32_t *cont_list = sh.cont_list, *nee_list = sh.nee_list, *counts = sh.counts;
  const StagedQueue qc{cont_list, counts}, qn{nee_list, counts + 2};
  const uint32_t n = B.cnt[b];
  const uint32_t* qin = (b & 1) ? B.q1 : B.q0;
  uint32_t* qout = (b & 1) ? B.q0 : B.q1;
  uint2* lds = stack + threadIdx.x;
  uint32_t rays = 0, segs = 0;
  // block-uniform loops: a chunk of up to kStageCap queue entries per block at
  // a time, shorter when the queue does not fill the grid that often
  const uint32_t chunk = blockDim.x * min(kStageIters, max(1u, n / (gridDim.x * blockDim.x)));
  for (uint32_t c0 = blockIdx.x * chunk; c0 < n; c0 += gridDim.x * chunk) {
  staged_reset(qc);
  if (MODE >= 1) staged_reset(qn);
  __syncthreads();
  for (uint32_t i0 = c0; i0 < c0 + chunk; i0 += blockDim.x) {
    const uint32_t i = i0 + threadIdx.x;
    bool cont = false, nee = false;
    uint32_t p = 0;
    if (i < n) {
      p = qin[i];
      segs++;
      shade_path<MODE, BATCH, VOL>(S0, B, P, b, GA, BA, p, i, lds, rays, cont, nee);
    }
    staged_push(qc, cont, p);
    if (MODE >= 1) staged_push(qn, nee, p);
  }
  __syncthreads();
  staged_claim(qc, &B.cnt[b + 1]);
  if (MODE >= 1) staged_claim(qn, &B.cnt[kShadowBase + b]);
  __syncthreads();
  staged_flush(qc, qout);
  if (MODE >= 1) staged_flush(qn, B.sq);
  __syncthreads();
  }
  wave_add(&B.stats[0], rays);
  wave_add(&B.stats[3], segs);
}

template <int MODE, bool BATCH, bool VOL>
__global__ void __launch_bounds__(kRBlock, (MODE == 0 && !VOL) ? RX_SHADE_MIN_WAVES : 1)
    k_shade(DevScene S0, PathBufs B, PassArgs P, uint32_t b, GuideArgs GA, BatchArgs BA) {
  __shared__ ShadeLds<MODE> sh;
  shade_body<MODE, BATCH, VOL>(S0, B, P, b, GA, BA, sh);
}

// ------------------------------------------------------------------ tail
// The last bounces of plain path tracing (RX_MODE_PT, no media, one scene) in
// one launch.  The queue of bounce b0 thins out after Russian roulette, and a
// bounce launch then lasts as long as its slowest ray's chain of dependent
// node fetches whatever the ray count; launched per bounce, those slowest
// rays add up.  Here a lane takes a path of the bounce-b0 queue and carries
// it through its remaining bounces: trace (trip by trip, as the extend
// kernel), then shade_path, then trace its continuation, until the path
// ends; idle lanes take the next path.  Lanes whose trace is done wait until
// kTailShadeMin of the wave are waiting, or none is still tracing, and shade
// together.  Hits, random numbers (keyed by pixel, sample and bounce) and
// the path state are the per-bounce kernels', so the film is bit-identical.
#ifndef RX_TAIL_SHADE_MIN
#define RX_TAIL_SHADE_MIN 32
#endif
constexpr int kTailShadeMin = RX_TAIL_SHADE_MIN;
constexpr uint32_t kDefaultTailBounce = 3;  // rx_render_tuning.tail_bounce default

static __global__ void __launch_bounds__(kRBlock) k_tail(DevScene S, PathBufs B, PassArgs P, uint32_t b0) {
  __shared__ uint2 stack[kLdsDepth * kRBlock];
  const uint32_t n = B.cnt[b0];
  const uint32_t* q = (b0 & 1) ? B.q1 : B.q0;
  const uint32_t gtid = blockIdx.x * blockDim.x + threadIdx.x;
  uint2* const lds = stack + threadIdx.x;
  uint2* const spill = B.spill + (size_t)gtid * B.spill_depth;
  const GuideArgs GA{};
  const BatchArgs BA{};
  bool ovf = false;
  int state = 0;  // 0 idle, 1 tracing bounce b of path p, 2 its hit waits to be shaded
  uint32_t p = 0, b = 0, traced = 0, rays = 0, segs = 0;
  WaveWork w = wave_work(n);
  TraceState ts;
  TraceStats st{0u, 0u};
  auto begin = [&]() {  // bounce b of path p, from the state the last shade wrote
    traced++;
    const int4 s4 = B.s[p];
    if (trace_begin<false>(tree_root(S.tree), ts, xyz(B.o[p]), xyz(B.d[p]), s4.x, s4.y, &st)) return 1;
    B.hit[p] = make_int2(__float_as_int(-1.0f), -1);  // misses the scene box
    return 2;
  };
  for (;;) {
    const unsigned long long idle = __ballot(state == 0);
    if (idle && !w.drained) {
      const uint32_t i = wave_take(w, idle, n, B.cnt + kWorkBase + b0 * kWorkWords);
      if (i < n) {
        p = q[i];
        b = b0;
        state = begin();
      }
    }
    const unsigned long long tracing = __ballot(state == 1), waiting = __ballot(state == 2);
    if (!tracing && !waiting && w.drained) break;
    if (waiting && (!tracing || __popcll(waiting) >= kTailShadeMin)) {
      if (state == 2) {
        bool cont = false, nee = false;
        segs++;
        shade_path<0, false, false>(S, B, P, b, GA, BA, p, 0, lds, rays, cont, nee);
        if (cont) {
          b++;
          state = begin();
        } else {
          state = 0;
        }
      }
    } else if (state == 1 && trace_trip<false, true, false, kLdsDepth>(S.tree, ts, lds, spill, B.spill_depth, &st,
                                                                       &ovf)) {
      B.hit[p] = ts.has ? make_int2(__float_as_int(ts.best), ts.bestPos) : make_int2(__float_as_int(-1.0f), -1);
      state = 2;
    }
  }
  if (ovf) *B.overflow = 1;
  wave_add(&B.stats[0], traced + rays);
  wave_add(&B.stats[1], traced);
  wave_add(&B.stats[3], segs);
}

// ------------------------------------------------------------------ shadow
// Visibility as MCRenderer::testVisibility (src/MCRenderer.cpp:108-120):
// visible when the first legal hit is not nearer than the light point
// (minus EPSILON).  Index-matched medium boundaries are crossed with the
// transmittance the PathTracer's distance sampling realises on average; each
// crossing is a bounce of the equivalent BSDF-sampled path, so a light reached
// after c crossings from the vertex of bounce b counts only when
// b + 1 + c <= max_depth, the paths the PathTracer itself can produce.
template <bool BATCH, bool VOL>
__device__ __forceinline__ void shadow_body(DevScene S0, PathBufs B, PassArgs P, uint32_t b, GuideArgs GA, BatchArgs BA,
                                            uint2* stack) {
  const uint32_t n = B.cnt[kShadowBase + b];
  const uint32_t gtid = blockIdx.x * blockDim.x + threadIdx.x;
  uint2* lds = stack + threadIdx.x;
  uint2* spill = B.spill + (size_t)gtid * B.spill_depth;
  uint32_t rays = 0, srays = 0;
  bool ovf = false, active = false;
  WaveWork ww = wave_work(n);
  // the lane's query: from o along w to the light point at distance `remaining`
  uint32_t p = 0, max_depth = 0;
  int c = 0, inside = NONE, contact = NONE;
  Vec3 o = v3(0, 0, 0), w = v3(0, 0, 0), T = v3(1, 1, 1);
  float remaining = 0.0f;
  DevScene S = S0;
  TraceState ts;
  TraceStats st{0u, 0u};
  auto finish = [&](bool visible) {
    if (VOL && visible && inside != NONE && S.objs[inside].type == OBJ_VPM) {
      const DevObject& M = S.objs[inside];
      Vec3 tr = vexp(v3(-M.st[0] * remaining, -M.st[1] * remaining, -M.st[2] * remaining));
      T = T * tr * (__expf(-M.lum_st * remaining) / lum(tr));
    }
    if (visible) {
      const float4 scv = B.sc[p];
      const float4 L = B.L[p];
      const Vec3 add = v3(scv.x * T.x, scv.y * T.y, scv.z * T.z);
      B.L[p] = make_float4(L.x + add.x, L.y + add.y, L.z + add.z, 0.0f);
      if (GA.tag > 0.0f) {  // the vertex's own NEE is not part of the radiance along its sampled direction
        float* r = GA.rec + ((size_t)p * GA.R + b) * kRecFloats;
        if (r[15] == GA.tag) r[14] += lum(add);
      }
    }
    return false;
  };
  // trace the next straight segment; false when the query is already decided
  auto start = [&]() {
    rays++;
    srays++;
    if (!trace_begin<false>(tree_root(S.tree), ts, o, w, inside, contact, &st))
      return finish(b + 1 + (uint32_t)c <= max_depth);  // no hit at all
    ts.limit = remaining * 1.001f;  // hits beyond the light point cannot block it
    ts.anyhit = remaining;
    // Without media the first blocking hit decides.  With one medium a
    // front-facing hit on any other object (legal under the current
    // condition) decides too: boundary crossings before it only change the
    // condition in ways that keep such a hit legal, so walking the crossings
    // would end at it or earlier.
    ts.use_anyhit = !VOL || S.n_vol <= 1;
    if (VOL && S.n_vol == 1) {
      ts.anyhit_front = true;
      ts.anyhit_skip = S.vol_obj[0];
    }
    return true;
  };
  for (;;) {
    const unsigned long long idle = __ballot(!active);
    if (idle && !ww.drained) {
      const uint32_t i = wave_take(ww, idle, n, B.cnt + kShadowWorkBase + b * kWorkWords);
      {
        if (i < n) {
          p = B.sq[i];
          if (BATCH) S = S0;  // path_ctx rebinds the scene's root and emitters
          max_depth = path_ctx<BATCH>(P, BA, p, S).max_depth;
          const float4 so = B.so[p], sd = B.sd[p];
          const int2 ctx = B.ss[p];
          o = xyz(so);
          w = xyz(sd);
          remaining = so.w;
          inside = ctx.x;
          contact = ctx.y;
          T = v3(1, 1, 1);
          c = 0;
          active = start();
        }
      }
    }
    if (ww.drained && !__any(active)) break;
    if (active && trace_trip<false, true, false, kLdsDepth>(S.tree, ts, lds, spill, B.spill_depth, &st, &ovf)) {
      // MCRenderer::testVisibility: visible when the first legal hit is not nearer than the light point
      if (!ts.has || ts.best - remaining >= -1e-5f) {
        active = finish(b + 1 + (uint32_t)c <= max_depth);
      } else {
        const float hd = ts.best;
        const int hobj = __float_as_int(S.tree.tv1[ts.bestPos].w);
        if (!VOL || S.objs[hobj].type != OBJ_VPM) {
          active = finish(false);
        } else {  // cross an index-matched medium boundary and go on
          const Vec3 xn = o + w * hd;
          const int next_inside = (inside != hobj) ? hobj : find_inside(S, xn, w, hobj, lds, B, rays);
          if (refr_coeff(S, inside) != refr_coeff(S, next_inside)) {  // would bend: not a straight shadow ray
            active = finish(false);
          } else {
            if (inside != NONE && S.objs[inside].type == OBJ_VPM) {
              const DevObject& M = S.objs[inside];
              Vec3 tr = vexp(v3(-M.st[0] * hd, -M.st[1] * hd, -M.st[2] * hd));
              T = T * tr * (__expf(-M.lum_st * hd) / lum(tr));
            }
            o = xn;
            remaining -= hd;
            inside = next_inside;
            contact = hobj;
            c++;
            active = c < kMaxCross ? start() : finish(false);
          }
        }
      }
    }
  }
  if (ovf) *B.overflow = 1;
  wave_add(&B.stats[0], rays);
  wave_add(&B.stats[2], srays);
}

template <bool BATCH, bool VOL>
__global__ void __launch_bounds__(kRBlock, kShadowMinWaves)
    k_shadow(DevScene S0, PathBufs B, PassArgs P, uint32_t b, GuideArgs GA, BatchArgs BA) {
  __shared__ uint2 stack[kLdsDepth * kRBlock];
  shadow_body<BATCH, VOL>(S0, B, P, b, GA, BA, stack);
}

// ------------------------------------------------------------------ training samples
// Deterministic selection of the pass's guided-vertex records: records are
// ranked in (path, bounce) order by a chunked count / scan / write, and an
// evenly strided subset of at most `cap` becomes the training batch with
// weight F/p_sample = lum(L after the vertex) * q_rr / lum(throughput at it).
constexpr int kRecChunk = 1024;

static __global__ void k_rec_count(const float* rec, uint32_t K, float tag, uint32_t* bcount) {
  __shared__ uint32_t red[256];
  uint32_t c = 0;
  const uint32_t base = blockIdx.x * kRecChunk + threadIdx.x * 4;
  for (int j = 0; j < 4; j++) {
    const uint32_t k = base + j;
    if (k < K && rec[(size_t)k * kRecFloats + 15] == tag) c++;
  }
  red[threadIdx.x] = c;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) bcount[blockIdx.x] = red[0];
}

// Exclusive scan of the chunk counts in place, one block of kScanThreads:
// each thread sums a contiguous run of counts, the block scans the run sums,
// and each thread writes its run's offsets.
constexpr int kScanThreads = 1024;
static __global__ void __launch_bounds__(kScanThreads) k_rec_scan(uint32_t* bcount, uint32_t nb, uint32_t* total) {
  __shared__ uint32_t part[kScanThreads];
  const uint32_t t = threadIdx.x;
  const uint32_t per = (nb + kScanThreads - 1) / kScanThreads;
  const uint32_t lo = min(nb, t * per), hi = min(nb, lo + per);
  uint32_t s = 0;
  for (uint32_t i = lo; i < hi; i++) s += bcount[i];
  part[t] = s;
  __syncthreads();
  for (int o = 1; o < kScanThreads; o <<= 1) {  // inclusive scan of the run sums (Hillis-Steele)
    const uint32_t v = t >= (uint32_t)o ? part[t - o] : 0u;
    __syncthreads();
    part[t] += v;
    __syncthreads();
  }
  uint32_t run = part[t] - s;
  for (uint32_t i = lo; i < hi; i++) {
    const uint32_t c = bcount[i];
    bcount[i] = run;
    run += c;
  }
  if (t == kScanThreads - 1) *total = part[t];
}

struct TrainOut {
  float4 *pos, *dir, *wdir;
  float *pb, *wt;
  uint32_t* rows;
};

static __global__ void k_rec_write(const float* rec, uint32_t K, float tag, const uint32_t* boff, const uint32_t* total,
                            uint32_t cap, const float4* L, uint32_t R, TrainOut T) {
  __shared__ uint32_t sc[256];
  uint32_t flags[4];
  uint32_t c = 0;
  const uint32_t base = blockIdx.x * kRecChunk + threadIdx.x * 4;
  for (int j = 0; j < 4; j++) {
    const uint32_t k = base + j;
    flags[j] = (k < K && rec[(size_t)k * kRecFloats + 15] == tag) ? 1u : 0u;
    c += flags[j];
  }
  sc[threadIdx.x] = c;
  __syncthreads();
  for (int o = 1; o < 256; o <<= 1) {  // inclusive scan (Hillis-Steele)
    uint32_t v = threadIdx.x >= (uint32_t)o ? sc[threadIdx.x - o] : 0u;
    __syncthreads();
    sc[threadIdx.x] += v;
    __syncthreads();
  }
  const uint32_t Ttot = *total;
  const uint32_t Bsel = min(Ttot, cap);
  uint32_t g = boff[blockIdx.x] + sc[threadIdx.x] - c;
  if (blockIdx.x == 0 && threadIdx.x == 0) *T.rows = Bsel;
  for (int j = 0; j < 4; j++) {
    if (!flags[j]) continue;
    const uint32_t k = base + j;
    const uint64_t lo = (uint64_t)g * Bsel / Ttot, hi = (uint64_t)(g + 1) * Bsel / Ttot;
    g++;
    if (hi == lo) continue;  // not in the strided subset
    const uint32_t slot = (uint32_t)lo;
    const float* r = rec + (size_t)k * kRecFloats;
    const uint32_t p = k / R;
    const float4 Lp = L[p];
    const float dl = lum(v3(Lp.x, Lp.y, Lp.z)) - r[11] - r[14];
    const float w = (r[12] > 1e-20f && dl > 0.0f) ? dl * r[13] / r[12] : 0.0f;
    T.pos[slot] = make_float4(r[0], r[1], r[2], 0.0f);
    T.dir[slot] = make_float4(r[3], r[4], r[5], 0.0f);
    T.wdir[slot] = make_float4(r[6], r[7], r[8], 0.0f);
    T.pb[slot] = r[10];
    T.wt[slot] = isfinite(w) ? w : 0.0f;
  }
}

// Mini-batch `step` of a training pass: rows r < min(batch, n) take stored
// sample ((step * batch + r) * 0x7fffffff) mod n, a permutation of the n
// stored samples (the prime multiplier is coprime to any n < 2^31), so
// consecutive steps sweep the store in a scattered, deterministic order.
// n is the pass's own samples, min(*total, store_cap), or with kept samples
// (kept > 0) the kept store's first `kept` rows.
static __global__ void k_guide_gather(TrainOut S, const uint32_t* total, uint32_t store_cap, uint32_t kept, uint32_t step,
                                      uint32_t batch, TrainOut D) {
  const uint32_t n = kept ? kept : min(*total, store_cap);
  const uint32_t rows = min(n, batch);
  if (blockIdx.x == 0 && threadIdx.x == 0) *D.rows = rows;
  for (uint32_t r = blockIdx.x * blockDim.x + threadIdx.x; r < rows; r += gridDim.x * blockDim.x) {
    const uint32_t k = (uint32_t)((((uint64_t)step * batch + r) * 0x7fffffffull) % n);
    D.pos[r] = S.pos[k];
    D.dir[r] = S.dir[k];
    D.wdir[r] = S.wdir[k];
    D.pb[r] = S.pb[k];
    D.wt[r] = S.wt[k];
  }
}

// ------------------------------------------------------------------ film
// Adds the pass's samples of each pixel in sample order (slot pass_slot(q, j)
// holds sample `sample + j` of pixel q), so the FP64 film is independent of
// how samples are grouped into passes.
constexpr uint32_t kAccPix = kRBlock;  // pixels per block of k_accumulate, one per thread
constexpr uint32_t kAccChunk = 8;      // samples per pixel staged in LDS at a time
static __global__ void __launch_bounds__(kAccPix) k_accumulate(PathBufs B, PassArgs P) {
  const uint32_t npix = P.npix, S = P.spass;
  if (!RX_PIXEL_MAJOR) {
    for (uint32_t q = blockIdx.x * blockDim.x + threadIdx.x; q < npix; q += gridDim.x * blockDim.x) {
      double s[3], s2[3];
      for (int k = 0; k < 3; k++) {
        s[k] = B.sum[3 * (size_t)q + k];
        s2[k] = B.sumsq[3 * (size_t)q + k];
      }
      for (uint32_t j = 0; j < S; j++) {
        const float4 L = B.L[pass_slot(P, q, j)];
        const double l[3] = {L.x, L.y, L.z};
        for (int k = 0; k < 3; k++) {
          s[k] += l[k];
          s2[k] += l[k] * l[k];
        }
      }
      for (int k = 0; k < 3; k++) {
        B.sum[3 * (size_t)q + k] = s[k];
        B.sumsq[3 * (size_t)q + k] = s2[k];
      }
    }
    return;
  }
  // Pixel-major slots: a block's pixels hold one contiguous run of slots.
  // It is read kAccChunk samples per pixel at a time into LDS, eight
  // neighbouring lanes on one pixel's neighbouring samples (coalesced), and
  // each thread then adds its own pixel's samples in order.
  __shared__ float4 tile[kAccPix * (kAccChunk + 1)];  // one float4 of padding per pixel row
  for (uint32_t q0 = blockIdx.x * kAccPix; q0 < npix; q0 += gridDim.x * kAccPix) {  // block-uniform
    const uint32_t q = q0 + threadIdx.x;
    double s[3] = {0, 0, 0}, s2[3] = {0, 0, 0};
    if (q < npix)
      for (int k = 0; k < 3; k++) {
        s[k] = B.sum[3 * (size_t)q + k];
        s2[k] = B.sumsq[3 * (size_t)q + k];
      }
    for (uint32_t j0 = 0; j0 < S; j0 += kAccChunk) {
      const uint32_t nj = min(kAccChunk, S - j0);
      for (uint32_t e = threadIdx.x; e < kAccPix * kAccChunk; e += blockDim.x) {
        const uint32_t r = e / kAccChunk, t = e - r * kAccChunk;
        if (q0 + r < npix && t < nj) tile[r * (kAccChunk + 1) + t] = B.L[(size_t)(q0 + r) * S + j0 + t];
      }
      __syncthreads();
      if (q < npix)
        for (uint32_t t = 0; t < nj; t++) {
          const float4 L = tile[threadIdx.x * (kAccChunk + 1) + t];
          const double l[3] = {L.x, L.y, L.z};
          for (int k = 0; k < 3; k++) {
            s[k] += l[k];
            s2[k] += l[k] * l[k];
          }
        }
      __syncthreads();
    }
    if (q < npix)
      for (int k = 0; k < 3; k++) {
        B.sum[3 * (size_t)q + k] = s[k];
        B.sumsq[3 * (size_t)q + k] = s2[k];
      }
  }
}

// Batch film: a pass holds at most one sample of any pixel and passes run in
// sample order, so each slot adds to its own pixel and the sums are formed in
// the same order as a single-scene render of that scene.
__device__ __forceinline__ void accumulate_batch_body(PathBufs B, BatchArgs BA) {
  const uint32_t n = BA.passes[*BA.pass_id].z;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const size_t q = (size_t)BA.info[BA.scn[i]].film_base + BA.pix[i];
    const float4 L = B.L[i];
    const double l[3] = {L.x, L.y, L.z};
    for (int k = 0; k < 3; k++) {
      B.sum[3 * q + k] += l[k];
      B.sumsq[3 * q + k] += l[k] * l[k];
    }
  }
}
static __global__ void k_accumulate_batch(PathBufs B, BatchArgs BA) { accumulate_batch_body(B, BA); }
static __global__ void k_pass_set(uint32_t* pass_id, uint32_t v) {
  if (blockIdx.x == 0 && threadIdx.x == 0) *pass_id = v;
}
static __global__ void k_pass_next(uint32_t* pass_id) {
  if (blockIdx.x == 0 && threadIdx.x == 0) *pass_id += 1;
}

static __global__ void k_fill(double* a, size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) a[i] = 0.0;
}
static __global__ void k_fill_u64(unsigned long long* a, int n, int32_t* ovf) {
  if (threadIdx.x < n) a[threadIdx.x] = 0ull;
  if (threadIdx.x == 0) *ovf = 0;
}
static __global__ void k_mean(const double* sum, float* mean, size_t n, double inv) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    mean[i] = (float)(sum[i] * inv);
}

// ------------------------------------------------------------------ host helpers
// Camera basis (src/Camera.cpp:39-43): front = normalize(focus - position),
// right = normalize(front x up), up' = right x front; integer focal length.
inline DevCam make_devcam(const rx_camera_desc* cam) {
  DevCam C;
  float f[3] = {cam->focus[0] - cam->position[0], cam->focus[1] - cam->position[1], cam->focus[2] - cam->position[2]};
  auto nrm = [](float* v) {
    float r = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    if (r > 0)
      for (int k = 0; k < 3; k++) v[k] /= r;
  };
  nrm(f);
  float rt[3] = {f[1] * cam->up[2] - f[2] * cam->up[1], f[2] * cam->up[0] - f[0] * cam->up[2],
                 f[0] * cam->up[1] - f[1] * cam->up[0]};
  nrm(rt);
  float u[3] = {rt[1] * f[2] - rt[2] * f[1], rt[2] * f[0] - rt[0] * f[2], rt[0] * f[1] - rt[1] * f[0]};
  memcpy(C.pos, cam->position, 12);
  memcpy(C.front, f, 12);
  memcpy(C.right, rt, 12);
  memcpy(C.up, u, 12);
  C.sight = (float)cam->sight_dist;
  C.W = cam->width;
  C.H = cam->height;
  return C;
}

// Philox key of a render seed.
inline uint32_t seed_key(uint64_t seed) { return (uint32_t)seed ^ (uint32_t)(seed >> 32) * 0x9E3779B9u; }

// Material constants of one object (sigma_t, y(sigma_t), albedo, RR maxima).
inline DevObject make_devobject(const rx_object_desc& od) {
  DevObject D;
  memset(&D, 0, sizeof(D));
  D.type = od.type;
  D.vol = -1;
  memcpy(D.c1, od.color, 12);
  memcpy(D.c2, od.color2, 12);
  D.expo = od.exponent;
  D.ior = od.ior;
  D.g = od.g;
  for (int k = 0; k < 3; k++) D.st[k] = od.color[k] + od.color2[k];
  auto y = [](const float* v) { return 0.212671f * v[0] + 0.715160f * v[1] + 0.072169f * v[2]; };
  D.lum_st = y(D.st);
  D.albedo = D.lum_st > 0 ? y(od.color) / D.lum_st : 0.0f;
  D.max1 = std::max(od.color[0], std::max(od.color[1], od.color[2]));
  D.max2 = std::max(od.color2[0], std::max(od.color2[1], od.color2[2]));
  return D;
}

// Emitter triangles of one scene with their area CDF (uniform area sampling,
// pdf 1/A_total; src/Scene.cpp:19-33, src/SceneObject.cpp:144-225): appends
// 3 float4 vertices per triangle (w of the first = object id + obj_base), the
// face normal and the normalised cumulative area; returns A_total.
inline double collect_emitters(const float* verts, const uint32_t* tris, uint32_t n_objects,
                               const rx_object_desc* objects, int32_t obj_base, std::vector<float>& etri,
                               std::vector<float>& enrm, std::vector<float>& ecdf) {
  double total = 0.0;
  std::vector<double> areas;
  for (uint32_t o = 0; o < n_objects; o++) {
    if (objects[o].type != RX_OBJ_EMISSIVE) continue;
    for (uint32_t t = 0; t < objects[o].tri_count; t++) {
      uint32_t g = objects[o].tri_begin + t;
      const float* a = verts + 3 * (size_t)tris[3 * (size_t)g];
      const float* bb = verts + 3 * (size_t)tris[3 * (size_t)g + 1];
      const float* cc = verts + 3 * (size_t)tris[3 * (size_t)g + 2];
      double e1[3], e2[3];
      for (int k = 0; k < 3; k++) {
        e1[k] = (double)bb[k] - a[k];
        e2[k] = (double)cc[k] - a[k];
      }
      double cx = e1[1] * e2[2] - e1[2] * e2[1], cy = e1[2] * e2[0] - e1[0] * e2[2], cz = e1[0] * e2[1] - e1[1] * e2[0];
      double area = 0.5 * std::sqrt(cx * cx + cy * cy + cz * cz);
      if (area <= 0) continue;
      int32_t oid = (int32_t)o + obj_base;
      float w0;
      memcpy(&w0, &oid, 4);
      const float* vs[3] = {a, bb, cc};
      for (int j = 0; j < 3; j++) {
        etri.insert(etri.end(), vs[j], vs[j] + 3);
        etri.push_back(j == 0 ? w0 : 0.0f);
      }
      float n[3];
      face_normal(a, bb, cc, n);
      enrm.insert(enrm.end(), n, n + 3);
      enrm.push_back(0.0f);
      total += area;
      areas.push_back(total);
    }
  }
  const size_t first = ecdf.size();
  for (double a : areas) ecdf.push_back((float)(a / total));
  if (ecdf.size() > first) ecdf.back() = 1.0f;
  return total;
}

}  // namespace rx
