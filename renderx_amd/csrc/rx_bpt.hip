// Bidirectional path tracing: BidirectionalPathTracer::renderPixels
// (src/BidirectionalPathTracer.cpp:6-117) with the CPU branch's semantics
// (:36-57): per pixel and sample one eye subpath (Camera::generateRay) and one
// light subpath (Scene::genEmissionSample), both traced by
// MCRenderer::samplePathIter (src/MCRenderer.cpp:151-267), then every
// (light prefix, eye prefix) pair joined by colorByConnectingPaths
// (:265-396) with connectRays (src/MCRenderer.cpp:50-106), the visibility of
// MCRenderer::testVisibility (:108-120), connectColorProb (:133-181) and the
// balance-heuristic weight of connectWeight (:197-263).
//
// One pass renders S samples of every pixel:
//   generate (eye slots [0, N), light slots [N, 2N))
//   -> { extend(b) -> record(b) } for b = 0..max_depth      subpath vertices
//   -> pairs -> visibility                                    one shadow query per connectable pair
//   -> connect                                                one thread per eye subpath
//   -> accumulate                                             own + splatted radiance, in sample order
// The subpaths reuse the path tracer's extend kernel and scatter functions;
// each vertex (the reference's Ray of a Path) is kept in HBM as 64 bytes.
// Light-tracing contributions (eye prefix of one vertex) land on other
// pixels: they are added as 2^-32 fixed point with 64-bit integer atomics,
// so the film stays independent of the order of the additions.
// VOL: scenes with homogeneous media (SceneVPMObject): subpaths scatter in
// and cross the media as the path tracer's paths do, and the connection code
// takes the media's Ray methods: getRadianceDecay (transmittance),
// getOriginSampleProbDensity (p_medium / P_surface, src/SceneVPMObject.cpp:266-273),
// the phase function as BSDF and direction density (:257-285) and no
// cosine term at medium vertices.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "rx_render_state.h"

namespace rx {

constexpr int kBptMaxVerts = 32;                  // vertices per subpath (max_depth + 2): one bit each in a mask
constexpr uint32_t kBptPathSlots = 1u << 20;      // eye + light subpath slots of a pass (default budget)
constexpr uint32_t kBptMaxPixels = 1u << 22;      // pair codes hold the eye slot in 22 bits
constexpr int kBptPairCount = kSpareWords;  // cnt word: connectable pairs of the pass
constexpr int kBptVisWork = kSpareWork;     // cnt block: work counters of the visibility kernel
static_assert(kBptVisWork + (int)kWorkWords <= kCntWords, "BPT counters fit the counter block");
constexpr double kSplatScale = 4294967296.0;      // fixed point of the light-tracing splats

// One subpath vertex: the fields of the reference's Ray that the connection
// code reads (include/Ray.h:8-51), with the contact normal resolved.
struct BVert {
  Vec3 o, d, c, n;  // origin, direction, color, contact normal (0 without contact)
  float dirProb, originProb;
  int contact, inside;
  bool random;  // directionSampleType == RANDOM
};
__device__ __forceinline__ void put_vert(float4* V, size_t idx, const BVert& v) {
  V[4 * idx + 0] = make_float4(v.o.x, v.o.y, v.o.z, v.dirProb);
  V[4 * idx + 1] = make_float4(v.d.x, v.d.y, v.d.z, v.originProb);
  V[4 * idx + 2] = make_float4(v.c.x, v.c.y, v.c.z, __int_as_float(v.contact));
  V[4 * idx + 3] = make_float4(v.n.x, v.n.y, v.n.z, __int_as_float((v.inside + 1) | (v.random ? (1 << 30) : 0)));
}
__device__ __forceinline__ BVert get_vert(const float4* V, size_t idx) {
  const float4 a = V[4 * idx + 0], b = V[4 * idx + 1], c = V[4 * idx + 2], d = V[4 * idx + 3];
  BVert v;
  v.o = xyz(a);
  v.dirProb = a.w;
  v.d = xyz(b);
  v.originProb = b.w;
  v.c = xyz(c);
  v.contact = __float_as_int(c.w);
  v.n = xyz(d);
  const int info = __float_as_int(d.w);
  v.inside = (info & ((1 << 30) - 1)) - 1;
  v.random = (info >> 30) & 1;
  return v;
}
// contact and sample type only (the pair filter)
__device__ __forceinline__ void vert_flags(const float4* V, size_t idx, int& contact, bool& random) {
  contact = __float_as_int(V[4 * idx + 2].w);
  random = (__float_as_int(V[4 * idx + 3].w) >> 30) & 1;
}

struct BptArgs {
  float4* V;         // (2N slots) x K vertices x 4 float4
  uint32_t* nv;      // vertices per subpath
  uint32_t* vis;     // N x K masks: bit k of (p, j) = eye vertex j of slot p sees light vertex k
  uint32_t* pairs;   // connectable pairs: p << 10 | j << 5 | k
  unsigned long long* splat;  // S x npix x 3 fixed-point light-tracing sums
  unsigned long long* own;    // N x 3 fixed-point sums of each eye slot's connected paths
  uint32_t N, K;
  uint32_t registered;  // RX_RENDER_BPT_REGISTERED
};

__device__ __forceinline__ bool is_emissive(const DevScene& S, int o) { return o >= 0 && S.objs[o].type == OBJ_EMISSIVE; }
// nvVector's normalize(): a zero vector stays zero (include/nvVector.h:522-527)
__device__ __forceinline__ Vec3 nrm0(Vec3 v) {
  const float r = sqrtf(dot(v, v));
  return r > 0 ? v3(v.x / r, v.y / r, v.z / r) : v;
}

// ------------------------------------------------------------------ subpaths
// Eye slots p < N: Camera::generateRay (the path tracer's camera ray and
// random numbers).  Light slots N + p: SceneObject::emit(true, false)
// (src/SceneObject.cpp:144-225): a point uniform in area over the emitters,
// a cosine direction about the light normal, colour Le, densities cosine/pi
// and 1 / (emitting area).  Vertex 0 of both is recorded; eye rays are traced
// without a condition (src/Camera.cpp:23), light rays with one (:212-223).
template <bool VOL>
__global__ void __launch_bounds__(kRBlock) k_bpt_generate(DevScene S, DevCam C, PathBufs B, PassArgs P, BptArgs A) {
  __shared__ uint2 stack[VOL ? kLdsDepth * kRBlock : 1];
  uint32_t rays = 0;
  const uint32_t stride = gridDim.x * blockDim.x;
  const uint32_t N = A.N, total = 2 * N;
  if (blockIdx.x == 0)
    for (int k = threadIdx.x; k < kCntWords; k += blockDim.x) B.cnt[k] = (k == 0) ? total : 0u;
  const Vec3 front = ld3(C.front), right = ld3(C.right), up = ld3(C.up), pos = ld3(C.pos);
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < total; i += stride) {
    const bool light = i >= N;
    uint32_t pix, smp;
    slot_pixel(P, light ? i - N : i, pix, smp);
    BVert v;
    v.c = v3(1, 1, 1);
    v.dirProb = 1.0f;
    v.originProb = 1.0f;
    v.random = true;
    if (!light) {
      const int x = (int)(pix % (uint32_t)C.W), y = (int)(pix / (uint32_t)C.W);
      U4 r = philox(U4{smp, 0xFFFFu, 0u, 0u}, pix, P.key1);
      float xp = ((float)x + u01(r.x)) - (float)C.W * 0.5f;
      float yp = (float)C.H * 0.5f - ((float)y + u01(r.y));
      v.o = pos;
      v.d = normalize(front * C.sight + right * xp + up * yp);
      v.contact = CAMERA;
      v.n = front;  // Camera::getWorldNormal
      v.inside = VOL ? find_inside(S, pos, v.d, NONE, stack + threadIdx.x, B, rays) : NONE;
      B.s[i] = make_int4(v.inside, kNoCondition, -1, 0);
    } else {
      U4 r = philox(U4{smp, 0xFFFFu, 0u, 1u}, pix, P.key1);
      int lo = 0, hi = S.n_etri - 1;
      const float us = u01(r.x);
      while (lo < hi) {
        int mid = (lo + hi) >> 1;
        if (us < S.ecdf[mid]) hi = mid;
        else lo = mid + 1;
      }
      const float4 e0 = S.etri[3 * lo], e1 = S.etri[3 * lo + 1], e2 = S.etri[3 * lo + 2];
      const Vec3 nL = xyz(S.enrm[lo]);
      const float su = sqrtf(u01(r.y)), bu = 1.0f - su, bv = u01(r.z) * su;  // SimpleShape::genRandTrianglePosition
      v.o = xyz(e0) * bu + xyz(e1) * bv + xyz(e2) * (1.0f - bu - bv);
      Vec3 w = sample_lobe(frame_from(nL), u01(r.w), u01(philox(U4{smp, 0xFFFFu, 1u, 1u}, pix, P.key1).x), 1.0f);
      if (dot(nL, w) < 0) w = -w;
      const int obj = __float_as_int(e0.w);
      v.d = w;
      v.contact = obj;
      v.n = nL;
      v.c = dot(nL, w) < 0 ? v3(0, 0, 0) : ld3(S.objs[obj].c1);  // SceneEmissiveObject::getBSDF
      v.dirProb = lobe_pdf(nL, w, 1.0f);
      v.originProb = 1.0f / S.total_area;
      v.inside = VOL ? find_inside(S, v.o, w, obj, stack + threadIdx.x, B, rays) : NONE;  // SceneObject::emit (:206)
      B.s[i] = make_int4(v.inside, obj, -1, 0);
    }
    put_vert(A.V, (size_t)i * A.K, v);
    A.nv[i] = 1;
    B.o[i] = f4(v.o, 0.0f);
    B.d[i] = f4(v.d, 0.0f);
    B.L[i] = make_float4(0, 0, 0, 0);
    B.q0[i] = i;
  }
  wave_add(&B.stats[0], rays);
}

// One scatter per queued subpath (samplePathIter's body after its hit is
// known): the scattered ray is the subpath's next vertex, recorded unless the
// depth limit replaces it by the (never connectable) terminating ray; a ray
// that goes on is queued for the next bounce.
template <bool VOL>
__global__ void __launch_bounds__(kRBlock) k_bpt_record(DevScene S, PathBufs B, PassArgs P, uint32_t b, BptArgs A) {
  __shared__ uint2 stack[kLdsDepth * kRBlock];
  __shared__ uint32_t cont_list[kStageCap], counts[2];
  const StagedQueue qc{cont_list, counts};
  const uint32_t n = B.cnt[b];
  const uint32_t* qin = (b & 1) ? B.q1 : B.q0;
  uint32_t* qout = (b & 1) ? B.q0 : B.q1;
  uint2* lds = stack + threadIdx.x;
  uint32_t rays = 0, segs = 0;
  const uint32_t chunk = blockDim.x * min(kStageIters, max(1u, n / (gridDim.x * blockDim.x)));
  for (uint32_t c0 = blockIdx.x * chunk; c0 < n; c0 += gridDim.x * chunk) {
  staged_reset(qc);
  __syncthreads();
  for (uint32_t i0 = c0; i0 < c0 + chunk; i0 += blockDim.x) {
    const uint32_t i = i0 + threadIdx.x;
    bool cont = false;
    uint32_t p = 0;
    if (i < n) {
      p = qin[i];
      const bool light = p >= A.N;
      uint32_t pix, smp;
      slot_pixel(P, light ? p - A.N : p, pix, smp);
      const bool russian = b >= P.rr_start;
      segs++;
      const Vec3 o = xyz(B.o[p]), d = xyz(B.d[p]);
      const int inside = B.s[p].x;
      const int2 h = B.hit[p];
      const float t = __int_as_float(h.x);
      const int pos = h.y;
      U4 r0 = philox(U4{smp, b, 0u, light ? 1u : 0u}, pix, P.key1);
      U4 r1 = philox(U4{smp, b, 1u, light ? 1u : 0u}, pix, P.key1);
      const float u[8] = {u01(r0.x), u01(r0.y), u01(r0.z), u01(r0.w), u01(r1.x), u01(r1.y), u01(r1.z), u01(r1.w)};
      int hitObj = NONE;
      Vec3 hn = v3(0, 0, 0);
      if (pos >= 0) {
        hitObj = __float_as_int(S.tree.tv1[pos].w);
        hn = xyz(S.tree.tn[pos]);
      }
      const int obj = inside != NONE ? inside : hitObj;
      if (obj != NONE) {  // otherwise the subpath ends with a terminating ray
        Scat sc;
        GuideHead head;
        float originScale;
        scatter_at<VOL, false>(S, o, d, t, pos, hitObj, hn, inside, russian, u, lds, B, rays, light, GuideArgs{}, i,
                               smp, b, pix, P.key1, sc, head, originScale);
        sc.originProb *= originScale;
        Vec3 nc = v3(0, 0, 0);
        if (sc.contact >= 0 && sc.contactPos >= 0) nc = xyz(S.tree.tn[sc.contactPos]);
        if (length(sc.dir) >= 0.5f && sc.inside == NONE && sc.contact != NONE) {  // light-leak guard (:183-201)
          float a = dot(nc, -d), c2 = dot(nc, sc.dir);
          if (a * a <= 0 || c2 * c2 <= 0) {
            sc.dir = v3(0, 0, 0);
            sc.color = v3(0, 0, 0);
          }
        }
        const bool alive = length(sc.dir) >= 0.5f;
        if (!alive || b + 1 <= P.max_depth) {
          BVert v;
          v.o = sc.x;
          v.d = sc.dir;
          v.c = sc.color;
          v.n = nc;
          v.dirProb = sc.dirProb;
          v.originProb = sc.originProb;
          v.contact = sc.contact;
          v.inside = sc.inside;
          v.random = sc.random;
          put_vert(A.V, (size_t)p * A.K + b + 1, v);
          A.nv[p] = b + 2;
        }
        if (alive && b + 1 <= P.max_depth) {
          B.o[p] = f4(sc.x, 0.0f);
          B.d[p] = f4(sc.dir, 0.0f);
          B.s[p] = make_int4(sc.inside, sc.contact, sc.contactPos, 0);
          cont = true;
        }
      }
    }
    staged_push(qc, cont, p);
  }
  __syncthreads();
  staged_claim(qc, &B.cnt[b + 1]);
  __syncthreads();
  staged_flush(qc, qout);
  __syncthreads();
  }
  wave_add(&B.stats[0], rays);
  wave_add(&B.stats[3], segs);
}

// ------------------------------------------------------------------ connections
// The (eye vertex j, light vertex k) pairs of every eye slot that
// colorByConnectingPaths would join (:327-339): neither on an emitter (the
// light vertex may be the emission point), both sampled RANDOM (connectRays).
constexpr uint32_t kPairIters = 8;  // consecutive (slot, eye vertex) entries per thread
__device__ __forceinline__ uint32_t pair_mask(const DevScene& S, const BptArgs& A, uint32_t p, uint32_t j) {
  uint32_t mask = 0;
  if (j < A.nv[p]) {
    int ce;
    bool re;
    vert_flags(A.V, (size_t)p * A.K + j, ce, re);
    if (re && !is_emissive(S, ce)) {
      const uint32_t nL = A.nv[A.N + p];
      for (uint32_t k = 0; k < nL; k++) {
        int cl;
        bool rl;
        vert_flags(A.V, (size_t)(A.N + p) * A.K + k, cl, rl);
        if (rl && !(k >= 1 && is_emissive(S, cl))) mask |= 1u << k;
      }
    }
  }
  return mask;
}
__global__ void __launch_bounds__(kRBlock) k_bpt_pairs(DevScene S, PathBufs B, BptArgs A) {
  __shared__ uint32_t wsum[kRBlock / 64], base;
  const uint32_t K = A.K, total = A.N * K;
  const uint32_t chunk = blockDim.x * kPairIters;
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  for (uint32_t c0 = blockIdx.x * chunk; c0 < total; c0 += gridDim.x * chunk) {
    // thread t: entries c0 + t*kPairIters + [0, kPairIters), i.e. eye vertices of one or two slots
    const uint32_t e0 = c0 + threadIdx.x * kPairIters;
    uint32_t masks[kPairIters], cnt = 0;
#pragma unroll
    for (uint32_t k = 0; k < kPairIters; k++) {
      const uint32_t i = e0 + k;
      const uint32_t p = i / K;
      masks[k] = i < total ? pair_mask(S, A, p, i - p * K) : 0u;
      cnt += (uint32_t)__popc(masks[k]);
    }
    // block-wide exclusive prefix of the counts, one atomic per chunk
    uint32_t x = cnt;
    for (int off = 1; off < 64; off <<= 1) {
      const uint32_t y = __shfl_up(x, off);
      if (lane >= off) x += y;
    }
    if (lane == 63) wsum[wid] = x;
    __syncthreads();
    uint32_t before = 0, all = 0;
    for (int w = 0; w < (int)(blockDim.x >> 6); w++) {
      if (w < wid) before += wsum[w];
      all += wsum[w];
    }
    if (threadIdx.x == 0) base = all ? atomicAdd(&B.cnt[kBptPairCount], all) : 0u;
    __syncthreads();
    uint32_t slot = base + before + x - cnt;
#pragma unroll
    for (uint32_t k = 0; k < kPairIters; k++) {
      const uint32_t i = e0 + k;
      const uint32_t p = i / K, j = i - p * K;
      for (uint32_t m = masks[k]; m; m &= m - 1) A.pairs[slot++] = (p << 10) | (j << 5) | (uint32_t)(__ffs(m) - 1);
    }
    __syncthreads();
  }
}

// The eye vertex of a connection after connectRays (src/MCRenderer.cpp:88-104):
// turned toward the light vertex, with the inside object that implies.
template <bool VOL>
__device__ __forceinline__ void eye_connect_state(const DevScene& S, const BVert& E, const BVert& Eprev, bool has_prev,
                                                  Vec3 dir, int& inside, uint2* lds, const PathBufs& B,
                                                  uint32_t& rays) {
  inside = E.inside;
  if (E.contact != NONE && has_prev) {
    if (dot(E.n, dir) <= 0) inside = E.contact;
    else if (dot(E.n, Eprev.d) <= 0) inside = Eprev.inside;
    else inside = VOL ? find_inside(S, E.o, dir, E.contact, lds, B, rays) : NONE;  // Scene::findInsideObject
  }
}

// MCRenderer::testVisibility for every connectable pair: the first legal hit
// from the eye vertex (with its connected inside object and its contact
// object's condition) is not nearer than the light vertex minus EPSILON.
template <bool VOL>
__global__ void __launch_bounds__(kRBlock) k_bpt_vis(DevScene S, PathBufs B, BptArgs A) {
  __shared__ uint2 stack[kLdsDepth * kRBlock];
  const uint32_t n = B.cnt[kBptPairCount];
  const uint32_t gtid = blockIdx.x * blockDim.x + threadIdx.x;
  uint2* lds = stack + threadIdx.x;
  uint2* spill = B.spill + (size_t)gtid * B.spill_depth;
  uint32_t srays = 0;
  bool ovf = false, active = false;
  WaveWork ww = wave_work(n);
  uint32_t code = 0;
  float dist = 0.0f;
  TraceState ts;
  TraceStats st{0u, 0u};
  for (;;) {
    const unsigned long long idle = __ballot(!active);
    if (idle && !ww.drained) {
      const uint32_t i = wave_take(ww, idle, n, B.cnt + kBptVisWork);
      if (i < n) {
        code = A.pairs[i];
        const uint32_t p = code >> 10, j = (code >> 5) & 31u, k = code & 31u;
        const BVert E = get_vert(A.V, (size_t)p * A.K + j);
        const BVert L = get_vert(A.V, (size_t)(A.N + p) * A.K + k);
        const Vec3 dir = -nrm0(E.o - L.o);
        int inside;
        if (j >= 1) eye_connect_state<VOL>(S, E, get_vert(A.V, (size_t)p * A.K + j - 1), true, dir, inside, lds, B, srays);
        else inside = E.inside;
        dist = length(E.o - L.o);
        srays++;
        active = trace_begin<false>(tree_root(S.tree), ts, E.o, dir, inside, E.contact, &st);
        if (active) {
          ts.limit = dist * 1.001f;  // hits beyond the light vertex cannot block it
          ts.anyhit = dist;
          ts.use_anyhit = true;
        } else {
          atomicOr(&A.vis[(size_t)p * A.K + j], 1u << k);  // no hit at all
        }
      }
    }
    if (ww.drained && !__any(active)) break;
    if (active && trace_trip<false, true, false, kLdsDepth>(S.tree, ts, lds, spill, B.spill_depth, &st, &ovf)) {
      if (!ts.has || ts.best - dist >= -1e-5f) {
        const uint32_t p = code >> 10, j = (code >> 5) & 31u, k = code & 31u;
        atomicOr(&A.vis[(size_t)p * A.K + j], 1u << k);
      }
      active = false;
    }
  }
  if (ovf) *B.overflow = 1;
  wave_add(&B.stats[2], srays);
  wave_add(&B.stats[0], srays);
}

// One connected path of colorByConnectingPaths: vertex i is light vertex i
// for i < lp, else eye vertex W - 1 - i; the two connection vertices carry
// connectRays' changes.
struct BView {
  const float4* V;
  size_t eb, lb;  // first vertex of the eye and of the light subpath
  int W, lp;
  BVert lc, ec;
  __device__ __forceinline__ BVert get(int i) const {
    if (lp >= 1 && i == lp - 1) return lc;
    if (lp >= 1 && i == lp) return ec;
    return i < lp ? get_vert(V, lb + i) : get_vert(V, eb + (W - 1 - i));
  }
  __device__ __forceinline__ bool random(int i) const {
    if (lp >= 1 && (i == lp - 1 || i == lp)) return true;
    int c;
    bool r;
    vert_flags(V, i < lp ? lb + i : eb + (W - 1 - i), c, r);
    return r;
  }
};

// Ray::getCosineTerm (src/Ray.cpp:71-78): media have no cosine term (hasCosineTerm)
__device__ __forceinline__ float bpt_cos(const DevScene& S, const BVert& r) {
  float d = 1.0f;
  const bool has = r.contact == CAMERA || (r.contact >= 0 && S.objs[r.contact].type != OBJ_VPM);
  if (has && length(r.d) > 0.5f) d = fabsf(dot(r.n, r.d));
  return fmaxf(d, 1e-6f);
}
// SceneObject material sums / averages of getBSDF and
// getDirectionSampleProbDensity (src/SceneObject.cpp:234-255,349-359) with
// the material rules of src/DiffuseMaterial.cpp:79-102 and
// src/GlossyMaterial.cpp:59-81; emitters: src/SceneEmissiveObject.cpp:21-26
__device__ Vec3 obj_bsdf(const DevObject& O, Vec3 in_d, const BVert& out) {
  if (O.type == OBJ_EMISSIVE) return dot(out.n, out.d) < 0 ? v3(0, 0, 0) : ld3(O.c1);
  // SceneVPMObject::getBSDF (src/SceneVPMObject.cpp:257-264): sigma_s times the phase function at a
  // scattering point; the dispatch below reaches a medium with a contact only on its own boundary
  if (O.type == OBJ_VPM) return out.contact == NONE ? ld3(O.c1) * hg(O.g, dot(in_d, out.d)) : v3(0, 0, 0);
  if (out.contact == NONE || O.type == OBJ_REFRACTIVE) return v3(0, 0, 0);
  const Vec3 n = out.n;
  const bool front = dot(n, in_d) <= 0 && dot(n, out.d) >= 0;
  Vec3 f = v3(0, 0, 0);
  if (O.type == OBJ_GLOSSY || O.type == OBJ_PHONG) {
    if (front) {
      const Vec3 r = mirror(n, in_d);
      const float rw = dot(r, out.d);
      if (rw > 0) {
        const Vec3 ks = ld3(O.c1);
        f = ks * powf(rw, O.expo) * maxc(ks);
      }
    }
  }
  if (O.type == OBJ_DIFFUSE || O.type == OBJ_PHONG) {
    const Vec3 rho = O.type == OBJ_DIFFUSE ? ld3(O.c1) : ld3(O.c2);
    if (front) f = f + v3(rho.x / 3.14159265358979323846f, rho.y / 3.14159265358979323846f,
                          rho.z / 3.14159265358979323846f);
  }
  return f;
}
__device__ float obj_pdf(const DevObject& O, Vec3 in_d, const BVert& out) {
  if (O.type == OBJ_REFRACTIVE) return 1.0f;  // Material::getDirectionSampleProbDensity
  // SceneVPMObject::getDirectionSampleProbDensity (:275-285): its own boundary crossings are DEFINITE
  if (O.type == OBJ_VPM) return (!out.random || out.contact != NONE) ? 0.0f : O.albedo * hg(O.g, dot(in_d, out.d));
  if (O.type == OBJ_EMISSIVE || out.contact == NONE) return 0.0f;
  float pg = 0.0f, pd = 0.0f;
  if (O.type == OBJ_GLOSSY || O.type == OBJ_PHONG) pg = lobe_pdf(mirror(out.n, in_d), out.d, O.expo);
  if (O.type == OBJ_DIFFUSE || O.type == OBJ_PHONG)
    pd = (O.type == OBJ_DIFFUSE ? O.max1 : O.max2) * lobe_pdf(out.n, out.d, 1.0f);
  return O.type == OBJ_PHONG ? (pg + pd) / 2 : pg + pd;
}
// Ray::getBSDF / Ray::getDirectionSampleProbDensity (src/Ray.cpp:14-40); the
// camera's own (include/Camera.h:31, src/Camera.cpp:99-108)
__device__ __forceinline__ Vec3 bpt_bsdf(const DevScene& S, Vec3 in_d, const BVert& out) {
  if (out.contact == CAMERA) return v3(1, 1, 1);
  if (out.contact >= 0) return obj_bsdf(S.objs[out.contact], in_d, out);
  if (out.inside >= 0) return obj_bsdf(S.objs[out.inside], in_d, out);
  return v3(0, 0, 0);
}
__device__ __forceinline__ float bpt_pdf(const DevScene& S, Vec3 in_d, const BVert& out) {
  if (out.contact == CAMERA) return 1.0f;
  if (out.contact >= 0) return obj_pdf(S.objs[out.contact], in_d, out);
  if (out.inside >= 0) return obj_pdf(S.objs[out.inside], in_d, out);
  return 0.0f;
}
// a ray's density of its own direction: the emitted-ray branch of
// SceneObject::getDirectionSampleProbDensity (src/SceneObject.cpp:234-247)
__device__ __forceinline__ float self_pdf(const BVert& r) {
  if (r.contact == CAMERA) return 1.0f;
  if (r.contact < 0) return 0.0f;
  return lobe_pdf(r.n, r.d, 1.0f);
}

// Ray::getOriginSampleProbDensity (src/Ray.cpp:42-50) of the ray `in` for the
// vertex `out` at distance dist: 1 unless `in` travels in a medium, then
// P_surface (a surface vertex) or p_medium (a scattering point)
__device__ __forceinline__ float origin_pdf(const DevScene& S, const BVert& in, const BVert& out, float dist) {
  if (in.inside < 0 || S.objs[in.inside].type != OBJ_VPM) return 1.0f;
  const DevObject& O = S.objs[in.inside];
  return out.contact != NONE ? P_surface(O, dist) : p_medium(O, dist);
}

__device__ __forceinline__ bool must_use_pt(const BView& v) {
  for (int i = 0; i + 1 < v.W; i++)
    if (v.random(i) && v.random(i + 1)) return false;
  return true;
}
// connectColorProb (:133-181)
__device__ void connect_color_prob(const DevScene& S, const BView& v, int ci, Vec3& color, float& prob) {
  color = v3(1, 1, 1);
  prob = 1.0f;
  BVert prev{}, cur = v.get(0);
  for (int i = 0; i < v.W; i++) {
    const BVert next = (i + 1 < v.W) ? v.get(i + 1) : cur;
    float dist = 0.0f;
    if (i <= ci) {
      dist = fmaxf(length(next.o - cur.o), 1e-5f);
      color = color * decay(S, cur.inside, dist);
    } else if (i > ci + 1) {
      dist = fmaxf(length(prev.o - cur.o), 1e-5f);
      color = color * decay(S, cur.inside, dist);
    }
    if (i == ci && i < v.W - 1) color = color * (bpt_cos(S, cur) * bpt_cos(S, next) / (dist * dist));
    if (i != ci && i != ci + 1) color = color * bpt_cos(S, cur);
    prob *= cur.dirProb * cur.originProb;
    color = color * cur.c;
    prev = cur;
    cur = next;
  }
}
// BidirectionalPathTracer::link (:183-195): vertex i (src) as seen from the
// connected path, re-aimed at vertex j (dst) where the path runs against its
// own sampling direction
__device__ __forceinline__ BVert link_to(BVert src, const BVert& dst, int ci, int i, int j) {
  if ((j > i && i <= ci) || (j < i && i > ci)) return src;
  src.d = nrm0(dst.o - src.o);
  src.inside = dst.inside;
  return src;
}
// connectWeight (:197-263) with expTerm = 1, in double as the reference.  The
// reference fills the backward densities pb[i] (the path's density with the
// light subpath starting at vertex i) and then sums pf[i] pb[i + 1] over the
// neighbours i, i + 1 that are both RANDOM, pf[i] the forward densities
// (mustUsePT: pb[0] when there is no such pair).  With pb[i - 1] = pb[i] G_i
// and pf[i + 1] = pf[i] F_i the sum is pf[0] T_0 for the recurrence
//   T_j = [pair at j] pb[j + 1] + F_j T_{j + 1},   T_{W-1} = 0,
// and the strategy's own term pf[ci] pb[ci + 1] = pf[0] U_0 with U_j =
// [j == ci] pb[j + 1] + F_j U_{j + 1}; one backward pass carries pb, T and U
// in registers (products and sums only, as the reference), sliding a window
// of four vertices j - 1 .. j + 2, so every vertex is read once.
__device__ float connect_weight(const DevScene& S, const BView& v, int ci) {
  const int W = v.W;
  // G_i (vertices i + 1, i, i - 1 as a, b, c) applied to x
  auto back = [&](double x, const BVert& a, const BVert& b, const BVert& c, int i) {
    const float dl = fmaxf(length(b.o - c.o), 1e-5f);
    double r = x * (double)bpt_cos(S, link_to(c, b, ci, i - 1, i));
    r /= (double)dl * (double)dl;
    if (b.random) {
      const BVert bc = link_to(b, c, ci, i, i - 1);
      if (i < W - 1) r *= (double)bpt_pdf(S, link_to(a, b, ci, i + 1, i).d, bc);
      else r *= (double)self_pdf(bc);
      r *= (double)origin_pdf(S, bc, c, dl);
    }
    return r;
  };
  // F_j (vertices j - 1, j, j + 1 as a, b, c)
  auto fwd = [&](const BVert& a, const BVert& b, const BVert& c, int j) {
    const float dl = fmaxf(length(c.o - b.o), 1e-5f);
    double r = (double)bpt_cos(S, link_to(c, b, ci, j + 1, j));
    r /= (double)dl * (double)dl;
    if (b.random) {
      const BVert bc = link_to(b, c, ci, j, j + 1);
      if (j > 0) r *= (double)bpt_pdf(S, link_to(a, b, ci, j - 1, j).d, bc);
      else r *= (double)self_pdf(bc);
      r *= (double)origin_pdf(S, bc, c, dl);
    }
    return r;
  };
  BVert d{}, c = v.get(W - 1), b = W > 1 ? v.get(W - 2) : c;  // vertices j + 2, j + 1, j
  double pb = c.originProb;                                    // pb[j + 1]
  double T = 0.0, U = 0.0;
  bool any_pair = false;
  for (int j = W - 2; j >= 0; j--) {
    const BVert a = j > 0 ? v.get(j - 1) : BVert{};
    const bool pair = b.random && c.random;
    any_pair |= pair;
    const double F = fwd(a, b, c, j);
    T = (pair ? pb : 0.0) + F * T;
    U = (j == ci ? pb : 0.0) + F * U;
    pb = back(pb, d, c, b, j + 1);  // pb[j]
    d = c;
    c = b;
    b = a;
  }
  const double pf0 = c.originProb;  // vertex 0
  const double sum = pf0 * T + (any_pair ? 0.0 : pb);
  const double self = ci == -1 ? pb : pf0 * U;
  const double res = self / sum;
  if (!(res >= 0)) return 0.0f;
  return (float)res;
}

// One split of colorByConnectingPaths for eye slot p: the W-vertex path of
// light prefix lp and eye prefix W - lp, after connectRays
// (src/MCRenderer.cpp:50-106); false when it contributes nothing, else its
// color / prob * weight.  Splits with lp >= 1 are the pair list's, already
// filtered as the reference's loop filters them (:327-339); the visibility
// bit is k_bpt_vis's.
template <bool VOL>
__device__ bool bpt_split(const DevScene& S, const BptArgs& A, uint32_t p, int W, int lp, BView& v, Vec3& out,
                          uint2* lds, const PathBufs& B, uint32_t& rays) {
  const int ep = W - lp, ci = lp - 1;
  v.V = A.V;
  v.eb = (size_t)p * A.K;
  v.lb = (size_t)(A.N + p) * A.K;
  v.W = W;
  v.lp = lp;
  if (lp == 0) {
    int cf;
    bool rf;
    vert_flags(A.V, v.eb + (W - 1), cf, rf);
    if (!is_emissive(S, cf)) return false;
    if (!must_use_pt(v)) return false;
  } else {
    if (!((A.vis[(size_t)p * A.K + ep - 1] >> (lp - 1)) & 1u)) return false;
    BVert lv = get_vert(A.V, v.lb + lp - 1), ev = get_vert(A.V, v.eb + ep - 1);
    // a missing previous ray is the ray itself
    const BVert pl = lp >= 2 ? get_vert(A.V, v.lb + lp - 2) : lv;
    const BVert pe = ep >= 2 ? get_vert(A.V, v.eb + ep - 2) : ev;
    lv.d = nrm0(ev.o - lv.o);
    ev.d = -lv.d;
    lv.dirProb = 1.0f;
    lv.c = bpt_bsdf(S, lp >= 2 ? pl.d : lv.d, lv);
    lv.random = true;
    if (lv.contact != NONE && lp >= 2) {
      if (dot(lv.n, lv.d) <= 0) lv.inside = lv.contact;
      else if (dot(lv.n, pl.d) <= 0) lv.inside = pl.inside;
      else lv.inside = VOL ? find_inside(S, lv.o, lv.d, lv.contact, lds, B, rays) : NONE;
    }
    ev.dirProb = 1.0f;
    ev.c = bpt_bsdf(S, ep >= 2 ? pe.d : ev.d, ev);
    ev.random = true;
    eye_connect_state<VOL>(S, ev, pe, ep >= 2, ev.d, ev.inside, lds, B, rays);
    v.lc = lv;
    v.ec = ev;
  }
  Vec3 color;
  float prob;
  connect_color_prob(S, v, ci, color, prob);
  if (!(prob > 0)) return false;
  if (!(length(color) > 0)) return false;
  const float weight = connect_weight(S, v, ci);
  out = v3(color.x / prob, color.y / prob, color.z / prob) * weight;
  return true;
}

// 2^-32 fixed point into a 64-bit integer sum: additions in any order give
// the same bits.
__device__ __forceinline__ void add_fixed(unsigned long long* dst, Vec3 c) {
  const float cc[3] = {c.x, c.y, c.z};
  for (int ch = 0; ch < 3; ch++) {
    const double fx = (double)cc[ch] * kSplatScale + 0.5;
    if (fx >= 1.0 && fx < 9.0e18) atomicAdd(dst + ch, (unsigned long long)fx);
  }
}

// Paths that are the eye subpath alone, ending on an emitter (lp = 0): one
// thread per eye slot, into L[p].
template <bool VOL>
__global__ void __launch_bounds__(kRBlock) k_bpt_connect(DevScene S, PathBufs B, BptArgs A) {
  const uint32_t stride = gridDim.x * blockDim.x;
  uint32_t rays = 0;  // lp = 0 joins no vertices: no containment queries
  for (uint32_t p = blockIdx.x * blockDim.x + threadIdx.x; p < A.N; p += stride) {
    const int nE = (int)A.nv[p];
    Vec3 own = v3(0, 0, 0);
    BView v;
    for (int W = 2; W <= nE; W++) {
      Vec3 c;
      if (bpt_split<VOL>(S, A, p, W, 0, v, c, nullptr, B, rays)) own = own + c;
    }
    B.L[p] = make_float4(own.x, own.y, own.z, 0.0f);
  }
}

// colorByConnectingPaths' connected splits (light prefix lp >= 1), one
// thread per entry of the pair list: paths with two or more eye vertices add
// to their eye slot's own fixed-point sum; those with the camera vertex alone
// go to the pixel Camera::transToPixel gives the light vertex
// (src/Camera.cpp:136-152), scaled by the light-tracing factors (:357-373),
// into the pass's splat image of the slot's sample.
template <bool VOL>
__global__ void __launch_bounds__(kRBlock) k_bpt_connect_pairs(DevScene S, DevCam C, PathBufs B, PassArgs P,
                                                               BptArgs A) {
  __shared__ uint2 stack[VOL ? kLdsDepth * kRBlock : 1];
  uint32_t rays = 0;
  const uint32_t n = B.cnt[kBptPairCount];
  const uint32_t stride = gridDim.x * blockDim.x;
  const Vec3 front = ld3(C.front), right = ld3(C.right), up = ld3(C.up), cpos = ld3(C.pos);
  const uint32_t npix = P.npix;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    const uint32_t code = A.pairs[i];
    const uint32_t p = code >> 10, j = (code >> 5) & 31u, k = code & 31u;
    const int ep = (int)j + 1, lp = (int)k + 1;
    BView v;
    Vec3 color;
    if (!bpt_split<VOL>(S, A, p, ep + lp, lp, v, color, stack + threadIdx.x, B, rays)) continue;
    if (ep >= 2) {
      add_fixed(A.own + 3 * (size_t)p, color);
      continue;
    }
    uint32_t jpix, jsmp;
    slot_pixel(P, p, jpix, jsmp);
    jsmp -= P.sample;  // the sample's index within the pass
    const Vec3 lo = v.lc.o;
    const Vec3 dir = nrm0(lo - v.ec.o);
    const float cosAtCamera = fabsf(dot(front, dir));
    const uint32_t sd = (uint32_t)C.sight;
    const float imageToSolidAngleFactor = (float)(sd * sd) / (cosAtCamera * cosAtCamera * cosAtCamera);
    const float imageToSurfaceFactor = imageToSolidAngleFactor / cosAtCamera;
    color = color * (imageToSurfaceFactor / (float)npix);
    color = color * cosAtCamera;
    const Vec3 q = lo - cpos;
    const float ratio = C.sight / dot(q, front);
    const float px = dot(q, right) * ratio + ((float)C.W / 2.0f - 0.5f);
    const float py = (float)C.H / 2.0f - dot(q, up) * ratio - 0.5f;
    int ix = -1, iy = -1;
    if (A.registered) {  // the pixel whose camera rays cover the raster position
      const float fx = floorf(px + 0.5f), fy = floorf(py + 0.5f);
      if (fx >= 0.0f && fx < (float)C.W && fy >= 0.0f && fy < (float)C.H) {
        ix = (int)fx;
        iy = (int)fy;
      }
    } else if (px > -1.0f && px < (float)C.W && py > -1.0f && py < (float)C.H) {  // (int) truncates toward 0
      ix = (int)px;
      iy = (int)py;
    }
    if (ix >= 0) add_fixed(A.splat + 3 * ((size_t)jsmp * npix + (uint32_t)(iy * C.W + ix)), color);
  }
  wave_add(&B.stats[0], rays);
}

// Iteration images in sample order: own + splatted radiance of sample j,
// divided by the vignetting factor (MCRenderer::eliminateVignetting), added
// to the FP64 sums.
__global__ void k_bpt_accumulate(PathBufs B, PassArgs P, DevCam C, const unsigned long long* splat,
                                 const unsigned long long* ownfix) {
  const uint32_t stride = gridDim.x * blockDim.x;
  const Vec3 front = ld3(C.front), right = ld3(C.right), up = ld3(C.up);
  for (uint32_t q = blockIdx.x * blockDim.x + threadIdx.x; q < P.npix; q += stride) {
    const int x = (int)(q % (uint32_t)C.W), y = (int)(q / (uint32_t)C.W);
    const Vec3 dc = normalize(front * C.sight + right * ((float)x + 0.5f - (float)C.W * 0.5f) +
                              up * ((float)C.H * 0.5f - 0.5f - (float)y));
    const float vig = fabsf(dot(dc, front));
    double s[3], s2[3];
    for (int k = 0; k < 3; k++) {
      s[k] = B.sum[3 * (size_t)q + k];
      s2[k] = B.sumsq[3 * (size_t)q + k];
    }
    for (uint32_t j = 0; j < P.spass; j++) {
      const uint32_t slot = pass_slot(P, q, j);
      const float4 L = B.L[slot];
      const unsigned long long* sp = splat + 3 * ((size_t)j * P.npix + q);
      const unsigned long long* ow = ownfix + 3 * (size_t)slot;
      const float own[3] = {L.x, L.y, L.z};
      for (int k = 0; k < 3; k++) {
        const float val = (own[k] + (float)((double)ow[k] / kSplatScale) + (float)((double)sp[k] / kSplatScale)) / vig;
        s[k] += val;
        s2[k] += (double)val * val;
      }
    }
    for (int k = 0; k < 3; k++) {
      B.sum[3 * (size_t)q + k] = s[k];
      B.sumsq[3 * (size_t)q + k] = s2[k];
    }
  }
}

static __global__ void k_fill_u32(uint32_t* a, size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) a[i] = 0u;
}

// ------------------------------------------------------------------ host
int render_bpt(rx_ctx* c, const rx_camera_desc* cam, const rx_render_params* prm, float* rgb_mean, double* sum,
               double* sumsq, rx_render_stats* stats) {
  RenderState& R = *c->render;
  if (prm->max_depth + 2 > (uint32_t)kBptMaxVerts)
    return set_err(c, RX_ERR_INVALID, "rx_render: the bidirectional tracer needs max_depth <= 30");
  const uint32_t npix = (uint32_t)(cam->width * cam->height);
  if (npix > kBptMaxPixels) return set_err(c, RX_ERR_INVALID, "rx_render: the bidirectional tracer renders at most 2^22 pixels");
  (void)hipSetDevice(c->device);
  hipStream_t s = c->stream;
  const uint32_t K = prm->max_depth + 2;
  uint32_t slots = std::min(c->tuning.path_slots ? c->tuning.path_slots : kBptPathSlots, kBptPathSlots);
  const uint32_t S_max = std::max<uint32_t>(1u, std::min<uint32_t>(std::max(prm->spp, 1u), slots / 2 / npix));
  const uint32_t N = npix * S_max;  // eye subpaths of a pass (as many light subpaths)
  const uint32_t N2 = 2 * N;
#define RXCK(x, w) \
  if (int r_ = hip_err(c, (x), w)) return r_;
  if (R.cap < N2) {
    RXCK(R.o.ensure(16 * (size_t)N2), "alloc");
    RXCK(R.d.ensure(16 * (size_t)N2), "alloc");
    RXCK(R.beta.ensure(16 * (size_t)N2), "alloc");
    RXCK(R.lastp.ensure(16 * (size_t)N2), "alloc");
    RXCK(R.L.ensure(16 * (size_t)N2), "alloc");
    RXCK(R.s.ensure(16 * (size_t)N2), "alloc");
    RXCK(R.hit.ensure(8 * (size_t)N2), "alloc");
    RXCK(R.q0.ensure(4 * (size_t)N2), "alloc");
    RXCK(R.q1.ensure(4 * (size_t)N2), "alloc");
    RXCK(R.sq.ensure(4 * (size_t)N2), "alloc");
    RXCK(R.so.ensure(16 * (size_t)N2), "alloc");
    RXCK(R.sd.ensure(16 * (size_t)N2), "alloc");
    RXCK(R.sc.ensure(16 * (size_t)N2), "alloc");
    RXCK(R.ss.ensure(8 * (size_t)N2), "alloc");
    R.cap = N2;
  }
  RXCK(R.bverts.ensure(64 * (size_t)N2 * K), "alloc vertices");
  RXCK(R.bnv.ensure(4 * (size_t)N2), "alloc");
  RXCK(R.bvis.ensure(4 * (size_t)N * K), "alloc");
  RXCK(R.bpairs.ensure(4 * (size_t)N * K * K), "alloc pairs");
  RXCK(R.bsplat.ensure(24 * (size_t)N), "alloc splats");
  RXCK(R.bown.ensure(24 * (size_t)N), "alloc own sums");
  RXCK(R.sum.ensure(24 * (size_t)npix), "alloc");
  RXCK(R.sumsq.ensure(24 * (size_t)npix), "alloc");
  RXCK(R.mean.ensure(12 * (size_t)npix), "alloc");
  RXCK(R.cnt.ensure(4 * kCntWords), "alloc");
  RXCK(R.stats.ensure(8 * 8), "alloc");
  if (!c->render_blocks_per_cu) {
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_extend<true, false, false>, kRBlock, 0) != hipSuccess || nb < 1)
      nb = 4;
    c->render_blocks_per_cu = nb;
  }
  const int per_cu = c->tuning.blocks_per_cu ? std::min<int>((int)c->tuning.blocks_per_cu, c->render_blocks_per_cu)
                                             : c->render_blocks_per_cu;
  const int blocks = (int)std::min<uint64_t>((N2 + kRBlock - 1) / kRBlock, (uint64_t)c->n_cu * per_cu);
  int sd = 0;
  if (int r = ensure_spill(c, blocks, kRBlock, (int)R.scene.max_depth, &sd)) return r;
  PathBufs B{};
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
  BptArgs A;
  A.V = R.bverts.as<float4>();
  A.nv = R.bnv.as<uint32_t>();
  A.vis = R.bvis.as<uint32_t>();
  A.pairs = R.bpairs.as<uint32_t>();
  A.splat = R.bsplat.as<unsigned long long>();
  A.own = R.bown.as<unsigned long long>();
  A.N = N;
  A.K = K;
  A.registered = (prm->flags & RX_RENDER_BPT_REGISTERED) ? 1u : 0u;
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
  const BatchArgs BA0{};
  hipLaunchKernelGGL(k_fill, dim3(1024), dim3(256), 0, s, B.sum, 3 * (size_t)npix);
  hipLaunchKernelGGL(k_fill, dim3(1024), dim3(256), 0, s, B.sumsq, 3 * (size_t)npix);
  hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(64), 0, s, B.stats, 8, B.overflow);
  RXCK(hipGetLastError(), "launch fill");
  RXCK(hipEventRecord(c->ev0, s), "event");
  uint64_t launches = 0;
  for (uint32_t done = 0; done < prm->spp;) {
    const uint32_t spass = std::min(S_max, prm->spp - done);
    P.sample = prm->sample_offset + done;
    P.spass = spass;
    P.n_paths = npix * spass;
    A.N = P.n_paths;
    done += spass;
    hipLaunchKernelGGL(k_fill_u32, dim3(1024), dim3(256), 0, s, A.vis, (size_t)A.N * K);
    hipLaunchKernelGGL(k_fill, dim3(1024), dim3(256), 0, s, reinterpret_cast<double*>(A.splat), 3 * (size_t)A.N);
    hipLaunchKernelGGL(k_fill, dim3(1024), dim3(256), 0, s, reinterpret_cast<double*>(A.own), 3 * (size_t)A.N);
    if (vol) hipLaunchKernelGGL(k_bpt_generate<true>, dim3(blocks), dim3(kRBlock), 0, s, S, C, B, P, A);
    else hipLaunchKernelGGL(k_bpt_generate<false>, dim3(blocks), dim3(kRBlock), 0, s, S, C, B, P, A);
    for (uint32_t b = 0; b <= prm->max_depth; b++) {
      hipLaunchKernelGGL((k_extend<true, false, false>), dim3(blocks), dim3(kRBlock), 0, s, S.tree, B, b, BA0);
      if (vol) hipLaunchKernelGGL(k_bpt_record<true>, dim3(blocks), dim3(kRBlock), 0, s, S, B, P, b, A);
      else hipLaunchKernelGGL(k_bpt_record<false>, dim3(blocks), dim3(kRBlock), 0, s, S, B, P, b, A);
      launches++;
    }
    hipLaunchKernelGGL(k_bpt_pairs, dim3(blocks), dim3(kRBlock), 0, s, S, B, A);
    if (vol) {
      hipLaunchKernelGGL(k_bpt_vis<true>, dim3(blocks), dim3(kRBlock), 0, s, S, B, A);
      hipLaunchKernelGGL(k_bpt_connect<true>, dim3(blocks), dim3(kRBlock), 0, s, S, B, A);
      hipLaunchKernelGGL(k_bpt_connect_pairs<true>, dim3(blocks), dim3(kRBlock), 0, s, S, C, B, P, A);
    } else {
      hipLaunchKernelGGL(k_bpt_vis<false>, dim3(blocks), dim3(kRBlock), 0, s, S, B, A);
      hipLaunchKernelGGL(k_bpt_connect<false>, dim3(blocks), dim3(kRBlock), 0, s, S, B, A);
      hipLaunchKernelGGL(k_bpt_connect_pairs<false>, dim3(blocks), dim3(kRBlock), 0, s, S, C, B, P, A);
    }
    hipLaunchKernelGGL(k_bpt_accumulate, dim3(blocks), dim3(kRBlock), 0, s, B, P, C,
                       (const unsigned long long*)A.splat, (const unsigned long long*)A.own);
    RXCK(hipGetLastError(), "launch pass");
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
    stats->samples = (uint64_t)npix * prm->spp;
    stats->ms_total = ms;
    stats->extend_launches = launches;
  }
#undef RXCK
  if (ovf) return set_err(c, RX_ERR_OVERFLOW, "rx_render: traversal stack overflow");
  return RX_OK;
}

}  // namespace rx
