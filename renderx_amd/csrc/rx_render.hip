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
#include <cmath>
#include <cstring>
#include <new>
#include <vector>

#include "rx_context.h"
#include "rx_guide.h"
#include "rx_kernels.h"

namespace rx {

enum { OBJ_DIFFUSE = 0, OBJ_GLOSSY = 1, OBJ_PHONG = 2, OBJ_REFRACTIVE = 3, OBJ_EMISSIVE = 4, OBJ_VPM = 5 };
constexpr int NONE = -1, CAMERA = -2;
constexpr int kRBlock = kTraceBlock;  // render kernels share the traversal block shape
constexpr int kMaxBounce = 64;
constexpr int kShadowBase = 96;       // cnt[kShadowBase + b]: shadow queue length of bounce b
constexpr int kCntWords = 192;
constexpr int kMaxCross = 4;          // index-matched boundary crossings per shadow ray
constexpr uint32_t FLAG_MIS = 1u;     // last scatter was RANDOM (only transparent crossings since)

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
  unsigned long long* stats;  // 0 rays, 1 extend rays, 2 shadow rays, 3 segments, 4 boxes, 5 tris
  uint2* spill;
  int32_t spill_depth;
  int32_t* overflow;
};

struct PassArgs {
  uint32_t n_paths, sample, key1, max_depth, rr_start, mode;
  uint32_t npix;  // pixels; path slot p renders pixel p % npix, sample `sample + p / npix`
};

// Guided mode: network outputs per queue row of the current bounce and the
// training records of the current pass (16 floats per (path, bounce) slot:
// x, -d, w, p_mix(w), p_bsdf(w), lum L before the vertex, lum throughput
// arriving at the vertex, RR probability, lum of the vertex's own NEE, tag).
struct GuideArgs {
  GuideNet net;
  const float* Z4;
  uint16_t* X;
  float* rec;
  uint32_t R;
  float tag;  // > 0: record this pass with that tag
};
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

__device__ __forceinline__ uint32_t wave_append(uint32_t* counter, bool pred) {
  const unsigned long long m = __ballot(pred);
  const int lane = threadIdx.x & 63;
  const int leader = m ? (__ffsll((long long)m) - 1) : 0;
  uint32_t base = 0;
  if (m && lane == leader) base = atomicAdd(counter, (uint32_t)__popcll(m));
  base = __shfl(base, leader);
  return base + (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
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
__global__ void __launch_bounds__(kRBlock) k_generate(DevScene S, DevCam C, PathBufs B, PassArgs P) {
  __shared__ uint2 stack[kLdsDepth * kRBlock];
  const uint32_t stride = gridDim.x * blockDim.x;
  uint32_t rays = 0;
  if (blockIdx.x == 0)
    for (int k = threadIdx.x; k < kCntWords; k += blockDim.x) B.cnt[k] = (k == 0) ? P.n_paths : 0u;
  const Vec3 front = ld3(C.front), right = ld3(C.right), up = ld3(C.up), pos = ld3(C.pos);
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < P.n_paths; i += stride) {
    const uint32_t pix = i % P.npix, smp = P.sample + i / P.npix;
    const int x = (int)(pix % (uint32_t)C.W), y = (int)(pix / (uint32_t)C.W);
    U4 r = philox(U4{smp, 0xFFFFu, 0u, 0u}, pix, P.key1);
    float xp = ((float)x + u01(r.x)) - (float)C.W * 0.5f;
    float yp = (float)C.H * 0.5f - ((float)y + u01(r.y));
    Vec3 d = normalize(front * C.sight + right * xp + up * yp);
    Vec3 dc = normalize(front * C.sight + right * ((float)x + 0.5f - (float)C.W * 0.5f) +
                        up * ((float)C.H * 0.5f - 0.5f - (float)y));
    float beta = fabsf(dot(front, d)) / fabsf(dot(dc, front));
    int inside = S.n_vol ? find_inside(S, pos, d, NONE, stack + threadIdx.x, B, rays) : NONE;
    B.o[i] = f4(pos, 0.0f);
    B.d[i] = f4(d, 0.0f);
    B.s[i] = make_int4(inside, CAMERA, -1, 0);
    B.beta[i] = make_float4(beta, beta, beta, 0.0f);
    B.L[i] = make_float4(0, 0, 0, 0);
    B.q0[i] = i;
  }
  wave_add(&B.stats[0], rays);
}

// ------------------------------------------------------------------ extend
// Nearest hit of every queued path.  Block k runs the contiguous queue range
// [k*per, (k+1)*per) with trace_range (lanes refill from the block's LDS
// counter as their rays end); every ray's result is independent of the lane
// that ran it.
template <bool LEG, bool STATS>
__global__ void __launch_bounds__(kRBlock) k_extend(DevTree T, PathBufs B, uint32_t b) {
  __shared__ uint2 stack[kLdsDepth * kRBlock];
  __shared__ uint32_t blk_next;
  const uint32_t n = B.cnt[b];
  const uint32_t* q = (b & 1) ? B.q1 : B.q0;
  const uint32_t per = (n + gridDim.x - 1) / gridDim.x;
  const uint32_t lo = min(n, blockIdx.x * per), hi = min(n, lo + per);
  if (threadIdx.x == 0) blk_next = lo;
  __syncthreads();
  const uint32_t gtid = blockIdx.x * blockDim.x + threadIdx.x;
  uint32_t boxes = 0, tris = 0;
  bool ovf = false;
  trace_range<false, LEG, STATS, kLdsDepth>(
      T, hi, &blk_next, stack + threadIdx.x, B.spill + (size_t)gtid * B.spill_depth, B.spill_depth, &ovf,
      [&](uint32_t i, Vec3& o, Vec3& d, int32_t& ins, int32_t& con) {
        const uint32_t p = q[i];
        o = xyz(B.o[p]);
        d = xyz(B.d[p]);
        if (LEG) {
          const int4 s = B.s[p];
          ins = s.x;
          con = s.y;
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

// ------------------------------------------------------------------ guide queries
// Encode (hit point, view direction) of every queued path of bounce b; row i
// of the network batch belongs to queue entry i.
__global__ void __launch_bounds__(kRBlock) k_guide_query(PathBufs B, GuideArgs GA, uint32_t b) {
  const uint32_t n = B.cnt[b];
  const uint32_t* q = (b & 1) ? B.q1 : B.q0;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const uint32_t p = q[i];
    const float4 o = B.o[p], d = B.d[p];
    const int2 h = B.hit[p];
    const float t = h.y >= 0 ? __int_as_float(h.x) : 0.0f;
    Vec3 dd = xyz(d);
    guide_encode(GA.net, xyz(o) + dd * t, -dd, GA.X + (size_t)i * kGIn);
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
__device__ __forceinline__ float guided_rr(const DevObject& O) {
  if (O.type == OBJ_PHONG)
    return fminf(1.0f, fmaxf(O.c1[0] + O.c2[0], fmaxf(O.c1[1] + O.c2[1], O.c1[2] + O.c2[2])));
  return O.max1;
}

// Guided surface scatter: with probability alpha sample the learned vMF
// mixture, otherwise the reference's BSDF strategy; weight f cos / p_mix with
// p_mix = alpha q + (1 - alpha) p_bsdf (one-sample MIS, unbiased because
// alpha <= 0.95 keeps the BSDF density in the mixture).
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
__device__ void surface_scatter(const DevScene& S, int obj, Vec3 x, Vec3 d, Vec3 n, int hitPos, int insideIn,
                                bool russian, const float* u, Scat& sc, uint2* lds, const PathBufs& B,
                                uint32_t& rays) {
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
    if (insideIn == obj) outside = find_inside(S, x, d, obj, lds, B, rays);
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
      sc.color = sf * ((n_i * n_i) / (n_t * n_t)) * ((1 - er) / fmaxf(fabsf(dot(n, t)), 1e-6f));
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

template <int MODE>
__global__ void __launch_bounds__(kRBlock) k_shade(DevScene S, PathBufs B, PassArgs P, uint32_t b, GuideArgs GA) {
  __shared__ uint2 stack[kLdsDepth * kRBlock];
  const uint32_t n = B.cnt[b];
  const uint32_t* qin = (b & 1) ? B.q1 : B.q0;
  uint32_t* qout = (b & 1) ? B.q0 : B.q1;
  const uint32_t stride = gridDim.x * blockDim.x;
  uint2* lds = stack + threadIdx.x;
  uint32_t rays = 0, segs = 0;
  const bool russian = b >= P.rr_start;
  for (uint32_t i0 = blockIdx.x * blockDim.x; i0 < n; i0 += stride) {
    const uint32_t i = i0 + threadIdx.x;
    bool cont = false, nee = false;
    uint32_t p = 0;
    if (i < n) {
      p = qin[i];
      const uint32_t pix = p % P.npix, smp = P.sample + p / P.npix;
      segs++;
      const float4 o4 = B.o[p], d4 = B.d[p], be = B.beta[p];
      const int4 s = B.s[p];
      const int2 h = B.hit[p];
      const Vec3 o = xyz(o4), d = xyz(d4);
      Vec3 beta = xyz(be);
      const int inside = s.x;
      const float t = __int_as_float(h.x);
      const int pos = h.y;
      U4 r0 = philox(U4{smp, b, 0u, 0u}, pix, P.key1);
      U4 r1 = philox(U4{smp, b, 1u, 0u}, pix, P.key1);
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
        GuideHead head;
        const Vec3 xh = o + d * t;  // hit point (valid when pos >= 0)
        int cur = obj;
        float originScale = 1.0f;
        bool done = false;
        for (int level = 0; level < 3 && !done; level++) {
          const DevObject& O = S.objs[cur];
          if (O.type == OBJ_VPM) {  // SceneVPMObject::scatter (src/SceneVPMObject.cpp:4-251)
            const float sampDist = -__logf(1.0f - u[5]) / O.lum_st;
            const bool go_in = (hitObj == cur) && (inside != cur);
            const bool be_in = (inside == cur);
            const bool out_of_vol = !(pos >= 0) || sampDist >= t;
            if (go_in || (be_in && out_of_vol && hitObj == cur)) {
              // cross the own boundary, entering or leaving
              int outside = go_in ? cur : find_inside(S, xh, d, cur, lds, B, rays);
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
                  float f = (n_i * n_i) / (n_t * n_t) * (1 - er);
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
              Frame lf = frame_from(d);
              const float g = O.g;
              float cosT;
              if (fabsf(g) < 1e-5f) cosT = 1 - 2 * u[6];
              else {
                float sq = (1 - g * g) / (1 - g + 2 * g * u[6]);
                cosT = (1 + g * g - sq * sq) / (2 * g);
              }
              float sinT = sqrtf(fmaxf(0.0f, 1 - cosT * cosT));
              float sp, cp;
              __sincosf(6.28318530717958647692f * u[7], &sp, &cp);
              Vec3 w = normalize(to_world(lf, v3(sinT * cp, cosT, sinT * sp)));
              float ph = hg(g, dot(d, w));
              sc.inside = cur;
              sc.contact = NONE;
              sc.random = true;
              sc.neeKind = 2;
              sc.neeObj = cur;
              sc.originProb = p_medium(O, sampDist);
              if (u[2] < O.albedo || !russian) {
                sc.dir = w;
                sc.color = ld3(O.c1) * ph;
                sc.dirProb = russian ? O.albedo * ph : ph;
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
            } else if (MODE == 2 && dot(d, hn) <= 0 &&
                       (S.objs[cur].type == OBJ_DIFFUSE || S.objs[cur].type == OBJ_GLOSSY ||
                        S.objs[cur].type == OBJ_PHONG)) {
              guide_head(GA.Z4 + (size_t)i * kGOut, head);
              U4 r3 = philox(U4{smp, b, 3u, 0u}, pix, P.key1);
              const float ug[4] = {u01(r3.x), u01(r3.y), u01(r3.z), u01(r3.w)};
              guided_scatter(S.objs[cur], cur, xh, d, hn, pos, inside, russian, head, ug, u[2], sc);
              done = true;
            } else {
              surface_scatter(S, cur, xh, d, hn, pos, inside, russian, u, sc, lds, B, rays);
              done = true;
            }
          }
        }
        sc.originProb *= originScale;
        // segment into the new vertex: decay by the incoming ray's medium, 1/originProb
        const float seg = length(sc.x - o);
        beta = beta * decay(S, inside, seg);
        // vec3 / 0 leaves the vector unchanged in the reference (nvVector.h operator/=)
        if (sc.originProb != 0.0f) beta = beta * (1.0f / sc.originProb);
        if (MODE >= 1 && sc.neeKind && b < P.max_depth) {
          // next-event estimation to an area light: one light point, MIS with the continuation density.
          // The light point is the path's vertex b + 1 (+ the boundary crossings of the shadow
          // ray), so like a continuation it needs a bounce left.
          U4 r2 = philox(U4{smp, b, 2u, 0u}, pix, P.key1);
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
              pB = (russian ? M.albedo : 1.0f) * ph;
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
          if (alive && b + 1 <= P.max_depth) {
            float cosT = 1.0f;
            if (sc.contact != NONE && S.objs[sc.contact].type != OBJ_VPM)
              cosT = fmaxf(fabsf(dot(nc, sc.dir)), 1e-6f);
            const Vec3 beta_arr = beta;  // throughput arriving at the vertex
            beta = beta * sc.color * (sc.dirProb != 0.0f ? cosT / sc.dirProb : cosT);
            uint32_t flags = (uint32_t)s.w;
            float lastPdf = be.w;
            if (sc.random) {
              flags = FLAG_MIS;
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
            } else if (!sc.transparent) {
              flags = 0;
            }
            B.o[p] = f4(sc.x, 0.0f);
            B.d[p] = f4(sc.dir, 0.0f);
            B.s[p] = make_int4(sc.inside, sc.contact, sc.contactPos, (int)flags);
            B.beta[p] = f4(beta, lastPdf);
            cont = true;
          }
        }
      }
    }
    const uint32_t slot = wave_append(&B.cnt[b + 1], cont);
    if (cont) qout[slot] = p;
    if (MODE >= 1) {
      const uint32_t sslot = wave_append(&B.cnt[kShadowBase + b], nee);
      if (nee) B.sq[sslot] = p;
    }
  }
  wave_add(&B.stats[0], rays);
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
__global__ void __launch_bounds__(kRBlock) k_shadow(DevScene S, PathBufs B, uint32_t b, uint32_t max_depth,
                                                    GuideArgs GA) {
  __shared__ uint2 stack[kLdsDepth * kRBlock];
  const uint32_t n = B.cnt[kShadowBase + b];
  const uint32_t gtid = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t stride = gridDim.x * blockDim.x;
  uint2* lds = stack + threadIdx.x;
  uint2* spill = B.spill + (size_t)gtid * B.spill_depth;
  uint32_t rays = 0, srays = 0;
  bool ovf = false;
  for (uint32_t i = gtid; i < n; i += stride) {
    const uint32_t p = B.sq[i];
    const float4 so = B.so[p], sd = B.sd[p], scv = B.sc[p];
    int2 ctx = B.ss[p];
    Vec3 o = xyz(so), w = xyz(sd);
    float remaining = so.w;
    int inside = ctx.x, contact = ctx.y;
    Vec3 T = v3(1, 1, 1);
    bool visible = false;
    for (int c = 0; c < kMaxCross; c++) {
      TraceStats st;
      Hit h = trace<false, true, false, kLdsDepth>(S.tree, o, w, inside, contact, lds, spill, B.spill_depth, &st,
                                                   &ovf);
      rays++;
      srays++;
      if (h.pos < 0 || h.dist - remaining >= -1e-5f) {
        visible = b + 1 + (uint32_t)c <= max_depth;
        break;
      }
      const int hobj = __float_as_int(S.tree.tv1[h.pos].w);
      if (S.objs[hobj].type != OBJ_VPM) break;
      Vec3 xn = o + w * h.dist;
      int next_inside = (inside != hobj) ? hobj : find_inside(S, xn, w, hobj, lds, B, rays);
      if (refr_coeff(S, inside) != refr_coeff(S, next_inside)) break;  // would bend: not a straight shadow ray
      if (inside != NONE && S.objs[inside].type == OBJ_VPM) {
        const DevObject& M = S.objs[inside];
        Vec3 tr = vexp(v3(-M.st[0] * h.dist, -M.st[1] * h.dist, -M.st[2] * h.dist));
        T = T * tr * (__expf(-M.lum_st * h.dist) / lum(tr));
      }
      o = xn;
      remaining -= h.dist;
      inside = next_inside;
      contact = hobj;
    }
    if (visible && inside != NONE && S.objs[inside].type == OBJ_VPM) {
      const DevObject& M = S.objs[inside];
      Vec3 tr = vexp(v3(-M.st[0] * remaining, -M.st[1] * remaining, -M.st[2] * remaining));
      T = T * tr * (__expf(-M.lum_st * remaining) / lum(tr));
    }
    if (visible) {
      float4 L = B.L[p];
      const Vec3 add = v3(scv.x * T.x, scv.y * T.y, scv.z * T.z);
      B.L[p] = make_float4(L.x + add.x, L.y + add.y, L.z + add.z, 0.0f);
      if (GA.tag > 0.0f) {  // the vertex's own NEE is not part of the radiance along its sampled direction
        float* r = GA.rec + ((size_t)p * GA.R + b) * kRecFloats;
        if (r[15] == GA.tag) r[14] += lum(add);
      }
    }
  }
  if (ovf) *B.overflow = 1;
  wave_add(&B.stats[0], rays);
  wave_add(&B.stats[2], srays);
}

// ------------------------------------------------------------------ training samples
// Deterministic selection of the pass's guided-vertex records: records are
// ranked in (path, bounce) order by a chunked count / scan / write, and an
// evenly strided subset of at most `cap` becomes the training batch with
// weight F/p_sample = lum(L after the vertex) * q_rr / lum(throughput at it).
constexpr int kRecChunk = 1024;

__global__ void k_rec_count(const float* rec, uint32_t K, float tag, uint32_t* bcount) {
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

__global__ void k_rec_scan(uint32_t* bcount, uint32_t nb, uint32_t* total) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  uint32_t s = 0;
  for (uint32_t i = 0; i < nb; i++) {
    const uint32_t c = bcount[i];
    bcount[i] = s;
    s += c;
  }
  *total = s;
}

struct TrainOut {
  float4 *pos, *dir, *wdir;
  float *pb, *wt;
  uint32_t* rows;
};

__global__ void k_rec_write(const float* rec, uint32_t K, float tag, const uint32_t* boff, const uint32_t* total,
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
__global__ void k_guide_gather(TrainOut S, const uint32_t* total, uint32_t store_cap, uint32_t step, uint32_t batch,
                               TrainOut D) {
  const uint32_t n = min(*total, store_cap);
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
// Adds the pass's samples of each pixel in sample order (slot j*npix + pixel
// holds sample `sample + j`), so the FP64 film is independent of how samples
// are grouped into passes.
__global__ void k_accumulate(PathBufs B, uint32_t npix, uint32_t S) {
  const uint32_t stride = gridDim.x * blockDim.x;
  for (uint32_t q = blockIdx.x * blockDim.x + threadIdx.x; q < npix; q += stride) {
    double s[3], s2[3];
    for (int k = 0; k < 3; k++) {
      s[k] = B.sum[3 * (size_t)q + k];
      s2[k] = B.sumsq[3 * (size_t)q + k];
    }
    for (uint32_t j = 0; j < S; j++) {
      const float4 L = B.L[(size_t)j * npix + q];
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
}

__global__ void k_fill(double* a, size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) a[i] = 0.0;
}
__global__ void k_fill_u64(unsigned long long* a, int n, int32_t* ovf) {
  if (threadIdx.x < n) a[threadIdx.x] = 0ull;
  if (threadIdx.x == 0) *ovf = 0;
}
__global__ void k_mean(const double* sum, float* mean, size_t n, double inv) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    mean[i] = (float)(sum[i] * inv);
}

// ------------------------------------------------------------------ host state
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
  DBuf rec, bcount, total;  // guided-mode training records
  uint32_t cap = 0;
  float box_lo[3], box_hi[3];
};

static void leaf_verts_of(const HostTree& t, const float* verts, const uint32_t* tris, std::vector<float>& lv) {
  lv.resize(9 * t.leaf_tris.size());
  for (size_t k = 0; k < t.leaf_tris.size(); k++)
    for (int j = 0; j < 3; j++) memcpy(&lv[9 * k + 3 * j], &verts[3 * (size_t)tris[3 * (size_t)t.leaf_tris[k] + j]], 12);
}

}  // namespace rx

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
  rx_render_state_free(c);
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
  // emitters: triangles with an area CDF (uniform area sampling, pdf 1/A_total)
  std::vector<float> etri, enrm, ecdf;
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
      int32_t oid = (int32_t)o;
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
  for (double a : areas) ecdf.push_back((float)(a / total));
  if (!ecdf.empty()) ecdf.back() = 1.0f;
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

int rx_render(rx_ctx* c, const rx_camera_desc* cam, const rx_render_params* prm, float* rgb_mean, double* sum,
              double* sumsq, rx_render_stats* stats) {
  if (!c || !cam || !prm) return set_err(c, RX_ERR_INVALID, "rx_render: null argument");
  if (!c->scene_loaded || !c->render) return set_err(c, RX_ERR_STATE, "rx_render: no scene loaded");
  if (cam->width <= 0 || cam->height <= 0 || (uint64_t)cam->width * cam->height > (1ull << 28))
    return set_err(c, RX_ERR_INVALID, "rx_render: bad resolution");
  if (prm->max_depth >= (uint32_t)kMaxBounce) return set_err(c, RX_ERR_INVALID, "rx_render: max_depth must be < 64");
  if (prm->mode > RX_MODE_GUIDED) return set_err(c, RX_ERR_INVALID, "rx_render: bad mode");
  if (prm->mode >= RX_MODE_NEE && c->render->scene.dev.n_etri == 0)
    return set_err(c, RX_ERR_INVALID, "rx_render: next-event estimation needs an emitter");
  (void)hipSetDevice(c->device);
  RenderState& R = *c->render;
  hipStream_t s = c->stream;
  const uint32_t npix = (uint32_t)(cam->width * cam->height);
  // samples per pass: enough path slots (~1M) to fill the GPU on small images
  const uint32_t S_max = std::max<uint32_t>(1u, std::min<uint32_t>(std::max(prm->spp, 1u), (1u << 20) / npix));
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
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_extend<true, false>, kRBlock, 0) != hipSuccess || nb < 1)
      nb = 4;
    c->render_blocks_per_cu = nb;
  }
  const int blocks =
      (int)std::min<uint64_t>((N + kRBlock - 1) / kRBlock, (uint64_t)c->n_cu * c->render_blocks_per_cu);
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
  // camera basis (src/Camera.cpp:39-43)
  DevCam C;
  {
    float f[3] = {cam->focus[0] - cam->position[0], cam->focus[1] - cam->position[1], cam->focus[2] - cam->position[2]};
    auto nrm = [](float* v) {
      float r = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
      if (r > 0) for (int k = 0; k < 3; k++) v[k] /= r;
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
  }
  PassArgs P;
  P.n_paths = N;
  P.npix = npix;
  P.key1 = (uint32_t)prm->seed ^ (uint32_t)(prm->seed >> 32) * 0x9E3779B9u;
  P.max_depth = prm->max_depth;
  P.rr_start = prm->rr_start_depth;
  P.mode = prm->mode;
  const DevScene S = R.scene.dev;
  const bool timing = (prm->flags & RX_RENDER_TIMING) != 0;
  const bool count = (prm->flags & RX_RENDER_COUNT_WORK) != 0;
  const int mode = (int)prm->mode;
  GuideArgs GA{};
  uint32_t train_passes = 0, batch = 0, steps = 0, store_cap = 0;
  if (mode == RX_MODE_GUIDED) {
    if (!c->guide) return set_err(c, RX_ERR_STATE, "rx_render: guided mode needs rx_guide_init");
    GuideState& G = *c->guide;
    for (int a = 0; a < 3; a++) {  // encode positions relative to the scene box (slightly padded)
      float pad = 1e-3f * (R.box_hi[a] - R.box_lo[a]) + 1e-6f;
      G.lo[a] = R.box_lo[a] - pad;
      G.hi[a] = R.box_hi[a] + pad;
    }
    // samples with index < train_passes train the network online (a render
    // resumed with sample_offset >= train_passes only uses it)
    train_passes = G.prm.train_passes;
    batch = std::max(256u, G.prm.batch_size);
    steps = std::max(1u, G.prm.steps_per_pass);
    store_cap = (uint32_t)std::min<uint64_t>((uint64_t)batch * steps, 1u << 26);
    if (int r = guide_reserve(c, G, std::max(N, batch))) return r;
    if (prm->sample_offset < train_passes && prm->spp)
      if (int r = guide_reserve_store(c, G, store_cap)) return r;
    const uint32_t Rs = prm->max_depth + 1;
    if (prm->sample_offset < train_passes && prm->spp) {
      RXCK(R.rec.ensure((size_t)N * Rs * kRecFloats * 4), "alloc records");
      RXCK(R.bcount.ensure(((size_t)N * Rs / kRecChunk + 2) * 4), "alloc records");
      RXCK(R.total.ensure(64), "alloc records");
      // tags restart at 1 every render: clear stale records once
      hipLaunchKernelGGL(k_fill, dim3(1024), dim3(256), 0, s, R.rec.as<double>(), (size_t)N * Rs * kRecFloats / 2);
    }
    GA.net = G.net();
    GA.Z4 = G.Z4.as<float>();
    GA.X = G.X.as<uint16_t>();
    GA.rec = R.rec.as<float>();
    GA.R = Rs;
  }
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
  uint64_t launches = 0;
  uint32_t pass = 0;
  for (uint32_t done = 0; done < prm->spp; pass++) {
    P.sample = prm->sample_offset + done;
    // a training pass renders one sample per pixel; otherwise up to S_max samples per pass
    const bool train = mode == RX_MODE_GUIDED && P.sample < train_passes;
    const uint32_t spass = train ? 1u : std::min(S_max, prm->spp - done);
    P.n_paths = npix * spass;
    done += spass;
    GA.tag = train ? (float)(pass + 1) : 0.0f;
    hipLaunchKernelGGL(k_generate, dim3(blocks), dim3(kRBlock), 0, s, S, C, B, P);
    for (uint32_t b = 0; b <= prm->max_depth; b++) {
      hipEvent_t e0 = timing ? mark() : nullptr;
      if (b == 0) {
        if (count) hipLaunchKernelGGL((k_extend<false, true>), dim3(blocks), dim3(kRBlock), 0, s, S.tree, B, b);
        else hipLaunchKernelGGL((k_extend<false, false>), dim3(blocks), dim3(kRBlock), 0, s, S.tree, B, b);
      } else {
        if (count) hipLaunchKernelGGL((k_extend<true, true>), dim3(blocks), dim3(kRBlock), 0, s, S.tree, B, b);
        else hipLaunchKernelGGL((k_extend<true, false>), dim3(blocks), dim3(kRBlock), 0, s, S.tree, B, b);
      }
      launches++;
      hipEvent_t e1 = timing ? mark() : nullptr;
      if (mode == RX_MODE_GUIDED) {  // network inference for this bounce's hit points
        hipLaunchKernelGGL(k_guide_query, dim3(blocks), dim3(kRBlock), 0, s, B, GA, b);
        RXCK(guide_forward(GA.net, GA.X, B.cnt + b, N, nullptr, nullptr, nullptr, c->guide->Z4.as<float>(), c->n_cu, s),
             "guide forward");
        hipLaunchKernelGGL((k_shade<2>), dim3(blocks), dim3(kRBlock), 0, s, S, B, P, b, GA);
      } else if (mode == RX_MODE_NEE) {
        hipLaunchKernelGGL((k_shade<1>), dim3(blocks), dim3(kRBlock), 0, s, S, B, P, b, GA);
      } else {
        hipLaunchKernelGGL((k_shade<0>), dim3(blocks), dim3(kRBlock), 0, s, S, B, P, b, GA);
      }
      hipEvent_t e2 = timing ? mark() : nullptr;
      hipEvent_t e3 = nullptr;
      if (mode) {
        hipLaunchKernelGGL(k_shadow, dim3(blocks), dim3(kRBlock), 0, s, S, B, b, prm->max_depth, GA);
        e3 = timing ? mark() : nullptr;
      }
      if (timing) {
        ext.push_back({e0, e1});
        shd.push_back({e1, e2});
        if (mode) sdw.push_back({e2, e3});
      }
    }
    hipLaunchKernelGGL(k_accumulate, dim3(blocks), dim3(kRBlock), 0, s, B, npix, spass);
    RXCK(hipGetLastError(), "launch pass");
    if (GA.tag > 0.0f) {  // online training on this pass's radiance samples
      GuideState& G = *c->guide;
      const uint32_t K = P.n_paths * GA.R;
      const uint32_t nb = (K + kRecChunk - 1) / kRecChunk;
      hipLaunchKernelGGL(k_rec_count, dim3(nb), dim3(256), 0, s, GA.rec, K, GA.tag, R.bcount.as<uint32_t>());
      hipLaunchKernelGGL(k_rec_scan, dim3(1), dim3(64), 0, s, R.bcount.as<uint32_t>(), nb, R.total.as<uint32_t>());
      const TrainOut ST{G.spos.as<float4>(), G.sdir.as<float4>(), G.swdir.as<float4>(), G.spb.as<float>(),
                        G.swt.as<float>(), G.srows.as<uint32_t>()};
      const TrainOut TB{G.pos.as<float4>(), G.dir.as<float4>(), G.wdir.as<float4>(), G.pb.as<float>(),
                        G.wt.as<float>(), G.rows.as<uint32_t>()};
      hipLaunchKernelGGL(k_rec_write, dim3(nb), dim3(256), 0, s, GA.rec, K, GA.tag, R.bcount.as<uint32_t>(),
                         R.total.as<uint32_t>(), store_cap, B.L, GA.R, ST);
      RXCK(hipGetLastError(), "launch record write");
      for (uint32_t k = 0; k < steps; k++) {
        hipLaunchKernelGGL(k_guide_gather, dim3(std::min<uint32_t>((batch + 255) / 256, 1024u)), dim3(256), 0, s, ST,
                           R.total.as<uint32_t>(), store_cap, k, batch, TB);
        RXCK(hipGetLastError(), "launch gather");
        if (int r = guide_train_rows(c, batch)) return r;
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
    auto tsum = [](const std::vector<std::pair<hipEvent_t, hipEvent_t>>& v) {
      double t = 0;
      for (auto& e : v) {
        float m = 0;
        if (e.first && e.second && hipEventElapsedTime(&m, e.first, e.second) == hipSuccess) t += m;
      }
      return t;
    };
    stats->ms_extend = tsum(ext);
    stats->ms_shade = tsum(shd);
    stats->ms_shadow = tsum(sdw);
  }
  for (auto e : evs) (void)hipEventDestroy(e);
#undef RXCK
  if (ovf) return set_err(c, RX_ERR_OVERFLOW, "rx_render: traversal stack overflow");
  return RX_OK;
}

}  // extern "C"
