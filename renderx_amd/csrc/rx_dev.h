// Device-side data layout and traversal for gfx950.
//
// Geometry arithmetic here must reproduce the reference's float results bit
// for bit, so every file including this header is compiled with
// -ffp-contract=off and correctly rounded division / square root; the
// operation order of each expression follows RenderX's vec3 template
// (include/nvVector.h:366-395, 491-527).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace rx {

// ------------------------------------------------------------------ layout
// Interior node, 64 bytes: both children's boxes and references.
//   f[0..5]  left child lo.xyz, hi.xyz   f[6..11] right child lo.xyz, hi.xyz
//   i[12]    left ref                    i[13]    right ref
// A reference >= 0 is an interior node index; < 0 is leaf ~ref.
struct DevTree {
  const float4* nodes;
  const int32_t* leaf_start;  // n_leaves + 1 offsets into the leaf-ordered triangles
  const float4* tv0;          // v0.xyz, w = triangle id bits
  const float4* tv1;          // v1.xyz, w = object id bits
  const float4* tv2;          // v2.xyz
  const float4* tn;           // face normal xyz, w = object-local triangle bits
  float root_lo[3], root_hi[3];
  int32_t root_ref;
  uint32_t n_leaf_tris;
};

struct Hit {
  float dist;   // -1 = miss
  int32_t pos;  // leaf-ordered triangle position, -1 = miss
};

struct Vec3 {
  float x, y, z;
};
__host__ __device__ __forceinline__ Vec3 v3(float a, float b, float c) { return Vec3{a, b, c}; }
__host__ __device__ __forceinline__ Vec3 operator+(Vec3 a, Vec3 b) { return v3(a.x + b.x, a.y + b.y, a.z + b.z); }
__host__ __device__ __forceinline__ Vec3 operator-(Vec3 a, Vec3 b) { return v3(a.x - b.x, a.y - b.y, a.z - b.z); }
__host__ __device__ __forceinline__ Vec3 operator-(Vec3 a) { return v3(-a.x, -a.y, -a.z); }
__host__ __device__ __forceinline__ Vec3 operator*(Vec3 a, float s) { return v3(a.x * s, a.y * s, a.z * s); }
__host__ __device__ __forceinline__ Vec3 operator*(float s, Vec3 a) { return v3(s * a.x, s * a.y, s * a.z); }
__host__ __device__ __forceinline__ Vec3 operator*(Vec3 a, Vec3 b) { return v3(a.x * b.x, a.y * b.y, a.z * b.z); }
__host__ __device__ __forceinline__ Vec3 cross(Vec3 l, Vec3 r) {
  return v3(l.y * r.z - l.z * r.y, l.z * r.x - l.x * r.z, l.x * r.y - l.y * r.x);
}
// ((0 + a.x*b.x) + a.y*b.y) + a.z*b.z: the leading 0 + turns -0 into +0 like the reference
__host__ __device__ __forceinline__ float dot(Vec3 a, Vec3 b) {
  float r = 0.0f;
  r += a.x * b.x;
  r += a.y * b.y;
  r += a.z * b.z;
  return r;
}
__host__ __device__ __forceinline__ float length(Vec3 a) { return __builtin_sqrtf(dot(a, a)); }
__host__ __device__ __forceinline__ Vec3 normalize(Vec3 a) {
  float r = length(a);
  if (r > 0) return v3(a.x / r, a.y / r, a.z / r);
  return a;
}
__device__ __forceinline__ Vec3 xyz(float4 f) { return v3(f.x, f.y, f.z); }
__device__ __forceinline__ float comp(Vec3 a, int i) { return i == 0 ? a.x : (i == 1 ? a.y : a.z); }

// ------------------------------------------------------------------ predicates
// KDTree::Node::BoundingBox::intersect (src/KDTree.cpp:34-52) reduced to the
// only thing the reference reads from it: minDist >= 0, i.e. some face passes.
// Faces are independent, so the loop may stop at the first passing face.
__host__ __device__ __forceinline__ bool face_pass(float plane, float o_a, float d_a, float o_b, float d_b, float lo_b,
                                          float hi_b, float o_c, float d_c, float lo_c, float hi_c) {
  if (d_a == 0) return false;
  float dist = (plane - o_a) / d_a;
  float pb = o_b + dist * d_b;
  float pc = o_c + dist * d_c;
  return dist >= 0 && lo_b <= pb && hi_b >= pb && lo_c <= pc && hi_c >= pc;
}
__host__ __device__ __forceinline__ bool box_hit(const float lo[3], const float hi[3], Vec3 o, Vec3 d) {
  // faces i = 0..5: plane ((float*)this)[i], axis a = i%3, checks (i+1)%3, (i+2)%3
  if (face_pass(lo[0], o.x, d.x, o.y, d.y, lo[1], hi[1], o.z, d.z, lo[2], hi[2])) return true;
  if (face_pass(lo[1], o.y, d.y, o.z, d.z, lo[2], hi[2], o.x, d.x, lo[0], hi[0])) return true;
  if (face_pass(lo[2], o.z, d.z, o.x, d.x, lo[0], hi[0], o.y, d.y, lo[1], hi[1])) return true;
  if (face_pass(hi[0], o.x, d.x, o.y, d.y, lo[1], hi[1], o.z, d.z, lo[2], hi[2])) return true;
  if (face_pass(hi[1], o.y, d.y, o.z, d.z, lo[2], hi[2], o.x, d.x, lo[0], hi[0])) return true;
  return face_pass(hi[2], o.z, d.z, o.x, d.x, lo[0], hi[0], o.y, d.y, lo[1], hi[1]);
}

// Per-ray constants of the slab computations: inv = 1 / d (correctly
// rounded) and oi = o * inv (rounded), so that a slab end is one FMA,
// t = fma(plane, inv, -oi), instead of (plane - o) * inv.  Besides its
// relative rounding (2u), t then carries an absolute error of at most
// u |oi_a|; slab_cull_margin bounds it for the ordered traversal's culling,
// and the filter's beta_a (16u B_a / |d_a| >= 16u |oi_a|) covers it.
__host__ __device__ __forceinline__ Vec3 slab_oi(Vec3 o, Vec3 inv) { return v3(o.x * inv.x, o.y * inv.y, o.z * inv.z); }
// Absolute margin added to every cull threshold: twice the largest |oi_a|
// rounding bound over the axes the key uses (d_a != 0); INF (no culling)
// when one is not finite.
__host__ __device__ __forceinline__ float slab_cull_margin(Vec3 d, Vec3 oi) {
  const float INF = __builtin_inff();
  float ax = d.x != 0 ? fabsf(oi.x) : 0.0f, ay = d.y != 0 ? fabsf(oi.y) : 0.0f, az = d.z != 0 ? fabsf(oi.z) : 0.0f;
  ax = ax <= INF ? ax : INF;
  ay = ay <= INF ? ay : INF;
  az = az <= INF ? az : INF;
  float m = ax;
  m = ay > m ? ay : m;
  m = az > m ? az : m;
  return m * (2.0f * 0x1p-24f);
}

// Slab entry distance with per-ray reciprocals: ordering key only (not a hit
// test).  One FMA per slab end (slab_oi), compare-select; identical to the
// oracle's key.
__host__ __device__ __forceinline__ float slab_tnear(const float lo[3], const float hi[3], Vec3 oi, Vec3 d, Vec3 inv) {
  const float INF = __builtin_inff();
  float t0, t1, a, b, c;
  if (d.x == 0) { t0 = -INF; t1 = INF; } else { t0 = fmaf(lo[0], inv.x, -oi.x); t1 = fmaf(hi[0], inv.x, -oi.x); }
  a = t0 < t1 ? t0 : t1;
  if (d.y == 0) { t0 = -INF; t1 = INF; } else { t0 = fmaf(lo[1], inv.y, -oi.y); t1 = fmaf(hi[1], inv.y, -oi.y); }
  b = t0 < t1 ? t0 : t1;
  if (d.z == 0) { t0 = -INF; t1 = INF; } else { t0 = fmaf(lo[2], inv.z, -oi.z); t1 = fmaf(hi[2], inv.z, -oi.z); }
  c = t0 < t1 ? t0 : t1;
  float t = a;
  t = b > t ? b : t;
  t = c > t ? c : t;
  return t;
}

// Certain-miss / certain-hit filter for box_hit.  If box_hit passes a face
// of axis c at dist t, the real hit point t* = (plane - o_c) / d_c >= 0 lies
// in every slab a widened by the rounding of the face check, |p_a - p_a*| <=
// 4u (|t*| |d_a| + B_a) (u = 2^-24, B_a = max |coordinate| on axis a of the
// scene box and the ray origin):
// in t units e_a - 4u t* - beta_a <= t* <= x_a + 4u t* + beta_a with
// beta_a = 4u B_a / |d_a| (+ an absolute floor).  The slab ends e_a, x_a are
// computed here with reciprocals as fma(plane, inv, -oi) (relative error <=
// 2u, bounded by 8u with this function's own rounding, plus an absolute
// u |oi_a| <= u B_a / |d_a| that beta's 16u B_a / |d_a| covers with the face
// check's 4u B_a / |d_a|), so when no t >= 0 satisfies the widened
// bounds, box_hit is certainly false (kBoxMiss).  Conversely, if the ray
// enters the box (or, from inside, leaves it) through the face of axis A at
// t* > 0 and t* lies inside every other slab by more than those bounds, the
// face check of that face passes and box_hit is certainly true (kBoxHit).
// Needs |d_a| >= 2^-30 on every axis (no overflow, no infinities); `beta`
// holds 16u B_a / |d_a| + 2^-120 / |d_a|.  tnear receives slab_tnear's value.
constexpr int kBoxUnknown = 0, kBoxMiss = 1, kBoxHit = 2;
// median of three NaN-free floats (one v_med3_f32 on the device)
__host__ __device__ __forceinline__ float med3f(float a, float b, float c) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __builtin_amdgcn_fmed3f(a, b, c);
#else
  return fmaxf(fminf(a, b), fminf(fmaxf(a, b), c));
#endif
}
#ifndef RX_CERTAIN_HIT
#define RX_CERTAIN_HIT 1  // 0: only the miss half (fewer registers in the traversal loop)
#endif
// slab_classify's two verdicts as lane masks: certain miss, certain hit
// (never both); neither means the exact predicate must decide.
__host__ __device__ __forceinline__ void slab_verdicts(const float lo[3], const float hi[3], Vec3 oi, Vec3 inv,
                                                       Vec3 beta, float& tnear, bool& certain_miss,
                                                       bool& certain_hit) {
  // the set-up (|d_a| >= 2^-30, finite boxes) keeps NaN out of every value
  // below, so min/max equal the compare-selects of slab_tnear
  const float U8 = 8.0f * 0x1p-24f, U24 = 24.0f * 0x1p-24f;
  const float t0x = fmaf(lo[0], inv.x, -oi.x), t1x = fmaf(hi[0], inv.x, -oi.x);
  const float t0y = fmaf(lo[1], inv.y, -oi.y), t1y = fmaf(hi[1], inv.y, -oi.y);
  const float t0z = fmaf(lo[2], inv.z, -oi.z), t1z = fmaf(hi[2], inv.z, -oi.z);
  const float ex = fminf(t0x, t1x), xx = fmaxf(t0x, t1x);
  const float ey = fminf(t0y, t1y), xy = fmaxf(t0y, t1y);
  const float ez = fminf(t0z, t1z), xz = fmaxf(t0z, t1z);
  const float t = fmaxf(ex, fmaxf(ey, ez));
  tnear = t;
  // The rounding bound of each slab end, U8 |end| + beta_a, rounded once;
  // the end is widened outwards by it for the miss test and narrowed inwards
  // for the hit test, each with one more rounding: no larger error than
  // rounding end -+ U8 |end| and then subtracting / adding beta_a.
  const float wex = fmaf(U8, fabsf(ex), beta.x), wxx = fmaf(U8, fabsf(xx), beta.x);
  const float wey = fmaf(U8, fabsf(ey), beta.y), wxy = fmaf(U8, fabsf(xy), beta.y);
  const float wez = fmaf(U8, fabsf(ez), beta.z), wxz = fmaf(U8, fabsf(xz), beta.z);
  // widened slab ends: true e_a - beta_a >= le_a, true x_a + beta_a <= hx_a
  const float L = fmaxf(fmaxf(0.0f, ex - wex), fmaxf(ey - wey, ez - wez));
  const float R = fminf(xx + wxx, fminf(xy + wxy, xz + wxz));
  // L (1 - 16u) > R (1 + 16u), with one rounding: L >= 0, and for R >= 0
  // (1 + 16u) / (1 - 16u) < 1 + 33u, so this is the stronger condition
  const bool miss = L > R * (1.0f + 33.0f * 0x1p-24f);
  certain_miss = miss;
  certain_hit = false;
#if RX_CERTAIN_HIT == 0
  return;
#endif
  // candidate face: the entry face if the entry is ahead (t > 0: the slab
  // end has the sign of (plane - o) / d, so the real entry is ahead too),
  // else the exit face (the first axis whose entry, or exit, equals it); its
  // point must lie inside the other two slabs with the rounding bounds to
  // spare: the candidate axis's own bounds are replaced by -inf / +inf and
  // the three are reduced, which keeps the loop free of branches.
  const float INF = __builtin_inff();
  const float eux = ex + wex, xdx = xx - wxx;
  const float euy = ey + wey, xdy = xy - wxy;
  const float euz = ez + wez, xdz = xz - wxz;
  const bool entry = t > 0.0f;
  const float tx = fminf(xx, fminf(xy, xz));
  const float tf = entry ? t : tx;
  const bool fx = tf == (entry ? ex : xx);
  const bool fy = !fx && tf == (entry ? ey : xy);
  const bool fz = !fx && !fy;
  const float eo = fmaxf(fx ? -INF : eux, fmaxf(fy ? -INF : euy, fz ? -INF : euz));
  const float xo = fminf(fx ? INF : xdx, fminf(fy ? INF : xdy, fz ? INF : xdz));
  const bool ok = tf > 0.0f && eo <= tf * (1.0f - U24) && tf * (1.0f + U24) <= xo;
  certain_hit = ok && !miss;
}
__host__ __device__ __forceinline__ int slab_classify(const float lo[3], const float hi[3], Vec3 oi, Vec3 inv,
                                                      Vec3 beta, float& tnear) {
  bool m, h;
  slab_verdicts(lo, hi, oi, inv, beta, tnear, m, h);
  return m ? kBoxMiss : (h ? kBoxHit : kBoxUnknown);
}
// The miss half of slab_classify.
__host__ __device__ __forceinline__ bool slab_certain_miss(const float lo[3], const float hi[3], Vec3 oi, Vec3 inv,
                                                           Vec3 beta, float& tnear) {
  return slab_classify(lo, hi, oi, inv, beta, tnear) == kBoxMiss;
}
// Per-ray set-up of slab_certain_miss; false when the filter must not be used.
// B_a bounds every coordinate the face check adds or subtracts on axis a:
// the scene box's and the ray origin's (a camera may sit outside the box).
__host__ __device__ __forceinline__ bool slab_filter_setup(const float root_lo[3], const float root_hi[3], Vec3 o,
                                                           Vec3 d, Vec3 inv, Vec3& beta) {
  const float m = 0x1p-30f;
  if (!(fabsf(d.x) >= m && fabsf(d.y) >= m && fabsf(d.z) >= m)) return false;
  const float c = 16.0f * 0x1p-24f;
  const float bx = fmaxf(fmaxf(fabsf(root_lo[0]), fabsf(root_hi[0])), fabsf(o.x));
  const float by = fmaxf(fmaxf(fabsf(root_lo[1]), fabsf(root_hi[1])), fabsf(o.y));
  const float bz = fmaxf(fmaxf(fabsf(root_lo[2]), fabsf(root_hi[2])), fabsf(o.z));
  beta = v3((c * bx + 0x1p-120f) * fabsf(inv.x) * 1.0001f, (c * by + 0x1p-120f) * fabsf(inv.y) * 1.0001f,
            (c * bz + 0x1p-120f) * fabsf(inv.z) * 1.0001f);
  return true;
}

// KDTree::intersect(ray, triangle) (src/KDTree.cpp:267-293) without the
// condition.  The edge-sign masks decide the hit before the normalisation and
// division, which only run for candidates (no side effects, same result).
__device__ __forceinline__ float tri_dist(Vec3 v0, Vec3 v1, Vec3 v2, Vec3 o, Vec3 d) {
  Vec3 l0 = v1 - v0, l1 = v2 - v1, l2 = v0 - v2;
  Vec3 r0 = v0 - o, r1 = v1 - o, r2 = v2 - o;
  float s0 = dot(d, cross(l0, r0));
  float s1 = dot(d, cross(l1, r1));
  float s2 = dot(d, cross(l2, r2));
  bool pos = s0 >= 0 && s1 >= 0 && s2 >= 0;
  bool neg = s0 <= 0 && s1 <= 0 && s2 <= 0;
  if (!pos && !neg) return -1.0f;
  Vec3 n = normalize(cross(l0, l1));
  float orthoDist = dot(n, r2);
  float den = dot(n, d);
  if (den == 0) return -1.0f;
  float dist = orthoDist / den;
  if (dist < 0) return -1.0f;
  return dist;  // may be NaN for degenerate triangles: the caller's dist >= 0 rejects it
}

// A contact id that stands for "no condition": the query takes the nearest hit
// as Scene::intersect without a Condition does (camera rays in a traversal
// that applies the condition to the other rays of its queue).
constexpr int32_t kNoCondition = -3;
// NoSelfIntersectionCondition::legal (src/NoSelfIntersectionCondition.cpp:12-27)
__device__ __forceinline__ bool legal(Vec3 d, Vec3 n, int32_t obj, int32_t inside, int32_t contact) {
  if (contact == kNoCondition) return true;
  float dn = dot(d, n);
  bool in = dn <= 0, out = dn >= 0;
  if (inside == contact) return !(in && contact == obj);
  return !(out && inside != obj);
}

struct TraceStats {
  uint32_t boxes, tris;
};

// Nearest hit of the reference tree.  EXHAUSTIVE visits every child whose
// box passes (the reference recursion's visit set); otherwise the nearer
// child (slab key) goes first and subtrees whose key exceeds best*1.001 are
// culled.  Both return the reference's winner: among equal distances the
// larger DFS leaf index wins (right subtree wins, src/KDTree.cpp:317-321),
// inside a leaf the first triangle wins (:305).
//
// stack: per-lane LDS column of LDS_DEPTH entries (stride = blockDim.x),
// spill: per-thread global overflow of SPILL_DEPTH entries.
// Traversal state of one ray, so that a lane can run it in bounded trips
// (descend to a leaf, test the leaf, pop) and take a new ray when it ends.
struct TraceState {
  Vec3 o, d, inv, oi, beta;
  float cm;     // slab_cull_margin: absolute term of every cull threshold
  bool fast, has;
  float best;
  float limit;  // subtrees whose entry key exceeds it are culled (INF: none)
  float anyhit; // with use_anyhit, a legal hit with dist - anyhit < -1e-5 ends the traversal,
  bool use_anyhit;
  bool anyhit_front;     // ... and with anyhit_front only a front-facing hit on an object other
  int32_t anyhit_skip;   // than anyhit_skip does
  int32_t bestLeaf, bestPos, ref, inside, contact;
  int32_t skip;  // with SKIP: the triangle id the ray never hits (IntersectionGPU::Ray::last_tid), -1 none
  int sp;
};

// Root of one tree inside a node array shared by several trees.
struct TreeRoot {
  int32_t ref;
  float lo[3], hi[3];
};
__device__ __forceinline__ TreeRoot tree_root(const DevTree& T) {
  return TreeRoot{T.root_ref, {T.root_lo[0], T.root_lo[1], T.root_lo[2]}, {T.root_hi[0], T.root_hi[1], T.root_hi[2]}};
}

// Root box test and set-up; false when the ray misses the scene box.
template <bool STATS>
__device__ __forceinline__ bool trace_begin(const TreeRoot& R, TraceState& s, Vec3 o, Vec3 d, int32_t inside,
                                            int32_t contact, TraceStats* st) {
  if (STATS) { st->boxes = 1; st->tris = 0; }
  s.o = o;
  s.d = d;
  s.inside = inside;
  s.contact = contact;
  s.has = false;
  s.best = -1.0f;
  s.limit = __builtin_inff();
  s.anyhit = 0.0f;
  s.use_anyhit = false;
  s.anyhit_front = false;
  s.anyhit_skip = INT32_MIN;
  s.skip = -1;
  s.bestLeaf = s.bestPos = -1;
  s.sp = 0;
  s.ref = R.ref;
  s.inv = v3(1.0f / d.x, 1.0f / d.y, 1.0f / d.z);
  s.oi = slab_oi(o, s.inv);
  s.cm = slab_cull_margin(d, s.oi);
  s.beta = v3(0, 0, 0);
  s.fast = slab_filter_setup(R.lo, R.hi, o, d, s.inv, s.beta);
  if (s.fast) {  // the root box as any other: the exact predicate only if undecided
    float tn;
    const int c = slab_classify(R.lo, R.hi, s.oi, s.inv, s.beta, tn);
    if (c != kBoxUnknown) return c == kBoxHit;
  }
  return box_hit(R.lo, R.hi, o, d);
}

// One trip: descend interior nodes to a leaf (or a dead end), test the
// leaf's triangles, pop the next subtree.  Returns true when the traversal
// has ended (s.has / s.best / s.bestPos hold the result).  EXHAUSTIVE visits
// every child whose box passes (the reference recursion's visit set);
// otherwise the nearer child (slab key) goes first and subtrees whose key
// exceeds best*1.001 are culled.  Both give the reference's winner: among
// equal distances the larger DFS leaf index wins (right subtree wins,
// src/KDTree.cpp:317-321), inside a leaf the first triangle wins (:305).
//
// lds_stack: per-lane LDS column of LDS_DEPTH entries (stride = blockDim.x),
// spill: per-thread global overflow of spill_depth entries.  SKIP: the
// triangle whose id is s.skip is passed over before its test, the GLSL
// seam's last_tid rule (Shader/Intersection.frag:63-64).
template <bool EXHAUSTIVE, bool LEGALITY, bool STATS, int LDS_DEPTH, bool ANYHIT = true, bool SKIP = false>
__device__ __forceinline__ bool trace_trip(const DevTree& T, TraceState& s, uint2* lds_stack, uint2* spill,
                                           int spill_depth, TraceStats* st, bool* overflow) {
  const float INF = __builtin_inff();
  const int stride = blockDim.x;
  // The lane's stack column lies in LDS: an LDS-qualified pointer keeps its
  // accesses ds_read / ds_write (a pointer chosen between the column and the
  // global spill would be a flat access), and the column offset of the top
  // entry, sp * stride, moves by one stride per push and pop.
#if defined(__HIP_DEVICE_COMPILE__)
  typedef __attribute__((address_space(3))) unsigned long long LdsU64;
  LdsU64* const col = (LdsU64*)lds_stack;
#else
  unsigned long long* const col = reinterpret_cast<unsigned long long*>(lds_stack);
#endif
  // an entry (ref, key) as one 64-bit LDS word: ref in the low half
  auto pack = [](uint2 e) { return (unsigned long long)e.x | ((unsigned long long)e.y << 32); };
  int top = s.sp * stride;
  int32_t ref = s.ref;
  while (ref >= 0) {
    const float4* np = T.nodes + 4 * (size_t)ref;
    float4 a = np[0], b = np[1], c = np[2];
    int4 r = *reinterpret_cast<const int4*>(np + 3);
    const float llo[3] = {a.x, a.y, a.z}, lhi[3] = {a.w, b.x, b.y};
    const float rlo[3] = {b.z, b.w, c.x}, rhi[3] = {c.y, c.z, c.w};
    if (STATS) st->boxes += 2;
    bool hl, hr;
    float tl = 0.0f, tr = 0.0f;
    if (s.fast) {  // the exact predicate only where the slab filter cannot decide
      bool ml, cl, mr, cr;
      slab_verdicts(llo, lhi, s.oi, s.inv, s.beta, tl, ml, cl);
      slab_verdicts(rlo, rhi, s.oi, s.inv, s.beta, tr, mr, cr);
      // a child the ordered traversal culls (entry beyond the cull distance)
      // needs no exact test: its answer would not be used
      const float cull = EXHAUSTIVE ? INF : fminf(s.has ? fmaf(s.best, 1.001f, s.cm) : INF, s.limit + s.cm);
      hl = cl || (!ml && (EXHAUSTIVE || tl <= cull) && box_hit(llo, lhi, s.o, s.d));
      hr = cr || (!mr && (EXHAUSTIVE || tr <= cull) && box_hit(rlo, rhi, s.o, s.d));
    } else {
      hl = box_hit(llo, lhi, s.o, s.d);
      hr = box_hit(rlo, rhi, s.o, s.d);
      if (!EXHAUSTIVE) {
        tl = slab_tnear(llo, lhi, s.oi, s.d, s.inv);
        tr = slab_tnear(rlo, rhi, s.oi, s.d, s.inv);
      }
    }
    int32_t next = INT32_MIN;
    if (EXHAUSTIVE) {
      if (hl && hr) {
        const uint2 e = make_uint2((uint32_t)r.y, 0u);
        if (s.sp < LDS_DEPTH) col[top] = pack(e);
        else if (s.sp - LDS_DEPTH < spill_depth) spill[s.sp - LDS_DEPTH] = e;
        else { *overflow = true; s.has = false; return true; }
        s.sp++;
        top += stride;
        next = r.x;
      } else if (hl) next = r.x;
      else if (hr) next = r.y;
    } else {
      const float thr = fminf(s.has ? fmaf(s.best, 1.001f, s.cm) : INF, s.limit + s.cm);
      const bool vl = hl && tl <= thr, vr = hr && tr <= thr;
      if (vl && vr) {
        const int32_t nearRef = (tr < tl) ? r.y : r.x;
        const int32_t farRef = (tr < tl) ? r.x : r.y;
        const float farT = (tr < tl) ? tl : tr;
        const uint2 e = make_uint2((uint32_t)farRef, __float_as_uint(farT));
        if (s.sp < LDS_DEPTH) col[top] = pack(e);
        else if (s.sp - LDS_DEPTH < spill_depth) spill[s.sp - LDS_DEPTH] = e;
        else { *overflow = true; s.has = false; return true; }
        s.sp++;
        top += stride;
        next = nearRef;
      } else if (vl) next = r.x;
      else if (vr) next = r.y;
    }
    ref = next;
  }
  if (ref != INT32_MIN) {  // a leaf
    const int32_t leaf = ~ref;
    const int32_t p0 = T.leaf_start[leaf], p1 = T.leaf_start[leaf + 1];
    for (int32_t p = p0; p < p1; p++) {
      if (STATS) st->tris++;
      float4 a = T.tv0[p], b = T.tv1[p], c = T.tv2[p];
      if (SKIP && __float_as_int(a.w) == s.skip) continue;
      float dist = tri_dist(xyz(a), xyz(b), xyz(c), s.o, s.d);
      if (!(dist >= 0)) continue;
      bool win = !s.has || dist < s.best || (dist == s.best && leaf > s.bestLeaf);
      if (!win) continue;
      float4 n = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
      if (LEGALITY) {
        n = T.tn[p];
        if (!legal(s.d, xyz(n), __float_as_int(b.w), s.inside, s.contact)) continue;
      }
      s.has = true;
      s.best = dist;
      s.bestLeaf = leaf;
      s.bestPos = p;
      if (ANYHIT && s.use_anyhit && dist - s.anyhit < -1e-5f &&
          (!s.anyhit_front || (__float_as_int(b.w) != s.anyhit_skip && dot(s.d, xyz(n)) < 0.0f))) {
        s.ref = INT32_MIN;  // the caller only needs to know that such a hit exists
        return true;
      }
    }
  }
  // pop
  ref = INT32_MIN;
  while (s.sp > 0) {
    s.sp--;
    top -= stride;
    // the LDS entry is read unconditionally (an in-range one below the
    // spill), the spill entry only where the stack has spilled (a volatile
    // read, which the compiler cannot merge with the LDS read into one
    // flat access through a selected pointer)
    const unsigned long long w = col[s.sp < LDS_DEPTH ? top : (LDS_DEPTH - 1) * stride];
    uint2 e = make_uint2((uint32_t)w, (uint32_t)(w >> 32));
    if (s.sp >= LDS_DEPTH) {
      const volatile uint2* sv = spill + (s.sp - LDS_DEPTH);
      e.x = sv->x;
      e.y = sv->y;
    }
    const float key = __uint_as_float(e.y);
    if (EXHAUSTIVE || (s.has ? key <= fminf(fmaf(s.best, 1.001f, s.cm), s.limit + s.cm) : !(key > s.limit + s.cm))) {
      ref = (int32_t)e.x;
      break;
    }
  }
  s.ref = ref;
  return ref == INT32_MIN;
}

// Nearest hit of the reference tree (trace_begin + trips until done).
template <bool EXHAUSTIVE, bool LEGALITY, bool STATS, int LDS_DEPTH>
__device__ __forceinline__ Hit trace(const DevTree& T, Vec3 o, Vec3 d, int32_t inside, int32_t contact,
                                     uint2* lds_stack, uint2* spill, int spill_depth, TraceStats* st,
                                     bool* overflow) {
  TraceState s;
  if (!trace_begin<STATS>(tree_root(T), s, o, d, inside, contact, st)) return Hit{-1.0f, -1};
  while (!trace_trip<EXHAUSTIVE, LEGALITY, STATS, LDS_DEPTH>(T, s, lds_stack, spill, spill_depth, st, overflow)) {
  }
  return s.has ? Hit{s.best, s.bestPos} : Hit{-1.0f, -1};
}

#ifndef RX_CHUNK
#define RX_CHUNK 64
#endif
constexpr uint32_t kChunk = RX_CHUNK;  // most queue entries a wave takes from a work counter at a time

#ifndef RX_WORK_GROUPS
#define RX_WORK_GROUPS 8
#endif
constexpr uint32_t kWorkGroups = RX_WORK_GROUPS;  // counters of one work queue
constexpr uint32_t kWorkStride = 32;              // words between them: each on its own 128-byte line
constexpr uint32_t kWorkWords = kWorkGroups * kWorkStride;  // a work queue's counter block (zero at launch)

// A wave's share of a queue [0, n).  Every wave starts on its own chunk of
// neighbouring entries (wave w: [w*ch, (w+1)*ch)), then takes further chunks
// from shared counters over the rest of the queue, and hands a chunk's
// entries to its idle lanes.  Neighbouring entries stay on one wave, and the
// counters keep every wave busy until the queue is drained, however unevenly
// the costs fall over the queue.  A chunk is an even share of the queue per
// wave, capped at kChunk: a short queue is spread thinly over the whole grid,
// starts all of its rays at once and needs no atomics.  The rest of the queue
// is cut into kWorkGroups segments with a counter each; a wave starts on the
// segment of its block's group (blockIdx mod kWorkGroups, the XCD the block
// runs on when blocks are dealt to XCDs in turn) and moves to the next
// segment when one is drained, so the atomics spread over kWorkGroups
// addresses instead of queueing at one.
struct WaveWork {
  uint32_t next = 0, end = 0;  // the wave's chunk [next, end), wave-uniform
  uint32_t ch = 1;             // chunk size
  uint32_t dyn = 0;            // start of the counters' part of the queue
  uint32_t seg = 0;            // entries per counter segment
  uint32_t g = 0, left = 0;    // current segment, segments not yet found drained
  bool drained = false;        // nothing left for this wave
};
__device__ __forceinline__ WaveWork wave_work(uint32_t n) {
  WaveWork w;
  const uint32_t per_block = blockDim.x >> 6, waves = gridDim.x * per_block;
  const uint32_t wid = blockIdx.x * per_block + (threadIdx.x >> 6);
  w.ch = max(1u, min(kChunk, (n + waves - 1) / waves));
  w.next = min(n, wid * w.ch);
  w.end = min(n, w.next + w.ch);
  w.dyn = min(n, waves * w.ch);
  w.seg = (n - w.dyn + kWorkGroups - 1) / kWorkGroups;
  w.g = blockIdx.x % kWorkGroups;
  w.left = w.dyn < n ? kWorkGroups : 0u;
  return w;
}
// Called by every lane of a wave with the wave's ballot of idle lanes
// (nonzero) while !w.drained; returns the entry for this lane if it is idle,
// n if there is none.  work: the queue's kWorkWords counter block.
__device__ __forceinline__ uint32_t wave_take(WaveWork& w, unsigned long long idle, uint32_t n, uint32_t* work) {
  const int lane = threadIdx.x & 63;
  const uint32_t need = (uint32_t)__popcll(idle);
  const uint32_t rank = (uint32_t)__popcll(idle & ((1ull << lane) - 1ull));
  const uint32_t avail = w.end - w.next;
  uint32_t i = w.next + rank;
  if (need > avail) {  // the chunk runs short: the next idle lanes take a new one
    uint32_t base = n, lim = n;
    if (lane == 0) {
      while (w.left > 0) {
        const uint32_t s0 = w.dyn + w.g * w.seg, s1 = min(n, s0 + w.seg);
        if (s0 < s1) {
          const uint32_t off = atomicAdd(work + w.g * kWorkStride, w.ch);
          if (off < s1 - s0) {
            base = s0 + off;
            lim = s1;
            break;
          }
        }
        w.g = (w.g + 1) % kWorkGroups;
        w.left--;
      }
    }
    // lane 0's values, as wave-uniform (scalar) values
    base = __builtin_amdgcn_readfirstlane(base);
    lim = __builtin_amdgcn_readfirstlane(lim);
    w.g = __builtin_amdgcn_readfirstlane(w.g);
    w.left = __builtin_amdgcn_readfirstlane(w.left);
    const uint32_t nend = base < n ? min(lim, base + w.ch) : n;
    if (rank >= avail) i = (base < n && rank - avail < nend - base) ? base + (rank - avail) : n;
    if (base >= n) {
      w.drained = true;
      w.next = w.end = n;
    } else {
      w.end = nend;
      w.next = min(w.end, base + (need - avail));
    }
  } else {
    w.next += need;
  }
  return ((idle >> lane) & 1ull) && i < n ? i : n;
}

// Runs queue entries [0, n) over every wave of the grid (WaveWork): each lane
// holds one ray and advances it a trip at a time, taking a new entry when its
// ray has ended.  fetch(i, o, d, inside, contact, root) loads entry i (root
// starts as T's own and may name another tree of the node array; with SKIP a
// seventh argument receives the triangle id the ray skips); emit(i, hit,
// stats) stores its result.  A wave leaves once the queue is drained and its
// lanes are idle.
template <bool EXHAUSTIVE, bool LEGALITY, bool STATS, int LDS_DEPTH, bool SKIP = false, class Fetch, class Emit>
__device__ __forceinline__ void trace_range(const DevTree& T, uint32_t n, uint32_t* work, uint2* lds_stack,
                                            uint2* spill, int spill_depth, bool* overflow, Fetch fetch, Emit emit) {
  bool active = false;
  uint32_t cur = 0;
  WaveWork w = wave_work(n);
  TraceState ts;
  TraceStats st{0u, 0u};
  for (;;) {
    const unsigned long long idle = __ballot(!active);
    if (idle && !w.drained) {
      const uint32_t i = wave_take(w, idle, n, work);
      if (i < n) {
        cur = i;
        Vec3 o, d;
        int32_t ins = -1, con = -1;
        TreeRoot root = tree_root(T);
        int32_t skip = -1;
        if constexpr (SKIP) fetch(i, o, d, ins, con, root, skip);
        else fetch(i, o, d, ins, con, root);
        active = trace_begin<STATS>(root, ts, o, d, ins, con, &st);
        ts.skip = skip;
        if (!active) emit(i, Hit{-1.0f, -1}, st);  // missed the scene box
      }
    }
    if (w.drained && !__any(active)) break;
    if (active &&
        trace_trip<EXHAUSTIVE, LEGALITY, STATS, LDS_DEPTH, false, SKIP>(T, ts, lds_stack, spill, spill_depth, &st,
                                                                        overflow)) {
      emit(cur, ts.has ? Hit{ts.best, ts.bestPos} : Hit{-1.0f, -1}, st);
      active = false;
    }
  }
}

// ------------------------------------------------------------------ packets
// Nearest hit of up to 64 coherent rays, one per lane (camera rays of the
// same pixels), traced as one packet: the wave walks the tree once, node by
// node, carrying the 64-bit mask of the lanes whose rays reach the node.  At
// an interior node every lane of the mask classifies both child boxes as the
// per-lane traversal does (filter, exact predicate when undecided, culling by
// its own best distance); the children that some lane enters are visited,
// the one more lanes find nearer first, the other pushed with its mask on a
// wave stack in LDS.  At a leaf the lanes of the mask test its triangles.
// A lane thus visits every subtree whose boxes pass for its ray and are not
// culled by its own best hit, so its result is the per-lane traversal's: the
// winner rule (nearest, ties to the larger leaf, first triangle in a leaf)
// does not depend on the visiting order.  Node, leaf and triangle addresses
// are wave-uniform, so one fetch serves the packet.  Every lane of the wave
// runs this function (invalid lanes with valid = false); `wstack` is the
// wave's max_sp entries (a node and its mask per entry).  LEGALITY: each
// lane's hits must pass its ray's condition (inside, contact), as in
// trace_trip (packets of later bounces hit the same triangles but, their
// directions scattered, walk the union of their paths: slower than lanes).
template <bool STATS, bool LEGALITY>
__device__ __forceinline__ void trace_packet(const DevTree& T, const TreeRoot& root, bool valid, Vec3 o, Vec3 d,
                                             int32_t inside, int32_t contact, uint4* wstack, int max_sp,
                                             TraceState& s, TraceStats* st, bool* overflow) {
  const float INF = __builtin_inff();
  const int lane = threadIdx.x & 63;
  const bool in_root = valid && trace_begin<STATS>(root, s, o, d, inside, contact, st);
  unsigned long long mask = __ballot(in_root);
  int32_t node = root.ref;
  int sp = 0;
  for (;;) {
    while (mask && node >= 0) {
      const float4* np = T.nodes + 4 * (size_t)node;
      const float4 a = np[0], b = np[1], c = np[2];
      const int4 r = *reinterpret_cast<const int4*>(np + 3);
      bool vl = false, vr = false;
      float tl = 0.0f, tr = 0.0f;
      if ((mask >> lane) & 1ull) {
        const float llo[3] = {a.x, a.y, a.z}, lhi[3] = {a.w, b.x, b.y};
        const float rlo[3] = {b.z, b.w, c.x}, rhi[3] = {c.y, c.z, c.w};
        if (STATS) st->boxes += 2;
        const float thr = s.has ? fmaf(s.best, 1.001f, s.cm) : INF;
        bool hl, hr;
        if (s.fast) {
          bool ml, cl, mr, cr;
          slab_verdicts(llo, lhi, s.oi, s.inv, s.beta, tl, ml, cl);
          slab_verdicts(rlo, rhi, s.oi, s.inv, s.beta, tr, mr, cr);
          hl = cl || (!ml && tl <= thr && box_hit(llo, lhi, s.o, s.d));
          hr = cr || (!mr && tr <= thr && box_hit(rlo, rhi, s.o, s.d));
        } else {
          hl = box_hit(llo, lhi, s.o, s.d);
          hr = box_hit(rlo, rhi, s.o, s.d);
          tl = slab_tnear(llo, lhi, s.oi, s.d, s.inv);
          tr = slab_tnear(rlo, rhi, s.oi, s.d, s.inv);
        }
        vl = hl && tl <= thr;
        vr = hr && tr <= thr;
      }
      const unsigned long long ML = __ballot(vl), MR = __ballot(vr);
      const int32_t left = __builtin_amdgcn_readfirstlane(r.x), right = __builtin_amdgcn_readfirstlane(r.y);
      if (ML && MR) {
        const unsigned long long preferR = __ballot(vl && vr && tr < tl);
        const bool rightFirst = 2 * __popcll(preferR) > __popcll(ML & MR);
        const unsigned long long farM = rightFirst ? ML : MR;
        if (sp >= max_sp) {  // deeper than the wave stack: the caller reports an overflow
          *overflow = true;
          s.has = false;
          return;
        }
        wstack[sp] = make_uint4((uint32_t)(rightFirst ? left : right), (uint32_t)farM, (uint32_t)(farM >> 32), 0u);
        sp++;
        node = rightFirst ? right : left;
        mask = rightFirst ? MR : ML;
      } else if (ML) {
        node = left;
        mask = ML;
      } else if (MR) {
        node = right;
        mask = MR;
      } else {
        node = INT32_MIN;
      }
    }
    if (mask && node != INT32_MIN) {  // a leaf: its triangles for the lanes of the mask
      const int32_t leaf = ~node;
      const int32_t p0 = __builtin_amdgcn_readfirstlane(T.leaf_start[leaf]);
      const int32_t p1 = __builtin_amdgcn_readfirstlane(T.leaf_start[leaf + 1]);
      const bool act = (mask >> lane) & 1ull;
      for (int32_t p = p0; p < p1; p++) {
        const float4 ta = T.tv0[p], tb = T.tv1[p], tc = T.tv2[p];
        if (act) {
          if (STATS) st->tris++;
          const float dist = tri_dist(xyz(ta), xyz(tb), xyz(tc), s.o, s.d);
          if (dist >= 0 && (!s.has || dist < s.best || (dist == s.best && leaf > s.bestLeaf)) &&
              (!LEGALITY || legal(s.d, xyz(T.tn[p]), __float_as_int(tb.w), s.inside, s.contact))) {
            s.has = true;
            s.best = dist;
            s.bestLeaf = leaf;
            s.bestPos = p;
          }
        }
      }
    }
    if (sp == 0) break;
    sp--;
    const uint4 e = wstack[sp];
    node = __builtin_amdgcn_readfirstlane((int32_t)e.x);
    mask = ((unsigned long long)__builtin_amdgcn_readfirstlane(e.z) << 32) | __builtin_amdgcn_readfirstlane(e.y);
  }
}

// ------------------------------------------------------------------ RNG
// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11): counter-based, one
// independent stream per (pixel, sample) keyed by the seed.
struct U4 {
  uint32_t x, y, z, w;
};
__host__ __device__ __forceinline__ U4 philox(U4 c, uint32_t k0, uint32_t k1) {
  const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u, W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
#pragma unroll
  for (int i = 0; i < 10; i++) {
    // each 32 x 32 -> 64-bit product as one value: one v_mad_u64_u32 on the
    // device for both halves (a v_mul_hi_u32 and a v_mul_lo_u32 otherwise)
    const uint64_t p0 = (uint64_t)M0 * c.x, p1 = (uint64_t)M1 * c.z;
    const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
    const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
    c = U4{hi1 ^ c.y ^ k0, lo1, hi0 ^ c.w ^ k1, lo0};
    k0 += W0;
    k1 += W1;
  }
  return c;
}
// 24-bit uniform in [0, 1 - 2^-24], the reference's float form (src/RandGenerator.cpp:21-25)
__host__ __device__ __forceinline__ float u01(uint32_t u) { return (float)(u >> 8) * (1.0f / 16777216.0f); }

}  // namespace rx
