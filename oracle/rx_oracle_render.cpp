// Oracle renderer: restatement of RenderX's CPU PathTracer and PathTracerTest
// (NEE + MIS) with the scene objects, materials, medium and camera they use.
//
// TEST INFRASTRUCTURE ONLY: this is the checker (and the timed CPU baseline)
// for the HIP wavefront renderer in renderx_amd; it is never part of it.
//
// Differences from the reference, all deliberate and statistical-neutral:
//  * random numbers: one independent stream per (pixel, sample) instead of the
//    racy global MT19937 (src/RandGenerator.cpp:9) and libc rand()
//    (src/SceneObject.cpp:377); floats keep the reference's 24-bit form
//    (src/RandGenerator.cpp:21-25);
//  * explicit spp instead of a time budget (include/PathTracer.h:7,14);
//  * the film keeps FP64 sum and sum of squares per pixel instead of the FP32
//    running mean (src/PathTracer.cpp:43,67);
//  * the wasted intersection inside SceneObject::emit (src/SceneObject.cpp:198)
//    is skipped: its result is never read by PathTracerTest.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "rx_oracle_geom.h"
#include "rx_oracle_scene.h"

namespace orc {

static const float PI_F = (float)M_PI;
// Planted mistakes for the tests that show the oracle's pins catch them
// (orc_set_mutation); 0 in every other use.
int g_mutation = 0;
// traversal queries issued by this thread (camera, bounce, shadow, containment)
static thread_local uint64_t g_rays = 0;
static const float EPS = 1e-5f;          // EPSILON (include/macros.h:7)
static const float COS_MIN = 1e-6f;      // COS_TERM_MIN (include/macros.h:9)

// ---------------------------------------------------------------- RNG
// SplitMix64 stream keyed by (seed, pixel, sample).  Test RNG only.
struct Rng {
  uint64_t s;
  Rng(uint64_t seed, uint64_t pixel, uint64_t sample) {
    s = seed ^ (pixel * 0x9E3779B97F4A7C15ull) ^ (sample * 0xD1B54A32D192ED03ull);
    next(); next();
  }
  uint64_t next() {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }
  uint32_t u32() { return (uint32_t)(next() >> 32); }
  // RNG::randFloat: low 24 bits / 2^24 (src/RandGenerator.cpp:21-25)
  float f() { return (u32() & 0xffffff) / float(1 << 24); }
};

// MT19937 as RandGenerator uses it (src/RandGenerator.cpp:11-65), for the
// single-stream reference mode and its known-answer test.
struct MT {
  uint32_t mt[624];
  int mti;
  explicit MT(uint32_t seed = 5489u) {
    mt[0] = seed;
    for (mti = 1; mti < 624; mti++) mt[mti] = 1812433253u * (mt[mti - 1] ^ (mt[mti - 1] >> 30)) + mti;
  }
  uint32_t next() {
    static const uint32_t mag01[2] = {0u, 0x9908b0dfu};
    uint32_t y;
    if (mti >= 624) {
      int k;
      for (k = 0; k < 624 - 397; k++) {
        y = (mt[k] & 0x80000000u) | (mt[k + 1] & 0x7fffffffu);
        mt[k] = mt[k + 397] ^ (y >> 1) ^ mag01[y & 1u];
      }
      for (; k < 623; k++) {
        y = (mt[k] & 0x80000000u) | (mt[k + 1] & 0x7fffffffu);
        mt[k] = mt[k + (397 - 624)] ^ (y >> 1) ^ mag01[y & 1u];
      }
      y = (mt[623] & 0x80000000u) | (mt[0] & 0x7fffffffu);
      mt[623] = mt[396] ^ (y >> 1) ^ mag01[y & 1u];
      mti = 0;
    }
    y = mt[mti++];
    y ^= (y >> 11);
    y ^= (y << 7) & 0x9d2c5680u;
    y ^= (y << 15) & 0xefc60000u;
    y ^= (y >> 18);
    return y;
  }
};

// ---------------------------------------------------------------- helpers
static inline float maxc(V3 c) { return std::max(c.x, std::max(c.y, c.z)); }
// luminance y() (include/smallFuncs.h:41-44)
static inline float lum(V3 v) { return 0.212671f * v[0] + 0.715160f * v[1] + 0.072169f * v[2]; }

// LocalFrame::buildFromNormal / toWorld (include/LocalFrame.h:13-32)
struct Frame {
  V3 s, t, n;
};
static Frame frame_from(V3 normal) {
  Frame f;
  f.n = normal;
  V3 tmpT = (std::fabs(f.n.z) > 0.99f) ? V3(1, 0, 0) : V3(0, 0, 1);
  f.s = normalize(cross(f.n, tmpT));
  f.t = normalize(cross(f.s, f.n));
  return f;
}
static inline V3 to_world(const Frame& f, V3 a) { return f.s * a.x + f.n * a.y + f.t * a.z; }

// RandGenerator::genHemiCosDirection (src/RandGenerator.cpp:152-166)
static V3 hemi_cos(Rng& g, float expTerm) {
  float dx = g.f(), dy = g.f();
  float u1 = 2.f * M_PI * dx;
  float u2 = std::pow(dy, 1.f / (expTerm + 1.f));
  float u3 = std::sqrt(1.f - u2 * u2);
  return normalize(V3(std::cos(u1) * u3, u2, std::sin(u1) * u3));
}
// CosineSphericalSampler::getProbDensity (include/CosineSphericalSampler.h:29-34)
static inline float cos_pdf(const Frame& f, V3 dir, float coeff) {
  float c = dot(f.n, dir);
  float cosTheta = c > 0.f ? c : 0.f;
  return (float)(std::pow(cosTheta, coeff) * (coeff + 1.f) / (2.f * M_PI));
}

// rotMat(axis, angle) * vec4(v, 0) (include/smallFuncs.h:11-24; the transpose
// of the column-major constructor is the right-handed rotation by +angle)
static V3 rotate(V3 axis, float angle, V3 v) {
  if (g_mutation == MUT_ROTATION_SIGN) angle = -angle;
  float c = std::cos(angle), s = std::sin(angle), _c = 1 - c;
  float x = axis.x, y = axis.y, z = axis.z;
  float R[3][3] = {{c + _c * x * x, _c * x * y - s * z, _c * x * z + s * y},
                   {_c * x * y + s * z, c + _c * y * y, _c * y * z - s * x},
                   {_c * x * z - s * y, _c * y * z + s * x, c + _c * z * z}};
  V3 r;
  for (int i = 0; i < 3; i++) r[i] = v.x * R[i][0] + v.y * R[i][1] + v.z * R[i][2] + 0.0f * 0.0f;
  return r;
}

// The Snell / Fresnel step of RefractiveMaterial::scatter
// (src/RefractiveMaterial.cpp:189-232) and SceneVPMObject's boundary crossings
// (src/SceneVPMObject.cpp:44-75,176-209): theta = acos(d.n) (clamped to
// [-1, 1] in RefractiveMaterial only), sin(phi) = n_i / n_t sin(theta); false
// on total internal reflection; otherwise the normal rotated by phi about
// normalize(n x d) (phi -> pi - phi when the ray runs against the normal) and
// the unpolarised Fresnel reflectance er = (r_s^2 + r_p^2) / 2.
static bool refr_dir(V3 d, V3 normal, float current_n, float next_n, bool clamp_cos, V3& dir, float& er) {
  if (g_mutation == MUT_INVERTED_RATIO) std::swap(current_n, next_n);
  float cd = dot(d, normal);
  if (clamp_cos) cd = cd < -1 ? -1.f : cd > 1 ? 1.f : cd;
  float theta = std::acos(cd);
  float sin_phi = current_n / next_n * std::sin(theta);
  if (sin_phi >= 1) return false;
  float phi = std::asin(sin_phi);
  if (theta > M_PI / 2) phi = M_PI - phi;
  V3 axis = normalize(cross(normal, d));
  dir = normalize(rotate(axis, phi, normal));
  float cos_theta = std::fabs(std::cos(theta)), cos_phi = std::fabs(std::cos(phi));
  float esr = std::pow(std::fabs(current_n * cos_theta - next_n * cos_phi) / (current_n * cos_theta + next_n * cos_phi), 2.f);
  float epr = std::pow(std::fabs(next_n * cos_theta - current_n * cos_phi) / (next_n * cos_theta + current_n * cos_phi), 2.f);
  er = (esr + epr) / 2;
  return true;
}

// ---------------------------------------------------------------- Ray record
enum { DEFINITE = 0, RANDOM = 1 };
static const int NONE = -1, CAMERA = -2;

struct PRay {  // class Ray (include/Ray.h:8-51)
  V3 origin, direction, color{1, 1, 1};
  float directionProb = 1, originProb = 1;
  int insideObject = NONE, contactObject = NONE;
  unsigned contactTri = 0;
  int intersectObject = NONE;
  unsigned intersectTri = 0;
  float intersectDist = -1;
  int directionSampleType = RANDOM, originSampleType = DEFINITE;
};

struct Ctx {
  const SceneO& sc;
  Rng& g;
  unsigned maxDepth;
  unsigned rrStart;
  bool light = false;  // scatter's fixIsLight: a light subpath of the BidirectionalPathTracer
};

static inline bool is_emissive(const SceneO& sc, int o) { return o >= 0 && sc.objs[o].type == OBJ_EMISSIVE; }
static inline bool is_volumetric(const SceneO& sc, int o) { return o >= 0 && sc.objs[o].type == OBJ_VPM; }
static inline float refr_coeff(const SceneO& sc, int o) {
  if (o < 0) return 1;
  int t = sc.objs[o].type;
  return (t == OBJ_REFRACTIVE || t == OBJ_VPM) ? sc.objs[o].ior : 1.f;  // getRefrCoeff overrides
}
// SceneObject::getWorldNormal for face-normal meshes; Camera::getWorldNormal = front
static V3 world_normal(const SceneO& sc, int obj, unsigned localTri) {
  if (obj == CAMERA) return sc.cam_front;
  if (obj < 0) return V3(0, 0, 0);
  return sc.tri_normal[sc.objs[obj].tri_begin + localTri];
}
static inline V3 contact_normal(const SceneO& sc, const PRay& r) {  // Ray::getContactNormal
  return r.contactObject != NONE ? world_normal(sc, r.contactObject, r.contactTri) : V3(0, 0, 0);
}
static float cosine_term(const SceneO& sc, const PRay& r) {  // Ray::getCosineTerm (src/Ray.cpp:71-78)
  float d = 1.f;
  bool hasCos = r.contactObject == CAMERA || (r.contactObject >= 0 && sc.objs[r.contactObject].type != OBJ_VPM);
  if (r.contactObject != NONE && hasCos && length(r.direction) > 0.5f)
    d = std::fabs(dot(contact_normal(sc, r), r.direction));
  return std::max(d, COS_MIN);
}
static V3 radiance_decay(const SceneO& sc, const PRay& r, float dist) {  // Ray::getRadianceDecay
  if (r.insideObject < 0) return V3(1, 1, 1);
  const ObjO& o = sc.objs[r.insideObject];
  if (o.type == OBJ_VPM)  // SceneVPMObject::transmittance
    return V3(std::exp(-o.sigma_t.x * dist), std::exp(-o.sigma_t.y * dist), std::exp(-o.sigma_t.z * dist));
  if (o.type == OBJ_REFRACTIVE)  // RefractiveMaterial::getRadianceDecay
    return V3(std::pow(o.c2.x, dist), std::pow(o.c2.y, dist), std::pow(o.c2.z, dist));
  return V3(1, 1, 1);
}

// Scene::intersect with NoSelfIntersectionCondition built from the ray
static float scene_intersect(const SceneO& sc, const PRay& r, bool useCond, int& obj, unsigned& localTri) {
  Ray kr{r.origin, r.direction};
  Legality L{sc.tri_obj.data(), sc.tri_normal.data(), r.insideObject, r.contactObject};
  unsigned tid = 0;
  g_rays++;
  float d = intersect_exhaustive(sc.tree, kr, tid, useCond ? &L : nullptr, nullptr);
  if (d >= 0) {
    obj = sc.tri_obj[tid];
    localTri = tid - sc.objs[obj].tri_begin;
  }
  return d;
}

// Scene::findInsideObject (src/Scene.cpp:354-390)
static int find_inside(const SceneO& sc, V3 o, V3 d, int current) {
  Ray front{o, d}, back{o, -d};
  for (int i = (int)sc.objs.size() - 1; i >= 0; i--) {
    if (current == i || !is_volumetric(sc, i)) continue;
    const Tree& t = sc.obj_tree[i];
    unsigned tid = 0;
    g_rays++;
    float dist = intersect_exhaustive(t, back, tid, nullptr, nullptr);
    if (dist < 1e-6f) continue;
    if (dot(sc.tri_normal[sc.objs[i].tri_begin + tid], back.d) < 1e-6f) continue;
    g_rays++;
    dist = intersect_exhaustive(t, front, tid, nullptr, nullptr);
    if (dist < 1e-6f) continue;
    if (dot(sc.tri_normal[sc.objs[i].tri_begin + tid], front.d) < 1e-6f) continue;
    return i;
  }
  return NONE;
}

// ---------------------------------------------------------------- materials
// DiffuseMaterial::scatter (src/DiffuseMaterial.cpp:4-77)
static PRay diffuse_scatter(Ctx& c, int object, V3 rho, const PRay& in, bool russian) {
  PRay out;
  out.origin = in.origin + in.direction * in.intersectDist;
  Frame lf = frame_from(world_normal(c.sc, in.intersectObject, in.intersectTri));
  V3 normal = lf.n;
  out.insideObject = in.insideObject;
  out.intersectObject = NONE;
  out.contactObject = object;
  out.contactTri = in.intersectTri;
  if (dot(in.direction, normal) > 0) {  // leaving
    out.direction = in.direction;
    out.directionProb = 1;
    out.color = V3(1, 1, 1);
    out.originSampleType = DEFINITE;
    out.directionSampleType = DEFINITE;
    return out;
  }
  out.direction = to_world(lf, hemi_cos(c.g, 1.f));
  float p = maxc(rho);
  if (c.g.f() < p || !russian) {
    // DiffuseBSDF::evaluate (include/DiffuseBSDF.h:18-24)
    out.color = (dot(lf.n, in.direction) <= 0 && dot(lf.n, out.direction) >= 0) ? rho / PI_F : V3(0, 0, 0);
    out.directionProb = russian ? p * cos_pdf(lf, out.direction, 1.f) : cos_pdf(lf, out.direction, 1.f);
  } else {
    out.direction = V3(0, 0, 0);
    out.color = V3(0, 0, 0);
    out.directionProb = 1 - p;
  }
  out.originSampleType = DEFINITE;
  out.directionSampleType = RANDOM;
  return out;
}

static inline V3 mirror(V3 n, V3 d) { return -dot(n, d) * n * 2.0f + d; }

// GlossyBSDF::evaluate (include/GlossyBSDF.h:24-33)
static V3 glossy_eval(const Frame& lf, V3 in, V3 out, V3 color, float coeff) {
  if (dot(lf.n, in) <= 0 && dot(lf.n, out) >= 0) {
    V3 r = g_mutation == MUT_GLOSSY_AXIS ? -in : mirror(lf.n, in);
    if (dot(r, out) > 0) return color * std::pow(dot(r, out), coeff);
  }
  return V3(0, 0, 0);
}

// GlossyMaterial::scatter (src/GlossyMaterial.cpp:4-57); RR is unconditional there
static PRay glossy_scatter(Ctx& c, int object, V3 ks, float coeff, const PRay& in) {
  PRay out;
  V3 position = in.origin + in.direction * in.intersectDist;
  Frame lf = frame_from(world_normal(c.sc, in.intersectObject, in.intersectTri));
  V3 normal = lf.n;
  out.insideObject = in.insideObject;
  out.origin = position;
  out.contactObject = object;
  out.contactTri = in.intersectTri;
  Frame olf = frame_from(mirror(normal, in.direction));
  out.direction = to_world(olf, hemi_cos(c.g, coeff));
  out.color = glossy_eval(lf, in.direction, out.direction, ks, coeff);
  out.directionProb = cos_pdf(olf, out.direction, coeff);
  float p = maxc(ks);
  out.originSampleType = DEFINITE;
  out.directionSampleType = RANDOM;
  if (length(out.color) == 0) {
    out.direction = V3(0, 0, 0);
    out.color = V3(0, 0, 0);
    out.directionProb = 1;
    return out;
  }
  if (c.g.f() < p) {
    out.directionProb *= p;
  } else {
    out.direction = V3(0, 0, 0);
    out.color = V3(0, 0, 0);
    out.directionProb = 1 - p;
  }
  return out;
}

static PRay scatter(Ctx& c, int object, const PRay& in, bool russian);

// RefractiveMaterial::scatter (src/RefractiveMaterial.cpp:132-269)
static PRay refractive_scatter(Ctx& c, int object, const PRay& in, bool russian) {
  const ObjO& O = c.sc.objs[object];
  PRay out;
  V3 position = in.origin + in.direction * in.intersectDist;
  if (in.intersectObject == NONE) {
    out.direction = V3(0, 0, 0);
    out.contactObject = NONE;
    out.directionSampleType = DEFINITE;
    return out;
  }
  Frame lf = frame_from(world_normal(c.sc, in.intersectObject, in.intersectTri));
  V3 normal = lf.n;
  out.color = O.c1;  // surfColor
  out.origin = position;
  out.direction = in.direction;
  V3 reflDir = mirror(normal, in.direction);
  out.contactObject = object;
  out.contactTri = in.intersectTri;
  out.directionSampleType = DEFINITE;
  if (in.intersectObject != object && in.insideObject == object)
    return scatter(c, in.intersectObject, in, russian);
  int currentInside = in.insideObject;
  int outSide = object;
  if (in.insideObject == object) outSide = find_inside(c.sc, out.origin, out.direction, object);
  float current_n = refr_coeff(c.sc, currentInside);
  float next_n = refr_coeff(c.sc, outSide);
  out.intersectObject = NONE;
  out.directionProb = 1;
  float er;
  if (!refr_dir(in.direction, normal, current_n, next_n, true, out.direction, er)) {
    // total internal reflection terminates the path
    out.direction = V3(0, 0, 0);
    out.contactObject = NONE;
    out.directionSampleType = DEFINITE;
    return out;
  }
  float p = er < 0.f ? 0.f : er > 1.f ? 1.f : er;
  if (c.g.f() < p) {
    out.direction = reflDir;
    out.color = out.color * (er / cosine_term(c.sc, out));
    out.directionProb = p;
    out.insideObject = in.insideObject;
  } else {
    if (!c.light) out.color = out.color * ((current_n * current_n) / (next_n * next_n));  // eye paths only
    out.color = out.color * ((1 - er) / cosine_term(c.sc, out));
    out.directionProb = 1 - p;
    out.insideObject = outSide;
  }
  out.direction = normalize(out.direction);
  return out;
}

// SceneEmissiveObject::scatter (src/SceneEmissiveObject.cpp:4-19)
static PRay emissive_scatter(Ctx& c, int object, const PRay& in) {
  PRay out;
  V3 position = in.origin + in.direction * in.intersectDist;
  out.intersectObject = NONE;
  out.insideObject = NONE;
  out.contactObject = object;
  out.contactTri = in.intersectTri;
  out.origin = position;
  out.direction = V3(0, 0, 0);
  out.color = dot(in.direction, contact_normal(c.sc, out)) <= 0 ? c.sc.objs[object].c1 : V3(0, 0, 0);
  out.directionProb = 1;
  out.directionSampleType = RANDOM;
  out.originSampleType = DEFINITE;
  return out;
}

// HG phase function value (include/HGPhaseFunc.h:11-17 / HGPhaseSampler.h:63-69)
static inline float hg_value(float g, float cosTheta) {
  float temp = 1 + g * g - 2 * g * cosTheta;
  return (float)(1 / (4 * M_PI) * (1 - g * g) / (temp * std::sqrt(temp)));
}

// index-matched boundary crossing used by SceneVPMObject for enter and exit
static bool vpm_refract(Ctx& c, const PRay& in, V3 normal, float current_n, float next_n, PRay& out,
                        float& er, V3& reflDir) {
  reflDir = normalize(mirror(normal, in.direction));
  return refr_dir(in.direction, normal, current_n, next_n, false, out.direction, er);
}

// SceneVPMObject::p_medium (include/SceneVPMObject.h:39-47): density of
// scattering at distance s, y(sigma_t e^{-sigma_t s})
static float p_medium(V3 sigma_t, float s) {
  if (g_mutation == MUT_P_MEDIUM_NO_SIGMA)
    return lum(V3(std::exp(-sigma_t.x * s), std::exp(-sigma_t.y * s), std::exp(-sigma_t.z * s)));
  V3 pm(sigma_t.x * std::exp(-sigma_t.x * s), sigma_t.y * std::exp(-sigma_t.y * s), sigma_t.z * std::exp(-sigma_t.z * s));
  return lum(pm);
}
// SceneVPMObject::P_surface (include/SceneVPMObject.h:27-37): probability of
// reaching the surface at distance d, y(e^{-sigma_t d})
static float P_surface(V3 sigma_t, float d) {
  if (g_mutation == MUT_P_SURFACE_COMPLEMENT)
    return lum(V3(1 - std::exp(-sigma_t.x * d), 1 - std::exp(-sigma_t.y * d), 1 - std::exp(-sigma_t.z * d)));
  return lum(V3(std::exp(-sigma_t.x * d), std::exp(-sigma_t.y * d), std::exp(-sigma_t.z * d)));
}
static float p_medium(const ObjO& o, float s) { return p_medium(o.sigma_t, s); }
static float P_surface(const ObjO& o, float d) { return P_surface(o.sigma_t, d); }

// SceneVPMObject::scatter (src/SceneVPMObject.cpp:4-251)
static PRay vpm_scatter(Ctx& c, int self, const PRay& in, bool russian) {
  const ObjO& O = c.sc.objs[self];
  PRay out;
  out.directionSampleType = RANDOM;
  // HomoMediaDistSampler::sampleDist (include/HomoMediaDistSampler.h:9-11)
  float sampDist = (float)(-std::log(1.0 - c.g.f()) / lum(O.sigma_t));
  bool go_in_vol = (in.intersectObject == self) && (in.insideObject != self);
  bool be_in_vol = (in.insideObject == self);
  bool out_of_vol = (sampDist >= in.intersectDist);
  if (go_in_vol) {
    V3 position = in.origin + in.direction * in.intersectDist;
    V3 normal = world_normal(c.sc, in.intersectObject, in.intersectTri);
    out.origin = position;
    out.direction = in.direction;
    float current_n = refr_coeff(c.sc, in.insideObject), next_n = refr_coeff(c.sc, self);
    out.intersectObject = NONE;
    out.color = V3(1, 1, 1);
    out.directionProb = 1;
    out.contactObject = self;
    out.contactTri = in.intersectTri;
    float er;
    V3 reflDir;
    if (!vpm_refract(c, in, normal, current_n, next_n, out, er, reflDir)) {
      out.direction = V3(0, 0, 0);
      out.contactObject = NONE;
      out.directionSampleType = DEFINITE;
      return out;
    }
    float p = er < 0.f ? 0.f : er > 1.f ? 1.f : er;
    if (c.g.f() < p) {
      out.direction = reflDir;
      out.color = out.color * (er / cosine_term(c.sc, out));
      out.directionProb = p;
      out.insideObject = in.insideObject;
      out.directionSampleType = DEFINITE;
    } else {
      if (!c.light) out.color = out.color * ((current_n * current_n) / (next_n * next_n));  // fixIsLight
      out.color = out.color * ((1 - er) / cosine_term(c.sc, out));
      out.directionProb = 1 - p;
      out.contactObject = out.insideObject = self;
      out.directionSampleType = DEFINITE;
    }
    out.direction = normalize(out.direction);
    return out;
  }
  if (be_in_vol && !out_of_vol) {  // scatter inside the medium
    Frame lf = frame_from(in.direction);
    out.origin = in.origin + in.direction * sampDist;
    // HGPhaseSampler::genSample (include/HGPhaseSampler.h:13-34)
    float e1 = c.g.f(), e2 = c.g.f();
    float g = O.g, cosTheta;
    if (std::fabs(g) < EPS) {
      cosTheta = 1 - 2 * e1;
    } else {
      float sqrTerm = (1 - g * g) / (1 - g + 2 * g * e1);
      cosTheta = (1 + g * g - sqrTerm * sqrTerm) / (2 * g);
    }
    float sinTheta = std::sqrt(1 - cosTheta * cosTheta);
    float sinPhi = (float)std::sin(2.0 * M_PI * e2), cosPhi = (float)std::cos(2.0 * M_PI * e2);
    out.direction = to_world(lf, V3(sinTheta * cosPhi, cosTheta, sinTheta * sinPhi));
    out.color = V3(1, 1, 1) * hg_value(g, dot(in.direction, out.direction));
    out.insideObject = self;
    out.contactTri = in.intersectTri;
    float albedo = lum(O.c1) / lum(O.sigma_t);
    float rander = c.g.f();
    if (rander < albedo || !russian) {
      out.contactObject = NONE;
      float oPdfW = hg_value(g, dot(lf.n, out.direction));
      out.directionProb = russian ? albedo * oPdfW : oPdfW;
      out.originProb = p_medium(O, sampDist);
      out.directionSampleType = RANDOM;
      out.color = out.color * O.c1;  // *= ds
    } else {
      out.direction = V3(0, 0, 0);
      out.color = V3(0, 0, 0);
      out.directionProb = 1;
      out.originProb = p_medium(O, sampDist);
      out.insideObject = self;
      out.contactObject = NONE;
      out.directionSampleType = RANDOM;
    }
    return out;
  }
  if (be_in_vol && out_of_vol) {  // reach a surface before the sampled distance
    out = in;
    out.direction = in.direction;
    out.origin = in.origin + in.intersectDist * in.direction;
    out.contactObject = in.intersectObject;
    out.contactTri = in.intersectTri;
    out.insideObject = self;
    out.directionProb = 1;
    out.color = V3(1, 1, 1);
    if (in.intersectObject == self) {  // leave through the own boundary
      V3 normal = world_normal(c.sc, in.intersectObject, in.intersectTri);
      int outSide = find_inside(c.sc, out.origin, out.direction, self);
      float current_n = refr_coeff(c.sc, self), next_n = refr_coeff(c.sc, outSide);
      out.intersectObject = NONE;
      float er;
      V3 reflDir;
      if (!vpm_refract(c, in, normal, current_n, next_n, out, er, reflDir)) {
        out.direction = V3(0, 0, 0);
        out.contactObject = NONE;
        out.directionSampleType = DEFINITE;
        return out;
      }
      float p = er < 0.f ? 0.f : er > 1.f ? 1.f : er;
      if (c.g.f() < p) {
        out.direction = reflDir;
        out.color = out.color * (er / cosine_term(c.sc, out));
        out.directionProb = p;
        out.originProb = P_surface(O, in.intersectDist);
        out.insideObject = in.insideObject;
        out.directionSampleType = DEFINITE;
      } else {
        if (!c.light) out.color = out.color * ((current_n * current_n) / (next_n * next_n));  // fixIsLight
        out.color = out.color * ((1 - er) / cosine_term(c.sc, out));
        out.directionProb = 1 - p;
        out.originProb = P_surface(O, in.intersectDist);
        out.insideObject = outSide;
        out.directionSampleType = DEFINITE;
      }
      out.direction = normalize(out.direction);
    } else {  // a surface inside the medium scatters
      out.contactObject = NONE;
      out.intersectDist = 0;
      if (in.intersectObject != NONE) out = scatter(c, in.intersectObject, out, russian);
      out.originProb *= P_surface(O, in.intersectDist);
    }
    return out;
  }
  return out;
}

// SceneObject::scatter / overrides: picks a material uniformly (Phong: glossy
// then diffuse lobe, include/ScenePhongObject.h:15-22)
static PRay scatter(Ctx& c, int object, const PRay& in, bool russian) {
  const ObjO& O = c.sc.objs[object];
  switch (O.type) {
    case OBJ_DIFFUSE: return diffuse_scatter(c, object, O.c1, in, russian);
    case OBJ_GLOSSY: return glossy_scatter(c, object, O.c1, O.expo, in);
    case OBJ_PHONG:
      if (c.g.u32() % 2 == 0) return glossy_scatter(c, object, O.c1, O.expo, in);
      return diffuse_scatter(c, object, O.c2, in, russian);
    case OBJ_REFRACTIVE: return refractive_scatter(c, object, in, russian);
    case OBJ_EMISSIVE: return emissive_scatter(c, object, in);
    case OBJ_VPM: return vpm_scatter(c, object, in, russian);
  }
  return PRay();
}

// ---------------------------------------------------------------- path
// MCRenderer::samplePathIter (src/MCRenderer.cpp:151-267), iterative form
static void sample_path(Ctx& c, PRay start, std::vector<PRay>& path) {
  path.clear();
  PRay prev = start;
  for (unsigned depth = 0;; depth++) {
    PRay term;
    term.origin = prev.origin;
    term.direction = V3(0, 0, 0);
    term.color = V3(0, 0, 0);
    term.directionProb = 1;
    term.directionSampleType = DEFINITE;
    path.push_back(prev);
    bool russian = depth >= c.rrStart;
    PRay next;
    if (prev.insideObject != NONE)
      next = scatter(c, prev.insideObject, prev, russian);
    else if (prev.intersectObject != NONE)
      next = scatter(c, prev.intersectObject, prev, russian);
    else {
      path.push_back(term);
      return;
    }
    if (length(next.direction) > 0.5f && next.insideObject == NONE && next.contactObject != NONE) {
      // shading-normal light-leak guard; geometric == shading normal here
      V3 geoN = contact_normal(c.sc, next), shdN = geoN;
      V3 wi = -prev.direction, wo = next.direction;
      float wiG = dot(geoN, wi), woG = dot(geoN, wo), wiS = dot(shdN, wi), woS = dot(shdN, wo);
      if (wiG * wiS <= 0 || woG * woS <= 0) {
        next.direction = V3(0, 0, 0);
        next.color = V3(0, 0, 0);
      }
    }
    if (length(next.direction) < 0.5f) {
      path.push_back(next);
      return;
    }
    if (depth + 1 > c.maxDepth) {
      path.push_back(term);
      return;
    }
    int obj = NONE;
    unsigned lt = 0;
    float dist = scene_intersect(c.sc, next, true, obj, lt);
    if (dist < 0) {
      path.push_back(next);
      path.push_back(term);
      return;
    }
    next.intersectObject = obj;
    next.intersectTri = lt;
    next.intersectDist = dist;
    prev = next;
  }
}

// Camera::generateRay / generateRayBase (src/Camera.cpp:38-47,4-36)
static PRay camera_ray(Ctx& c, unsigned pixel) {
  const CamO& cam = c.sc.cam;
  int x = pixel % cam.width, y = pixel / cam.width;
  float xs = (float)x + c.g.f();
  float ys = (float)y + c.g.f();
  V3 front = c.sc.cam_front, right = c.sc.cam_right, orthoUp = c.sc.cam_up;
  float xp = (float)(xs - (cam.width / 2.0));
  float yp = (float)((cam.height / 2.0) - ys);
  PRay r;
  r.direction = normalize(front * (float)cam.sightDist + right * xp + orthoUp * yp);
  r.color = V3(1, 1, 1);
  r.origin = cam.position;
  r.contactObject = CAMERA;
  int obj = NONE;
  unsigned lt = 0;
  r.intersectDist = scene_intersect(c.sc, r, false, obj, lt);
  r.intersectTri = lt;
  r.intersectObject = (r.intersectDist >= 0) ? obj : NONE;
  r.insideObject = find_inside(c.sc, r.origin, r.direction, NONE);
  r.directionProb = 1.f;
  r.directionSampleType = RANDOM;
  r.originSampleType = DEFINITE;
  return r;
}

// PathTracer::renderPixels per-sample estimate (src/PathTracer.cpp:29-67)
static V3 pt_sample(Ctx& c, unsigned pixel, std::vector<PRay>& path) {
  PRay cr = camera_ray(c, pixel);
  sample_path(c, cr, path);
  if (!is_emissive(c.sc, path.back().contactObject)) return V3(0, 0, 0);
  V3 color(1, 1, 1);
  for (size_t i = 0; i < path.size(); i++) {
    color = color * (path[i].color / path[i].directionProb / path[i].originProb);
    if (i != path.size() - 1) {
      color = color * cosine_term(c.sc, path[i]);
      float dist = length(path[i + 1].origin - path[i].origin);
      color = color * radiance_decay(c.sc, path[i], dist);
    }
  }
  return color / c.sc.vignette[pixel];  // Camera::eliminateVignetting
}

// ---- PathTracerTest (NEE + MIS), src/PathTracerTest.cpp:7-136
static V3 get_bsdf(const SceneO& sc, const PRay& in, const PRay& out);
static float dir_pdf(const SceneO& sc, const PRay& in, const PRay& out);

static V3 obj_bsdf(const SceneO& sc, int o, const PRay& in, const PRay& out) {
  const ObjO& O = sc.objs[o];
  auto diffuse = [&](V3 rho) -> V3 {
    if (out.contactObject == NONE) return V3(0, 0, 0);
    Frame lf = frame_from(contact_normal(sc, out));
    return (dot(lf.n, in.direction) <= 0 && dot(lf.n, out.direction) >= 0) ? rho / PI_F : V3(0, 0, 0);
  };
  auto glossy = [&](V3 ks, float n) -> V3 {
    if (out.contactObject == NONE) return V3(0, 0, 0);
    Frame lf = frame_from(contact_normal(sc, out));
    return maxc(ks) * glossy_eval(lf, in.direction, out.direction, ks, n);
  };
  switch (O.type) {
    case OBJ_DIFFUSE: return diffuse(O.c1);
    case OBJ_GLOSSY: return glossy(O.c1, O.expo);
    case OBJ_PHONG: return glossy(O.c1, O.expo) + diffuse(O.c2);
    case OBJ_EMISSIVE:
      if (dot(contact_normal(sc, out), out.direction) < 0) return V3(0, 0, 0);
      return O.c1;
    case OBJ_VPM:
      if (out.contactObject == NONE) return O.c1 * hg_value(O.g, dot(in.direction, out.direction));
      if (out.contactObject != o) return obj_bsdf(sc, out.contactObject, in, out);
      return V3(0, 0, 0);
  }
  return V3(0, 0, 0);  // refractive: Material::getBSDF
}
static V3 get_bsdf(const SceneO& sc, const PRay& in, const PRay& out) {  // Ray::getBSDF
  if (out.contactObject == CAMERA) return V3(1, 1, 1);  // Camera::getBSDF (include/Camera.h:31)
  if (out.contactObject >= 0) return obj_bsdf(sc, out.contactObject, in, out);
  if (out.insideObject >= 0) return obj_bsdf(sc, out.insideObject, in, out);
  return V3(0, 0, 0);
}
static float obj_dir_pdf(const SceneO& sc, int o, const PRay& in, const PRay& out) {
  const ObjO& O = sc.objs[o];
  auto diffuse = [&](V3 rho) -> float {
    if (out.contactObject == NONE) return 0;
    Frame lf = frame_from(contact_normal(sc, out));
    return maxc(rho) * cos_pdf(lf, out.direction, 1.f);
  };
  auto glossy = [&](float n) -> float {
    if (out.contactObject == NONE) return 0;
    Frame lf = frame_from(mirror(contact_normal(sc, out), in.direction));
    return cos_pdf(lf, out.direction, n);
  };
  switch (O.type) {
    case OBJ_DIFFUSE: return diffuse(O.c1);
    case OBJ_GLOSSY: return glossy(O.expo);
    case OBJ_PHONG: return (glossy(O.expo) + diffuse(O.c2)) / 2;
    case OBJ_REFRACTIVE: return 1;
    case OBJ_EMISSIVE: return 0;
    case OBJ_VPM: {
      if (out.directionSampleType == DEFINITE) return 0;
      if (out.contactObject >= 0) return obj_dir_pdf(sc, out.contactObject, in, out);
      float albedo = lum(O.c1) / lum(O.sigma_t);
      return albedo * hg_value(O.g, dot(in.direction, out.direction));
    }
  }
  return 0;
}
static float dir_pdf(const SceneO& sc, const PRay& in, const PRay& out) {
  if (out.contactObject == CAMERA) return 1;  // Camera::getDirectionSampleProbDensity (src/Camera.cpp:99-108)
  if (out.contactObject >= 0) return obj_dir_pdf(sc, out.contactObject, in, out);
  if (out.insideObject >= 0) return obj_dir_pdf(sc, out.insideObject, in, out);
  return 0;
}

// Scene::genEmissionSample -> SurfaceSampler::genSample -> SceneObject::emit
// (src/Scene.cpp:91-98; src/SceneObject.cpp:144-225), origin part
static PRay light_sample(Ctx& c) {
  const SceneO& sc = c.sc;
  PRay r;
  if (sc.emitters.empty()) {
    r.direction = V3(0, 0, 0);
    r.color = V3(0, 0, 0);
    return r;
  }
  float rnd = c.g.f() * sc.emit_total_weight;
  size_t idx = std::lower_bound(sc.emit_weight_cum.begin(), sc.emit_weight_cum.end(), rnd) - sc.emit_weight_cum.begin();
  if (idx >= sc.emit_weight_cum.size()) idx = sc.emit_weight_cum.size() - 1;
  const int o = sc.emitters[idx];
  const ObjO& O = sc.objs[o];
  float rnd2 = c.g.f() * O.area_total;
  size_t ti = std::lower_bound(O.area_cum.begin(), O.area_cum.end(), rnd2) - O.area_cum.begin();
  if (ti >= O.area_cum.size()) ti = O.area_cum.size() - 1;
  r.contactObject = o;
  r.contactTri = (unsigned)ti;
  // SimpleShape::genRandTrianglePosition (src/SimpleShape.cpp:28-42)
  unsigned gt = O.tri_begin + (unsigned)ti;
  V3 v0 = sc.tree.verts[sc.tree.tri_v[3 * gt]], v1 = sc.tree.verts[sc.tree.tri_v[3 * gt + 1]],
     v2 = sc.tree.verts[sc.tree.tri_v[3 * gt + 2]];
  float u1 = c.g.f(), u2 = c.g.f();
  float su1 = std::sqrt(u1);
  float u = 1.f - su1, v = u2 * su1;
  r.origin = v0 * u + v1 * v + v2 * (1.f - u - v);
  r.originProb = O.emit_weight / O.area_total;
  // cosine direction about the light normal (consumes two draws, then flips)
  Frame lf = frame_from(contact_normal(sc, r));
  r.direction = to_world(lf, hemi_cos(c.g, 1.f));
  if (dot(contact_normal(sc, r), r.direction) < 0) r.direction = -r.direction;
  r.insideObject = find_inside(sc, r.origin, r.direction, o);
  r.color = get_bsdf(sc, r, r);
  r.directionProb = cos_pdf(lf, r.direction, 1.f);
  r.directionSampleType = r.originSampleType = RANDOM;
  return r;
}

// MCRenderer::testVisibility (src/MCRenderer.cpp:108-120)
static bool test_visibility(const SceneO& sc, const PRay& start, const PRay& end) {
  int obj;
  unsigned lt;
  float d = scene_intersect(sc, start, true, obj, lt);
  return d < 0 || (d - length(start.origin - end.origin) >= -EPS);
}

// PathTracerTest::colorByConnectingLights (src/PathTracerTest.cpp:88-136)
static V3 connect_lights(Ctx& c, const PRay& ray, const PRay& lastRay) {
  const SceneO& sc = c.sc;
  PRay lightRay = light_sample(c);
  lightRay.direction = ray.origin - lightRay.origin;
  float dist = length(lightRay.direction);
  float dist2 = dist * dist;
  lightRay.direction = normalize(lightRay.direction);
  PRay outRay = ray;
  outRay.direction = -lightRay.direction;
  const PRay& inRay = lastRay;
  if (!test_visibility(sc, outRay, lightRay)) return V3(0, 0, 0);
  V3 bsdfFactor = get_bsdf(sc, lastRay, outRay);
  if (lum(bsdfFactor) < 1e-7f) return V3(0, 0, 0);
  float cosAtLight = std::min(std::max(0.f, std::fabs(dot(contact_normal(sc, lightRay), lightRay.direction))), 1.f);
  if (cosAtLight < 1e-6f) return V3(0, 0, 0);
  float cosToLight = 1;
  if (ray.contactObject != NONE)
    cosToLight = std::min(std::max(0.f, std::fabs(dot(contact_normal(sc, ray), -lightRay.direction))), 1.f);
  float bsdfToLightPdf = dir_pdf(sc, inRay, outRay);
  float dirPdfW = lightRay.originProb * dist2 / cosAtLight;
  float mis = dirPdfW / (dirPdfW + bsdfToLightPdf);
  V3 tmp = lightRay.color * cosAtLight * bsdfFactor * cosToLight / (lightRay.originProb * dist2);
  return tmp * mis;
}

static V3 pttest_sample(Ctx& c, unsigned pixel, std::vector<PRay>& path) {
  const SceneO& sc = c.sc;
  PRay cr = camera_ray(c, pixel);
  sample_path(c, cr, path);
  V3 throughput(1, 1, 1), color(0, 0, 0);
  bool lastSpecular = true;
  float lastPdfW = 1.f;
  for (size_t i = 1; i < path.size(); i++) {
    if (is_emissive(sc, path[i].contactObject)) {
      const ObjO& E = sc.objs[path[i].contactObject];
      V3 contrib = E.c1;
      float dirPdfA = E.emit_weight / E.area_total;
      float mis = 1.f;
      if (i > 1 && !lastSpecular) {
        float cosine = dot(contact_normal(sc, path[i]), -path[i - 1].direction);
        float dist = length(path[i].origin - path[i - 1].origin);
        float dirPdfW = dirPdfA * dist * dist / std::fabs(cosine);
        mis = lastPdfW / (lastPdfW + dirPdfW);
      }
      color = color + throughput * contrib * mis;
      break;
    }
    if (path[i].directionSampleType == RANDOM) color = color + throughput * connect_lights(c, path[i], path[i - 1]);
    if (i + 1 >= path.size() ||
        (path[i].origin.x == path[i + 1].origin.x && path[i].origin.y == path[i + 1].origin.y &&
         path[i].origin.z == path[i + 1].origin.z))
      break;
    lastSpecular = (path[i].directionSampleType == DEFINITE);
    lastPdfW = path[i].directionProb;
    throughput = throughput * (path[i].color * cosine_term(sc, path[i]) / path[i].directionProb);
    float dist = std::max(length(path[i].origin - path[i - 1].origin), 1e-5f);
    throughput = throughput * radiance_decay(sc, path[i - 1], dist);
  }
  return color;
}

// ---- BidirectionalPathTracer, CPU branch (src/BidirectionalPathTracer.cpp:6-398)
typedef std::vector<PRay> Path;

// A ray's density of its own direction, Ray::getDirectionSampleProbDensity(ray)
// called with itself: the emitted-ray branch of SceneObject's
// (src/SceneObject.cpp:234-247, cosine about the contact normal); Camera
// overrides it with 1 (src/Camera.cpp:99-108).
static float self_dir_pdf(const SceneO& sc, const PRay& r) {
  if (r.contactObject == CAMERA) return 1.f;
  if (r.contactObject < 0) return 0.f;
  return cos_pdf(frame_from(contact_normal(sc, r)), r.direction, 1.f);
}
// Ray::getOriginSampleProbDensity (src/Ray.cpp:42-50): 1 unless the ray
// travels in a medium, then SceneVPMObject::getOriginSampleProbDensity
// (src/SceneVPMObject.cpp:266-273): P_surface for a vertex with a contact
// object (the camera included), p_medium for a scattering point.
static float origin_pdf(const SceneO& sc, const PRay& in, const PRay& out) {
  if (in.insideObject < 0 || sc.objs[in.insideObject].type != OBJ_VPM) return 1.f;
  const float dist = std::max(length(in.origin - out.origin), EPS);
  const ObjO& o = sc.objs[in.insideObject];
  const bool surface = (out.contactObject != NONE) != (g_mutation == MUT_ORIGIN_PDF_SWAP);
  return surface ? P_surface(o, dist) : p_medium(o, dist);
}

// Density check of the balance weights (test hook, RenderParamsO::flags bit 1).
static bool g_density_check = false;
static std::atomic<uint64_t> g_dc_checked{0}, g_dc_skipped{0};
static std::atomic<uint64_t> g_dc_max_bits{0};  // bits of the largest relative difference (a non-negative double)
static void dc_record(double rel) {
  uint64_t bits;
  std::memcpy(&bits, &rel, 8);
  uint64_t cur = g_dc_max_bits.load();
  while (bits > cur && !g_dc_max_bits.compare_exchange_weak(cur, bits)) {
  }
}

// BidirectionalPathTracer::mustUsePT (:124-131)
static bool must_use_pt(const Path& path) {
  for (size_t i = 0; i + 1 < path.size(); i++)
    if (path[i].directionSampleType == path[i + 1].directionSampleType && path[i].directionSampleType == RANDOM)
      return false;
  return true;
}

// BidirectionalPathTracer::link (:183-195): the vertex i ray, turned toward
// vertex j unless it already points that way along the connected path.
static PRay bpt_link(const Path& path, int ci, int i, int j) {
  const int n = (int)path.size();
  const PRay& src = path[std::min(std::max(i, 0), n - 1)];
  const PRay& dst = path[std::min(std::max(j, 0), n - 1)];
  PRay r = src;
  if ((j > i && i <= ci) || (j < i && i > ci)) return r;
  r.direction = normalize(dst.origin - src.origin);
  r.insideObject = dst.insideObject;
  return r;
}

// MCRenderer::connectRays (src/MCRenderer.cpp:50-106), merged = false
static bool connect_rays(const SceneO& sc, Path& path, int ci) {
  PRay& lightRay = path[ci];
  PRay& eyeRay = path[ci + 1];
  if (lightRay.directionSampleType == DEFINITE || eyeRay.directionSampleType == DEFINITE) return false;
  const int n = (int)path.size();
  const PRay& prevLightRay = path[std::max(ci - 1, 0)];
  const PRay& prevEyeRay = path[std::min(ci + 2, n - 1)];
  lightRay.direction = normalize(eyeRay.origin - lightRay.origin);
  eyeRay.direction = -lightRay.direction;
  lightRay.directionProb = 1;
  lightRay.color = get_bsdf(sc, prevLightRay, lightRay);
  lightRay.directionSampleType = RANDOM;
  if (lightRay.contactObject != NONE && ci > 0) {
    if (dot(contact_normal(sc, lightRay), lightRay.direction) <= 0) lightRay.insideObject = lightRay.contactObject;
    else if (dot(contact_normal(sc, lightRay), prevLightRay.direction) <= 0) lightRay.insideObject = prevLightRay.insideObject;
    else lightRay.insideObject = find_inside(sc, lightRay.origin, lightRay.direction, lightRay.contactObject);
  }
  eyeRay.directionProb = 1;
  eyeRay.color = get_bsdf(sc, prevEyeRay, eyeRay);
  eyeRay.directionSampleType = RANDOM;
  if (eyeRay.contactObject != NONE && ci < n - 2) {
    if (dot(contact_normal(sc, eyeRay), eyeRay.direction) <= 0) eyeRay.insideObject = eyeRay.contactObject;
    else if (dot(contact_normal(sc, eyeRay), prevEyeRay.direction) <= 0) eyeRay.insideObject = prevEyeRay.insideObject;
    else eyeRay.insideObject = find_inside(sc, eyeRay.origin, eyeRay.direction, eyeRay.contactObject);
  }
  return true;
}

// BidirectionalPathTracer::connectColorProb (:133-181): (throughput, density)
static void connect_color_prob(const SceneO& sc, const Path& path, int ci, V3& color, float& prob) {
  color = V3(1, 1, 1);
  prob = 1;
  const int n = (int)path.size();
  for (int i = 0; i < n; i++) {
    float dist = 0;
    if (i <= ci) {
      dist = std::max(length(path[i + 1].origin - path[i].origin), EPS);
      color = color * radiance_decay(sc, path[i], dist);
    } else if (i > ci + 1) {
      dist = std::max(length(path[i - 1].origin - path[i].origin), EPS);
      color = color * radiance_decay(sc, path[i], dist);
    }
    if (i == ci && i < n - 1) color = color * (cosine_term(sc, path[i]) * cosine_term(sc, path[i + 1]) / (dist * dist));
    if (i != ci && i != ci + 1) color = color * cosine_term(sc, path[i]);
    prob *= path[i].directionProb * path[i].originProb;
    color = color * path[i].color;
  }
}

// BidirectionalPathTracer::connectWeight (:197-263) with expTerm = 1: the
// balance heuristic over the strategies that could have made the path
static float connect_weight(const SceneO& sc, const Path& path, int ci, bool usePT) {
  const int n = (int)path.size();
  std::vector<double> pf(n), pb(n), dl(n);
  pf[0] = path[0].originProb;
  for (int i = 0; i < n - 1; i++) {
    pf[i + 1] = pf[i] * cosine_term(sc, bpt_link(path, ci, i + 1, i));
    dl[i] = std::max(length(path[i + 1].origin - path[i].origin), EPS);
    pf[i + 1] /= dl[i] * dl[i];
    if (path[i].directionSampleType == RANDOM) {
      if (i > 0) pf[i + 1] *= dir_pdf(sc, bpt_link(path, ci, i - 1, i), bpt_link(path, ci, i, i + 1));
      else pf[i + 1] *= self_dir_pdf(sc, bpt_link(path, ci, i, i + 1));
      pf[i + 1] *= origin_pdf(sc, bpt_link(path, ci, i, i + 1), bpt_link(path, ci, i + 1, i + 2));
    }
  }
  pb[n - 1] = path[n - 1].originProb;
  for (int i = n - 1; i > 0; i--) {
    pb[i - 1] = pb[i] * cosine_term(sc, bpt_link(path, ci, i - 1, i));
    pb[i - 1] /= dl[i - 1] * dl[i - 1];
    if (path[i].directionSampleType == RANDOM) {
      if (i < n - 1) pb[i - 1] *= dir_pdf(sc, bpt_link(path, ci, i + 1, i), bpt_link(path, ci, i, i - 1));
      else pb[i - 1] *= self_dir_pdf(sc, bpt_link(path, ci, i, i - 1));
      pb[i - 1] *= origin_pdf(sc, bpt_link(path, ci, i, i - 1), bpt_link(path, ci, i - 1, i - 2));
    }
  }
  if (g_density_check && ci >= 0 && ci + 1 < n) {
    // The light subpath drew vertices 0..ci, the eye subpath n-1..ci+1.  Each
    // density factor connectWeight uses for those sampled steps is compared
    // with the one the step was drawn with: the origin density of the vertex
    // a step leads to (the ray's originProb: P_surface or p_medium of the
    // medium crossed, 1 outside media), and the direction density at a
    // vertex (the ray's directionProb) where the reference's two agree:
    // emission, the camera, diffuse surfaces and scattering points in a
    // medium.  (At a glossy or Phong vertex getDirectionSampleProbDensity
    // leaves out the roulette factor, or sums both lobes, that the scatter
    // includes; at a DEFINITE vertex connectWeight includes neither factor.)
    auto diffuse_or_medium = [&](const PRay& r) {
      if (r.contactObject == CAMERA) return true;
      if (r.contactObject >= 0) return sc.objs[r.contactObject].type == OBJ_DIFFUSE;
      return r.insideObject >= 0 && sc.objs[r.insideObject].type == OBJ_VPM;
    };
    double worst = 0;
    auto cmp = [&](double model, double drawn) {
      const double rel = std::fabs(model / drawn - 1.0);
      worst = std::max(worst, std::isfinite(rel) ? rel : 1e300);
    };
    bool definite = false;
    for (int i = 0; i < ci; i++) {
      if (path[i].directionSampleType != RANDOM) {
        definite = true;
        continue;
      }
      cmp(origin_pdf(sc, bpt_link(path, ci, i, i + 1), bpt_link(path, ci, i + 1, i + 2)), path[i + 1].originProb);
      if (i == 0) cmp(self_dir_pdf(sc, bpt_link(path, ci, 0, 1)), path[0].directionProb);
      else if (diffuse_or_medium(path[i]))
        cmp(dir_pdf(sc, bpt_link(path, ci, i - 1, i), bpt_link(path, ci, i, i + 1)), path[i].directionProb);
    }
    for (int i = n - 1; i > ci + 1; i--) {
      if (path[i].directionSampleType != RANDOM) {
        definite = true;
        continue;
      }
      cmp(origin_pdf(sc, bpt_link(path, ci, i, i - 1), bpt_link(path, ci, i - 1, i - 2)), path[i - 1].originProb);
      if (i == n - 1) cmp(self_dir_pdf(sc, bpt_link(path, ci, i, i - 1)), path[i].directionProb);
      else if (diffuse_or_medium(path[i]))
        cmp(dir_pdf(sc, bpt_link(path, ci, i + 1, i), bpt_link(path, ci, i, i - 1)), path[i].directionProb);
    }
    dc_record(worst);
    g_dc_checked++;
    if (definite) g_dc_skipped++;
  }
  double sum = 0;
  for (int i = 0; i < n - 1; i++)
    if (path[i].directionSampleType == RANDOM && path[i + 1].directionSampleType == RANDOM) sum += pf[i] * pb[i + 1];
  if (usePT || must_use_pt(path)) sum += pb[0];
  double self = ci == -1 ? pb[0] : pf[ci] * pb[ci + 1];
  if (!(self / sum >= 0)) return 0;
  return (float)(self / sum);
}

// Camera::transToPixel (src/Camera.cpp:136-152)
static void trans_to_pixel(const SceneO& sc, V3 point, float& px, float& py) {
  V3 v = point - sc.cam.position;
  float frontDist = dot(v, sc.cam_front);
  float ratio = sc.cam.sightDist / frontDist;
  px = dot(v, sc.cam_right) * ratio;
  py = dot(v, sc.cam_up) * ratio;
  px += sc.cam.width / 2.f - 0.5f;
  py = sc.cam.height / 2.f - py - 0.5f;
}

// BidirectionalPathTracer::colorByConnectingPaths (:265-396): every (light
// prefix, eye prefix) pair of one eye and one light subpath; contributions
// with an eye prefix of one vertex land on the pixel the light vertex projects
// to, the others on the eye path's own pixel.  img is the iteration's image.
// With `registered`, a light-tracing contribution lands on the pixel whose
// camera rays cover its raster position (floor, not the reference's
// (int)(x - 0.5) of transToPixel's pixel-centre coordinates); the estimator
// then has the PathTracer's expectation at every pixel, which is what the
// tests use it for.
static void connect_paths(Ctx& c, unsigned pixel, const Path& eyePath, const Path& lightPath, double* img,
                          bool usePT, bool registered) {
  const SceneO& sc = c.sc;
  const int maxE = (int)eyePath.size(), maxL = (int)lightPath.size();
  const int npix = sc.cam.width * sc.cam.height;
  for (int W = 2; W <= maxE + maxL; W++) {
    const int lplStart = W - maxE > 0 ? W - maxE : 0;
    const int lplEnd = W - 1 < maxL ? W - 1 : maxL;
    for (int lp = lplStart; lp <= lplEnd; lp++) {
      const int ep = W - lp;
      Path path(W);
      for (int li = 0; li < lp; li++) path[li] = lightPath[li];
      for (int ei = 0; ei < ep; ei++) path[W - 1 - ei] = eyePath[ei];
      const int lci = lp - 1, eci = lp;
      if (lp == 0 && !is_emissive(sc, path.front().contactObject)) continue;
      if (lp == 0 && !(usePT || must_use_pt(path))) continue;
      if (ep > 0 && lp > 0) {
        if (lp > 1 && is_emissive(sc, path[lci].contactObject)) continue;
        if (is_emissive(sc, path[eci].contactObject)) continue;
        if (!connect_rays(sc, path, lci)) continue;
        if (!test_visibility(sc, path[eci], path[lci])) continue;
      }
      V3 color;
      float prob;
      connect_color_prob(sc, path, lci, color, prob);
      if (!(prob > 0)) continue;
      if (!(length(color) > 0)) continue;
      float weight = connect_weight(sc, path, lci, usePT);
      color = color / prob * weight;
      if (ep == 1) {
        V3 dir = normalize(path[lci].origin - path[eci].origin);
        float cosAtCamera = std::fabs(dot(sc.cam_front, dir));
        float imageToSolidAngleFactor = (float)(sc.cam.sightDist * sc.cam.sightDist) / (cosAtCamera * cosAtCamera * cosAtCamera);
        float imageToSurfaceFactor = imageToSolidAngleFactor / cosAtCamera;
        color = color * (imageToSurfaceFactor / (float)npix);
        color = color * cosAtCamera;
        float fx, fy;
        trans_to_pixel(sc, path[W - 2].origin, fx, fy);
        int x = (int)fx, y = (int)fy;
        if (registered) {
          x = (int)std::floor(fx + 0.5f);
          y = (int)std::floor(fy + 0.5f);
        }
        if (x >= 0 && x < sc.cam.width && y >= 0 && y < sc.cam.height) {
          double* d = img + 3 * ((size_t)y * sc.cam.width + x);
          for (int k = 0; k < 3; k++) {
#pragma omp atomic
            d[k] += color[k];
          }
        }
      } else {
        double* d = img + 3 * (size_t)pixel;
        for (int k = 0; k < 3; k++) {
#pragma omp atomic
          d[k] += color[k];
        }
      }
    }
  }
}

// One BidirectionalPathTracer iteration per sample: per pixel one eye and
// one light subpath (src/BidirectionalPathTracer.cpp:36-57), the iteration
// image divided by the vignetting factor (:100) and added to the FP64 sums.
static void render_bpt(const SceneO& sc, const RenderParamsO& p, double* sum, double* sumsq, uint64_t* rays) {
  g_density_check = (p.flags & 2u) != 0;
  if (g_density_check) {
    g_dc_checked = 0;
    g_dc_skipped = 0;
    g_dc_max_bits = 0;
  }
  const unsigned npix = (unsigned)(sc.cam.width * sc.cam.height);
  std::vector<double> img(3 * (size_t)npix);
  uint64_t total_rays = 0;
  for (unsigned smp = 0; smp < p.spp; smp++) {
    std::fill(img.begin(), img.end(), 0.0);
#pragma omp parallel for schedule(dynamic, 16) reduction(+ : total_rays)
    for (int px = 0; px < (int)npix; px++) {
      g_rays = 0;
      Rng g(p.seed, (uint64_t)px, (uint64_t)(smp + p.sample_offset));
      Ctx c{sc, g, p.max_depth, p.rr_start_depth};
      Path eyePath, lightPath;
      PRay lightRay = light_sample(c);
      if (length(lightRay.direction) > 0.5f) {  // SceneObject::emit's own first hit, with the condition (:212-223)
        int obj = NONE;
        unsigned lt = 0;
        float dist = scene_intersect(sc, lightRay, true, obj, lt);
        if (dist > 0) {
          lightRay.intersectDist = dist;
          lightRay.intersectObject = obj;
          lightRay.intersectTri = lt;
        }
      }
      PRay cameraRay = camera_ray(c, px);
      sample_path(c, cameraRay, eyePath);
      c.light = true;
      sample_path(c, lightRay, lightPath);
      c.light = false;
      connect_paths(c, (unsigned)px, eyePath, lightPath, img.data(), false, (p.flags & 1u) != 0);
      total_rays += g_rays;
    }
    for (unsigned px = 0; px < npix; px++)
      for (int k = 0; k < 3; k++) {
        double v = (float)(img[3 * px + k]) / sc.vignette[px];
        sum[3 * px + k] += v;
        if (sumsq) sumsq[3 * px + k] += v * v;
      }
  }
  if (rays) *rays += total_rays;
}

void render(const SceneO& sc, const RenderParamsO& p, double* sum, double* sumsq, uint64_t* rays) {
  if (p.mode == 3) {
    render_bpt(sc, p, sum, sumsq, rays);
    return;
  }
  const unsigned npix = (unsigned)(sc.cam.width * sc.cam.height);
  uint64_t total_rays = 0;
#pragma omp parallel for schedule(static) reduction(+ : total_rays)
  for (int px = 0; px < (int)npix; px++) {
    g_rays = 0;
    std::vector<PRay> path;
    double s[3] = {0, 0, 0}, q[3] = {0, 0, 0};
    for (unsigned smp = 0; smp < p.spp; smp++) {
      Rng g(p.seed, (uint64_t)px, (uint64_t)(smp + p.sample_offset));
      Ctx c{sc, g, p.max_depth, p.rr_start_depth};
      V3 v = (p.mode == 1) ? pttest_sample(c, px, path) : pt_sample(c, px, path);
      for (int k = 0; k < 3; k++) {
        s[k] += v[k];
        q[k] += (double)v[k] * v[k];
      }
    }
    for (int k = 0; k < 3; k++) {
      sum[3 * px + k] += s[k];
      if (sumsq) sumsq[3 * px + k] += q[k];
    }
    total_rays += g_rays;
  }
  if (rays) *rays += total_rays;
}

// Draws from the restated samplers, for distribution tests:
// kind 0: cosine-power lobe about +y (genHemiCosDirection), out = (x,y,z)
// kind 1: HG phase about +y (HGPhaseSampler::genSample), out = (x,y,z)
// kind 2: homogeneous-medium distance (HomoMediaDistSampler), out = (s,0,0), param = sigma_t (grey)
// kind 3: Fresnel reflectance er of RefractiveMaterial at incidence angle acos(param2) into IOR param
void sampler_draws(int kind, int n, float param, float param2, uint64_t seed, float* out) {
  for (int i = 0; i < n; i++) {
    Rng g(seed, (uint64_t)i, 0);
    V3 r;
    if (kind == 0) {
      r = hemi_cos(g, param);
    } else if (kind == 1) {
      Frame lf = frame_from(V3(0, 1, 0));
      float e1 = g.f(), e2 = g.f(), gg = param, cosTheta;
      if (std::fabs(gg) < EPS) cosTheta = 1 - 2 * e1;
      else {
        float sqrTerm = (1 - gg * gg) / (1 - gg + 2 * gg * e1);
        cosTheta = (1 + gg * gg - sqrTerm * sqrTerm) / (2 * gg);
      }
      float sinTheta = std::sqrt(1 - cosTheta * cosTheta);
      r = to_world(lf, V3(sinTheta * (float)std::cos(2.0 * M_PI * e2), cosTheta, sinTheta * (float)std::sin(2.0 * M_PI * e2)));
    } else if (kind == 2) {
      r = V3((float)(-std::log(1.0 - g.f()) / lum(V3(param, param, param))), 0, 0);
    } else {
      float cur = 1.f, nxt = param;
      float theta = std::acos(-param2);  // incoming against a +z normal
      float sin_phi = cur / nxt * std::sin(theta);
      float phi = std::asin(sin_phi);
      if (theta > M_PI / 2) phi = M_PI - phi;
      float ct = std::fabs(std::cos(theta)), cp = std::fabs(std::cos(phi));
      float esr = std::pow(std::fabs(cur * ct - nxt * cp) / (cur * ct + nxt * cp), 2.f);
      float epr = std::pow(std::fabs(nxt * ct - cur * cp) / (nxt * ct + cur * cp), 2.f);
      r = V3((esr + epr) / 2, 0, 0);
    }
    out[3 * i] = r.x;
    out[3 * i + 1] = r.y;
    out[3 * i + 2] = r.z;
  }
}

bool pin_refract(int kind, const float* d, const float* n, float n_i, float n_t, float* dir, float* er) {
  V3 out(0, 0, 0);
  float e = 0;
  const bool ok = refr_dir(V3(d[0], d[1], d[2]), V3(n[0], n[1], n[2]), n_i, n_t, kind == 0, out, e);
  dir[0] = out.x;
  dir[1] = out.y;
  dir[2] = out.z;
  *er = e;
  return ok;
}

void pin_medium(const float* sigma_t, float x, float* p_med, float* P_surf) {
  const V3 st(sigma_t[0], sigma_t[1], sigma_t[2]);
  *p_med = p_medium(st, x);
  *P_surf = P_surface(st, x);
}

void pin_glossy(const float* n, const float* in, const float* out, const float* ks, float coeff, float* res) {
  const Frame lf = frame_from(V3(n[0], n[1], n[2]));
  const V3 r = glossy_eval(lf, V3(in[0], in[1], in[2]), V3(out[0], out[1], out[2]), V3(ks[0], ks[1], ks[2]), coeff);
  res[0] = r.x;
  res[1] = r.y;
  res[2] = r.z;
}

uint32_t mt_nth(uint32_t n) {
  MT m(5489u);
  uint32_t v = 0;
  for (uint32_t i = 0; i < n; i++) v = m.next();
  return v;
}

void bpt_density_stats(uint64_t* checked, uint64_t* skipped, double* max_rel) {
  *checked = g_dc_checked.load();
  *skipped = g_dc_skipped.load();
  const uint64_t bits = g_dc_max_bits.load();
  std::memcpy(max_rel, &bits, 8);
}

}  // namespace orc
