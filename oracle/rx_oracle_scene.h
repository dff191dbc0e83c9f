// Oracle scene model: the parts of Scene / SceneObject / Camera state the
// restated path tracer reads.  TEST INFRASTRUCTURE ONLY.
#pragma once
#include <cstdint>
#include <vector>

#include "rx_oracle_geom.h"

namespace orc {

enum ObjType { OBJ_DIFFUSE = 0, OBJ_GLOSSY = 1, OBJ_PHONG = 2, OBJ_REFRACTIVE = 3, OBJ_EMISSIVE = 4, OBJ_VPM = 5 };

// Flat C layout shared with the tests' scene description (64 bytes).
struct ObjDescC {
  int32_t type;
  uint32_t tri_begin, tri_count, pad0;
  float color[3];
  float exponent;
  float color2[3];
  float ior;
  float g;
  float pad1[3];
};
struct CamDescC {
  float position[3], focus[3], up[3];
  uint32_t sight_dist;
  int32_t width, height;
};
struct RenderParamsO {
  uint32_t spp, max_depth, rr_start_depth, mode;  // 0 PathTracer, 1 PathTracerTest, 3 BidirectionalPathTracer
  uint64_t seed;
  uint32_t sample_offset, flags;
};

struct ObjO {
  int type;
  unsigned tri_begin, tri_count;
  V3 c1, c2;
  float expo, ior, g;
  V3 sigma_t;
  // emitter sampler state (SceneObject::preprocessEmissionSampler, Scene.cpp:19-33)
  std::vector<float> area_cum;
  float area_total = 0, emit_weight = 0;
};
struct CamO {
  V3 position, focus, up;
  unsigned sightDist;
  int width, height;
};

struct SceneO {
  Tree tree;
  std::vector<int> tri_obj;
  std::vector<V3> tri_normal;
  std::vector<ObjO> objs;
  std::vector<Tree> obj_tree;  // per-object trees (Scene::buildObjKDTrees), volumetric only
  CamO cam;
  V3 cam_front, cam_right, cam_up;
  std::vector<float> vignette;  // |d_center . front| per pixel
  std::vector<int> emitters;
  std::vector<float> emit_weight_cum;
  float emit_total_weight = 0;
};

void render(const SceneO& sc, const RenderParamsO& p, double* sum, double* sumsq, uint64_t* rays);

// Planted mistakes (test hooks only): each must make one of the oracle's pins fail.
enum Mutation {
  MUT_NONE = 0,
  MUT_ROTATION_SIGN = 1,         // rotMat by -phi: the refracted ray bends the wrong way
  MUT_INVERTED_RATIO = 2,        // n_t / n_i instead of n_i / n_t
  MUT_P_SURFACE_COMPLEMENT = 3,  // P_surface = y(1 - e^{-sigma_t d})
  MUT_P_MEDIUM_NO_SIGMA = 4,     // p_medium without its sigma_t factor
  MUT_GLOSSY_AXIS = 5,           // glossy lobe about -d instead of the mirror direction
  MUT_ORIGIN_PDF_SWAP = 6        // the balance weights' origin density: P_surface and p_medium swapped
};
extern int g_mutation;
// Bidirectional renders with RenderParamsO::flags bit 1: every connection
// compares the densities its balance weight uses for the sampled prefixes
// (the products pf / pb of BidirectionalPathTracer::connectWeight) with the
// densities the subpaths were actually drawn with (each ray's directionProb
// and originProb), under the same geometry factors.  Counts (connections
// compared; those among them with a DEFINITE vertex, whose step neither side
// includes) and the largest relative difference since the last reset.
void bpt_density_stats(uint64_t* checked, uint64_t* skipped, double* max_rel);
// The oracle's own functions, for the tests that pin them:
// kind 0 = RefractiveMaterial's step, 1 = SceneVPMObject's; false on TIR.
bool pin_refract(int kind, const float* d, const float* n, float n_i, float n_t, float* dir, float* er);
void pin_medium(const float* sigma_t, float x, float* p_med, float* P_surf);
void pin_glossy(const float* n, const float* in, const float* out, const float* ks, float coeff, float* res);
uint32_t mt_nth(uint32_t n);
void sampler_draws(int kind, int n, float param, float param2, uint64_t seed, float* out);

}  // namespace orc
