// C entry points of the oracle library (liboracle.so), loaded by the tests
// with ctypes.  TEST INFRASTRUCTURE ONLY.
#include <omp.h>

#include <cmath>
#include <cstring>

#include "rx_oracle_geom.h"
#include "rx_oracle_scene.h"

using namespace orc;

namespace {

void scene_setup(SceneO& sc, int nv, const float* v, int nt, const uint32_t* tri, int nobj,
                 const ObjDescC* od, const CamDescC* cd, int abs_mode) {
  sc.tree.verts.resize(nv);
  for (int i = 0; i < nv; i++) sc.tree.verts[i] = V3(v[3 * i], v[3 * i + 1], v[3 * i + 2]);
  sc.tree.tri_v.assign(tri, tri + 3 * (size_t)nt);
  sc.tree.abs_mode = abs_mode;
  build_tree(sc.tree);  // Scene::buildKDTree (src/Scene.cpp:248-285)
  sc.tri_obj.assign(nt, -1);
  sc.tri_normal.resize(nt);
  for (int t = 0; t < nt; t++) {
    const V3& a = sc.tree.verts[tri[3 * t]];
    const V3& b = sc.tree.verts[tri[3 * t + 1]];
    const V3& c = sc.tree.verts[tri[3 * t + 2]];
    sc.tri_normal[t] = face_normal(a, b, c);
  }
  sc.objs.resize(nobj);
  sc.obj_tree.resize(nobj);
  for (int o = 0; o < nobj; o++) {
    ObjO& O = sc.objs[o];
    O.type = od[o].type;
    O.tri_begin = od[o].tri_begin;
    O.tri_count = od[o].tri_count;
    O.c1 = V3(od[o].color[0], od[o].color[1], od[o].color[2]);
    O.c2 = V3(od[o].color2[0], od[o].color2[1], od[o].color2[2]);
    O.expo = od[o].exponent;
    O.ior = od[o].ior;
    O.g = od[o].g;
    O.sigma_t = O.c1 + O.c2;  // setSigma: dt = ds + da (include/SceneVPMObject.h:84-88)
    for (unsigned t = 0; t < O.tri_count; t++) sc.tri_obj[O.tri_begin + t] = o;
    if (O.type == OBJ_VPM) {  // per-object tree for containment tests
      Tree& T = sc.obj_tree[o];
      T.verts = sc.tree.verts;
      T.tri_v.assign(tri + 3 * (size_t)O.tri_begin, tri + 3 * (size_t)(O.tri_begin + O.tri_count));
      T.abs_mode = abs_mode;
      build_tree(T);
    }
  }
  // emitter sampler (Scene::preprocessEmissionSampler, SurfaceSampler::preprocess)
  float totalWeight = 0;
  for (int o = 0; o < nobj; o++) {
    ObjO& O = sc.objs[o];
    if (O.type != OBJ_EMISSIVE) continue;
    O.area_total = 0;
    for (unsigned t = 0; t < O.tri_count; t++) {
      unsigned g = O.tri_begin + t;
      V3 a = sc.tree.verts[tri[3 * g]], b = sc.tree.verts[tri[3 * g + 1]], c = sc.tree.verts[tri[3 * g + 2]];
      float area = length(cross(b - a, c - b)) / 2;  // SimpleShape::getTriangleArea
      O.area_total += area;
      O.area_cum.push_back(O.area_total);
    }
    O.emit_weight = O.area_total;
    totalWeight += O.emit_weight;
    sc.emitters.push_back(o);
    sc.emit_weight_cum.push_back(totalWeight);
  }
  sc.emit_total_weight = totalWeight;
  for (int o : sc.emitters) sc.objs[o].emit_weight /= totalWeight;
  // camera (src/Camera.cpp:4-36,78-96)
  CamO& C = sc.cam;
  C.position = V3(cd->position[0], cd->position[1], cd->position[2]);
  C.focus = V3(cd->focus[0], cd->focus[1], cd->focus[2]);
  C.up = V3(cd->up[0], cd->up[1], cd->up[2]);
  C.sightDist = cd->sight_dist;
  C.width = cd->width;
  C.height = cd->height;
  sc.cam_front = normalize(C.focus - C.position);
  sc.cam_right = normalize(cross(sc.cam_front, C.up));
  sc.cam_up = cross(sc.cam_right, sc.cam_front);
  sc.vignette.resize((size_t)C.width * C.height);
  for (int y = 0; y < C.height; y++)
    for (int x = 0; x < C.width; x++) {
      V3 dir = sc.cam_front * (float)C.sightDist + sc.cam_right * (float)(x + 0.5 - (C.width / 2.0)) +
               sc.cam_up * (float)((C.height / 2.0) - 0.5 - y);
      dir = normalize(dir);
      sc.vignette[(size_t)y * C.width + x] = std::fabs(dot(dir, sc.cam_front));
    }
}

}  // namespace

extern "C" {

void* orc_tree_create(int nv, const float* v, int nt, const uint32_t* tri, int abs_mode) {
  Tree* t = new Tree;
  t->verts.resize(nv);
  for (int i = 0; i < nv; i++) t->verts[i] = V3(v[3 * i], v[3 * i + 1], v[3 * i + 2]);
  t->tri_v.assign(tri, tri + 3 * (size_t)nt);
  t->abs_mode = abs_mode;
  build_tree(*t);
  return t;
}

void orc_tree_free(void* h) { delete (Tree*)h; }

// Number of serialized nodes and leaf triangles.
void orc_tree_sizes(void* h, int* n_nodes, int* n_leaf_tris) {
  Serialized s;
  serialize(*(Tree*)h, s);
  *n_nodes = (int)s.nodes4.size() / 4;
  *n_leaf_tris = (int)s.v1_4.size() / 4;
}

// KDTree::serializeForGPU arrays, vec4f each.
void orc_tree_serialize(void* h, float* nodes4, float* min4, float* max4, float* v1, float* v2, float* v3) {
  Serialized s;
  serialize(*(Tree*)h, s);
  memcpy(nodes4, s.nodes4.data(), s.nodes4.size() * 4);
  memcpy(min4, s.min4.data(), s.min4.size() * 4);
  memcpy(max4, s.max4.data(), s.max4.size() * 4);
  memcpy(v1, s.v1_4.data(), s.v1_4.size() * 4);
  memcpy(v2, s.v2_4.data(), s.v2_4.size() * 4);
  memcpy(v3, s.v3_4.data(), s.v3_4.size() * 4);
}

// Face normal per triangle (SimpleShape::getWorldNormal, flat branch).
void orc_face_normals(int nv, const float* v, int nt, const uint32_t* tri, float* out) {
  (void)nv;
  for (int t = 0; t < nt; t++) {
    V3 a(v[3 * tri[3 * t]], v[3 * tri[3 * t] + 1], v[3 * tri[3 * t] + 2]);
    V3 b(v[3 * tri[3 * t + 1]], v[3 * tri[3 * t + 1] + 1], v[3 * tri[3 * t + 1] + 2]);
    V3 c(v[3 * tri[3 * t + 2]], v[3 * tri[3 * t + 2] + 1], v[3 * tri[3 * t + 2] + 2]);
    V3 n = face_normal(a, b, c);
    out[3 * t] = n.x;
    out[3 * t + 1] = n.y;
    out[3 * t + 2] = n.z;
  }
}

// Nearest hit for a batch.  mode 0 = exhaustive recursion (reference),
// 1 = ordered traversal (canonical counts), 2 = brute force.
// Legality arrays may be NULL (no condition); skip (may be NULL) names a
// triangle id per ray that is never hit (IntersectionGPU::Ray::last_tid).
void orc_intersect(void* h, int mode, int n, const float* org, const float* dir, const int* tri_obj,
                   const float* tri_normal, const int* inside, const int* contact, const int* skip, float* dist,
                   int* tid, uint32_t* boxes, uint32_t* tris, int threads) {
  const Tree& t = *(Tree*)h;
  if (threads > 0) omp_set_num_threads(threads);
#pragma omp parallel for schedule(dynamic, 256)
  for (int i = 0; i < n; i++) {
    Ray r{V3(org[3 * i], org[3 * i + 1], org[3 * i + 2]), V3(dir[3 * i], dir[3 * i + 1], dir[3 * i + 2]),
          skip ? skip[i] : -1};
    Legality L{tri_obj, (const V3*)tri_normal, inside ? inside[i] : -1, contact ? contact[i] : -1};
    const Legality* cond = (tri_obj && tri_normal && inside && contact) ? &L : nullptr;
    HitStats st;
    unsigned id = 0;
    float d;
    if (mode == 0)
      d = intersect_exhaustive(t, r, id, cond, &st);
    else if (mode == 1)
      d = intersect_ordered(t, r, id, cond, &st);
    else
      d = intersect_bruteforce(t, r, id, cond);
    dist[i] = d;
    tid[i] = d >= 0 ? (int)id : -1;
    if (boxes) boxes[i] = st.boxes;
    if (tris) tris[i] = st.tris;
  }
}

// FP64 barycentrics (u, v) of the ray/plane point: p = (1-u-v) v0 + u v1 + v v2.
void orc_barycentrics(int n, const float* org, const float* dir, const int* tid, const float* verts,
                      const uint32_t* tri, double* uv) {
  for (int i = 0; i < n; i++) {
    if (tid[i] < 0) {
      uv[2 * i] = uv[2 * i + 1] = -1;
      continue;
    }
    double P[3][3];
    for (int k = 0; k < 3; k++)
      for (int c = 0; c < 3; c++) P[k][c] = verts[3 * tri[3 * tid[i] + k] + c];
    double o[3] = {org[3 * i], org[3 * i + 1], org[3 * i + 2]}, d[3] = {dir[3 * i], dir[3 * i + 1], dir[3 * i + 2]};
    double e1[3], e2[3], s[3], p[3], q[3];
    for (int c = 0; c < 3; c++) {
      e1[c] = P[1][c] - P[0][c];
      e2[c] = P[2][c] - P[0][c];
      s[c] = o[c] - P[0][c];
    }
    p[0] = d[1] * e2[2] - d[2] * e2[1];
    p[1] = d[2] * e2[0] - d[0] * e2[2];
    p[2] = d[0] * e2[1] - d[1] * e2[0];
    q[0] = s[1] * e1[2] - s[2] * e1[1];
    q[1] = s[2] * e1[0] - s[0] * e1[2];
    q[2] = s[0] * e1[1] - s[1] * e1[0];
    double det = p[0] * e1[0] + p[1] * e1[1] + p[2] * e1[2];
    uv[2 * i] = (p[0] * s[0] + p[1] * s[1] + p[2] * s[2]) / det;
    uv[2 * i + 1] = (q[0] * d[0] + q[1] * d[1] + q[2] * d[2]) / det;
  }
}

void* orc_scene_create(int nv, const float* v, int nt, const uint32_t* tri, int nobj, const void* objs,
                       const void* cam, int abs_mode) {
  SceneO* sc = new SceneO;
  scene_setup(*sc, nv, v, nt, tri, nobj, (const ObjDescC*)objs, (const CamDescC*)cam, abs_mode);
  return sc;
}

void orc_scene_free(void* h) { delete (SceneO*)h; }

// Replace the camera (and its per-pixel vignetting factors) of a built scene.
void orc_scene_set_camera(void* h, const void* cam) {
  SceneO& sc = *(SceneO*)h;
  const CamDescC* cd = (const CamDescC*)cam;
  CamO& C = sc.cam;
  C.position = V3(cd->position[0], cd->position[1], cd->position[2]);
  C.focus = V3(cd->focus[0], cd->focus[1], cd->focus[2]);
  C.up = V3(cd->up[0], cd->up[1], cd->up[2]);
  C.sightDist = cd->sight_dist;
  C.width = cd->width;
  C.height = cd->height;
  sc.cam_front = normalize(C.focus - C.position);
  sc.cam_right = normalize(cross(sc.cam_front, C.up));
  sc.cam_up = cross(sc.cam_right, sc.cam_front);
  sc.vignette.resize((size_t)C.width * C.height);
  for (int y = 0; y < C.height; y++)
    for (int x = 0; x < C.width; x++) {
      V3 dir = sc.cam_front * (float)C.sightDist + sc.cam_right * (float)(x + 0.5 - (C.width / 2.0)) +
               sc.cam_up * (float)((C.height / 2.0) - 0.5 - y);
      dir = normalize(dir);
      sc.vignette[(size_t)y * C.width + x] = std::fabs(dot(dir, sc.cam_front));
    }
}

void* orc_scene_tree(void* h) { return &((SceneO*)h)->tree; }

void orc_render(void* h, const void* params, double* sum, double* sumsq, int threads, uint64_t* rays) {
  if (threads > 0) omp_set_num_threads(threads);
  render(*(SceneO*)h, *(const RenderParamsO*)params, sum, sumsq, rays);
}

uint32_t orc_mt_nth(uint32_t n) { return mt_nth(n); }

void orc_sampler_draws(int kind, int n, float param, float param2, uint64_t seed, float* out) {
  sampler_draws(kind, n, param, param2, seed, out);
}

int orc_max_threads(void) { return omp_get_max_threads(); }

void orc_set_mutation(int m) { g_mutation = m; }
void orc_bpt_density_stats(uint64_t* checked, uint64_t* skipped, double* max_rel) {
  bpt_density_stats(checked, skipped, max_rel);
}
int orc_refract_dir(int kind, const float* d, const float* n, float n_i, float n_t, float* dir, float* er) {
  return pin_refract(kind, d, n, n_i, n_t, dir, er) ? 1 : 0;
}
void orc_medium_densities(const float* sigma_t, float x, float* p_med, float* P_surf) {
  pin_medium(sigma_t, x, p_med, P_surf);
}
void orc_glossy_eval(const float* n, const float* in, const float* out, const float* ks, float coeff, float* res) {
  pin_glossy(n, in, out, ks, coeff, res);
}

}  // extern "C"
