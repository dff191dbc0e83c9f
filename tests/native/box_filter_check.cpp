// CPU check of the traversal's box filter: over random and adversarial
// (ray, box) pairs, slab_classify may only claim a miss where the reference
// box predicate (box_hit, src/KDTree.cpp:34-52) is false and a hit where it
// is true.
// Usage: box_filter_check <pairs per family> <seed>.  Prints one line of
// counts per family; exits 1 on the first violation.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <random>

#include "../../renderx_amd/csrc/rx_dev.h"

using namespace rx;

static float nudge(float x, int ulps) {
  for (int i = 0; i < ulps; i++) x = std::nextafterf(x, INFINITY);
  for (int i = 0; i > ulps; i--) x = std::nextafterf(x, -INFINITY);
  return x;
}

int main(int argc, char** argv) {
  const long n = argc > 1 ? atol(argv[1]) : 1000000;
  std::mt19937_64 rng(argc > 2 ? atoll(argv[2]) : 1);
  std::uniform_real_distribution<double> U(0.0, 1.0);
  auto unit = [&]() {
    double z = 2 * U(rng) - 1, p = 2 * M_PI * U(rng), r = std::sqrt(std::max(0.0, 1 - z * z));
    return v3((float)(r * std::cos(p)), (float)(r * std::sin(p)), (float)z);
  };
  long bad = 0;
  for (int fam = 0; fam < 8; fam++) {
    long fast = 0, filtered = 0, exact_miss = 0, exact_hit = 0, certain_hit = 0;
    for (long k = 0; k < n; k++) {
      const double R = std::pow(10.0, (int)(U(rng) * 5) - 1);  // scene half-size 0.1 .. 1000
      const float rlo[3] = {(float)-R, (float)-R, (float)-R}, rhi[3] = {(float)R, (float)R, (float)R};
      float lo[3], hi[3];
      for (int a = 0; a < 3; a++) {
        double c = (2 * U(rng) - 1) * R * 0.9, h = R * std::pow(10.0, -6 * U(rng)) * 0.1;
        lo[a] = (float)(c - h);
        hi[a] = (float)(c + h);
        if (fam == 5 && U(rng) < 0.3) hi[a] = lo[a];  // flat boxes (single-triangle leaves)
      }
      Vec3 o = v3((float)((2 * U(rng) - 1) * R), (float)((2 * U(rng) - 1) * R), (float)((2 * U(rng) - 1) * R));
      Vec3 d = unit();
      if (fam >= 1) {  // aim at a point on the box surface: a face, an edge or a corner, nudged by ulps
        float p[3];
        for (int a = 0; a < 3; a++) p[a] = (float)(lo[a] + (hi[a] - lo[a]) * U(rng));
        int nfix = 1 + (int)(U(rng) * 3);
        for (int j = 0; j < nfix; j++) {
          int a = (int)(U(rng) * 3);
          p[a] = U(rng) < 0.5 ? lo[a] : hi[a];
          p[a] = nudge(p[a], (int)(U(rng) * 9) - 4);
        }
        Vec3 t = v3(p[0], p[1], p[2]);
        d = normalize(t - o);
      }
      if (fam == 6) {  // aim just outside an edge or corner, at a relative gap 1e-8 .. 1e-1 of the box size
        float p[3];
        for (int a = 0; a < 3; a++) p[a] = (float)(lo[a] + (hi[a] - lo[a]) * U(rng));
        int nfix = 1 + (int)(U(rng) * 3);
        for (int j = 0; j < nfix; j++) {
          int a = (int)(U(rng) * 3);
          double gap = (hi[a] - lo[a] + std::fabs(lo[a]) * 1e-6) * std::pow(10.0, -8 + 7 * U(rng));
          p[a] = U(rng) < 0.5 ? (float)(lo[a] - gap) : (float)(hi[a] + gap);
        }
        d = normalize(v3(p[0], p[1], p[2]) - o);
      }
      if (fam == 7) {  // origin far outside the scene box (a distant camera), aimed at a face, edge or corner
        const Vec3 u = unit();
        const float far = (float)(R * std::pow(10.0, 1 + 3 * U(rng)));
        o = v3(u.x * far, u.y * far, u.z * far);
        float p[3];
        for (int a = 0; a < 3; a++) p[a] = (float)(lo[a] + (hi[a] - lo[a]) * U(rng));
        const int a = (int)(U(rng) * 3);
        p[a] = nudge(U(rng) < 0.5 ? lo[a] : hi[a], (int)(U(rng) * 9) - 4);
        d = normalize(v3(p[0], p[1], p[2]) - o);
      }
      if (fam == 2) {  // nearly axis-parallel directions
        int a = (int)(U(rng) * 3);
        float s = (float)std::pow(10.0, -7 * U(rng) - 1);
        if (a == 0) d.x *= s;
        if (a == 1) d.y *= s;
        if (a == 2) d.z *= s;
        d = normalize(d);
      }
      if (fam == 3) {  // origin exactly on a face plane of the box
        int a = (int)(U(rng) * 3);
        float v = U(rng) < 0.5 ? (a == 0 ? lo[0] : a == 1 ? lo[1] : lo[2]) : (a == 0 ? hi[0] : a == 1 ? hi[1] : hi[2]);
        if (a == 0) o.x = v;
        if (a == 1) o.y = v;
        if (a == 2) o.z = v;
      }
      if (fam == 4) {  // origin just inside / outside a face, ray along the face
        int a = (int)(U(rng) * 3);
        float v = nudge(U(rng) < 0.5 ? lo[a] : hi[a], (int)(U(rng) * 5) - 2);
        if (a == 0) { o.x = v; d.x = 0.0f; }
        if (a == 1) { o.y = v; d.y = 0.0f; }
        if (a == 2) { o.z = v; d.z = 0.0f; }
        d = normalize(d);
        if (U(rng) < 0.5) {  // and a slightly tilted twin
          if (a == 0) d.x = 1e-6f;
          if (a == 1) d.y = 1e-6f;
          if (a == 2) d.z = 1e-6f;
        }
      }
      const bool exact = box_hit(lo, hi, o, d);
      (exact ? exact_hit : exact_miss)++;
      const Vec3 inv = v3(1.0f / d.x, 1.0f / d.y, 1.0f / d.z);
      Vec3 beta;
      if (!slab_filter_setup(rlo, rhi, o, d, inv, beta)) continue;
      fast++;
      float tn;
      const Vec3 oi = slab_oi(o, inv);
      const int cls = slab_classify(lo, hi, oi, inv, beta, tn);
      const bool miss = cls == kBoxMiss;
      if (cls == kBoxHit) {
        certain_hit++;
        if (!exact) {
          printf("VIOLATION (certain hit) fam %d: lo %a %a %a hi %a %a %a o %a %a %a d %a %a %a\n", fam, lo[0], lo[1],
                 lo[2], hi[0], hi[1], hi[2], o.x, o.y, o.z, d.x, d.y, d.z);
          return 1;
        }
      }
      const float tref = slab_tnear(lo, hi, oi, d, inv);
      if (!(tn == tref || (std::isnan(tn) && std::isnan(tref)))) {
        printf("key mismatch fam %d: %.9g vs %.9g\n", fam, tn, tref);
        return 1;
      }
      if (miss) {
        filtered++;
        if (exact) {
          printf("VIOLATION fam %d: lo %a %a %a hi %a %a %a o %a %a %a d %a %a %a\n", fam, lo[0], lo[1], lo[2], hi[0],
                 hi[1], hi[2], o.x, o.y, o.z, d.x, d.y, d.z);
          bad++;
          return 1;
        }
      }
    }
    printf("family %d: pairs %ld exact_hit %ld exact_miss %ld filter_usable %ld certain_miss %ld certain_hit %ld\n", fam,
           n, exact_hit, exact_miss, fast, filtered, certain_hit);
  }
  return bad ? 1 : 0;
}
