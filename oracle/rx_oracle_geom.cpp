// Oracle geometry: restatement of src/KDTree.cpp and
// src/NoSelfIntersectionCondition.cpp.  TEST INFRASTRUCTURE ONLY (see header).
#include "rx_oracle_geom.h"

#include <algorithm>
#include <cstdlib>
#include <limits>

namespace orc {

// boundvec<true/false> (src/KDTree.cpp:4-20)
static inline V3 bound_min(V3 p, V3 c) {
  return V3(p.x < c.x ? p.x : c.x, p.y < c.y ? p.y : c.y, p.z < c.z ? p.z : c.z);
}
static inline V3 bound_max(V3 p, V3 c) {
  return V3(p.x > c.x ? p.x : c.x, p.y > c.y ? p.y : c.y, p.z > c.z ? p.z : c.z);
}

// |ratio - 0.5| under the two bindings of `abs` (src/KDTree.cpp:191)
static inline double split_badness(float ratio, int mode) {
  double v = (double)ratio - 0.5;
  if (mode == ABS_INT) return (double)std::abs((int)v);
  return std::fabs(v);
}

// KDTree::splitBest (src/KDTree.cpp:129-228), iterative over a work list.
void build_tree(Tree& t) {
  t.nodes.clear();
  const unsigned ntri = (unsigned)(t.tri_v.size() / 3);
  t.nodes.emplace_back();
  for (unsigned i = 0; i < ntri; i++) t.nodes[0].tris.push_back(i);  // KDTree::build :233-235
  std::vector<int> work{0};
  std::vector<unsigned> uniq;
  std::vector<float> coords;
  while (!work.empty()) {
    int ni = work.back();
    work.pop_back();
    std::vector<unsigned> tris;
    tris.swap(t.nodes[ni].tris);
    if (tris.empty()) continue;  // :135-136
    // tight box over the node's triangle vertices (:137-148)
    V3 lo = t.verts[t.tri_v[3 * tris[0]]], hi = lo;
    uniq.clear();
    for (unsigned tr : tris)
      for (int k = 0; k < 3; k++) {
        unsigned vi = t.tri_v[3 * tr + k];
        lo = bound_min(t.verts[vi], lo);
        hi = bound_max(t.verts[vi], hi);
        uniq.push_back(vi);
      }
    t.nodes[ni].lo = lo;
    t.nodes[ni].hi = hi;
    if (tris.size() <= 1) {  // maxLeafTriangleNum = 1 (:24, :153-154)
      t.nodes[ni].tris.swap(tris);
      continue;
    }
    // unique vertex ids (the unordered_set of :139-147); only the multiset of
    // their coordinates matters to nth_element below.
    std::sort(uniq.begin(), uniq.end());
    uniq.erase(std::unique(uniq.begin(), uniq.end()), uniq.end());
    float bestRatio = 10.0f;
    int bestDim = 0;
    float bestSplit = 0.0f;
    for (int dim = 0; dim < 3; dim++) {
      coords.clear();
      for (unsigned vi : uniq) coords.push_back(t.verts[vi][dim]);
      // nth_element at size>>1 (:161-168)
      std::nth_element(coords.begin(), coords.begin() + (coords.size() >> 1), coords.end());
      float split = coords[coords.size() >> 1];
      unsigned leftCount = 0;
      for (unsigned tr : tris) {  // :171-186: left if ANY vertex <= split
        unsigned mask = 0;
        for (int k = 0; k < 3; k++) mask += (t.verts[t.tri_v[3 * tr + k]][dim] <= split) << k;
        if (mask != 0) leftCount++;
      }
      float ratio = leftCount / float(tris.size());
      if (split_badness(ratio, t.abs_mode) < split_badness(bestRatio, t.abs_mode)) {
        bestRatio = ratio;
        bestDim = dim;
        bestSplit = split;
      }
    }
    if (bestRatio == 0 || bestRatio == 1) {  // :199-203
      t.nodes[ni].tris.swap(tris);
      continue;
    }
    std::vector<unsigned> L, R;
    for (unsigned tr : tris) {  // :208-224
      unsigned mask = 0;
      for (int k = 0; k < 3; k++) mask += (t.verts[t.tri_v[3 * tr + k]][bestDim] <= bestSplit) << k;
      (mask == 0 ? R : L).push_back(tr);
    }
    int li = (int)t.nodes.size();
    t.nodes.emplace_back();
    int ri = (int)t.nodes.size();
    t.nodes.emplace_back();
    t.nodes[ni].left = li;
    t.nodes[ni].right = ri;
    t.nodes[li].tris.swap(L);
    t.nodes[ri].tris.swap(R);
    work.push_back(ri);
    work.push_back(li);
  }
  // leaf order of the pre-order traversal (the serialization order)
  t.leaf_rank.assign(t.nodes.size(), -1);
  t.tri_leaf.assign(ntri, -1);
  t.tri_pos.assign(ntri, 0);
  std::vector<int> stk{0};
  int next = 0;
  while (!t.nodes.empty() && !stk.empty()) {
    int n = stk.back();
    stk.pop_back();
    if (t.nodes[n].left < 0) {
      int p = 0;
      for (unsigned tr : t.nodes[n].tris) {
        t.tri_leaf[tr] = next;
        t.tri_pos[tr] = p++;
      }
      t.leaf_rank[n] = next++;
    } else {
      stk.push_back(t.nodes[n].right);
      stk.push_back(t.nodes[n].left);
    }
  }
}

static void serialize_rec(const Tree& t, int ni, int parent, Serialized& s) {
  const Node& n = t.nodes[ni];
  auto push4 = [](std::vector<float>& v, float a, float b, float c, float d) {
    v.push_back(a); v.push_back(b); v.push_back(c); v.push_back(d);
  };
  push4(s.min4, n.lo.x, n.lo.y, n.lo.z, 0);
  push4(s.max4, n.hi.x, n.hi.y, n.hi.z, 0);
  if (n.left >= 0 && n.right >= 0) {
    push4(s.nodes4, (float)parent, -1, -1, -1);
    int me = (int)s.nodes4.size() / 4 - 1;
    serialize_rec(t, n.left, me, s);
    s.nodes4[4 * me + 1] = (float)(s.nodes4.size() / 4);  // right child index
    serialize_rec(t, n.right, me, s);
  } else {
    int si = (int)s.v1_4.size() / 4;
    for (unsigned tr : n.tris) {
      V3 a = t.verts[t.tri_v[3 * tr]], b = t.verts[t.tri_v[3 * tr + 1]], c = t.verts[t.tri_v[3 * tr + 2]];
      push4(s.v1_4, a.x, a.y, a.z, (float)tr);
      push4(s.v2_4, b.x, b.y, b.z, 0);
      push4(s.v3_4, c.x, c.y, c.z, 0);
    }
    push4(s.nodes4, (float)parent, -1, (float)si, (float)(s.v1_4.size() / 4));
  }
}

void serialize(const Tree& t, Serialized& s) {
  s = Serialized();
  if (!t.nodes.empty()) serialize_rec(t, 0, -1, s);
}

float box_intersect(V3 lo, V3 hi, const Ray& r) {
  const float bb[6] = {lo.x, lo.y, lo.z, hi.x, hi.y, hi.z};  // ((float*)this)[i]
  float minDist = -1;
  for (unsigned i = 0; i < 6; i++) {
    const int a = i % 3;
    if (r.d[a] == 0) continue;
    float dist = (bb[i] - r.o[a]) / r.d[a];
    V3 p = r.o + dist * r.d;
    bool inter = dist >= 0;
    const int b = (i + 1) % 3, c = (i + 2) % 3;
    inter &= lo[b] <= p[b];
    inter &= hi[b] >= p[b];
    inter &= lo[c] <= p[c];
    inter &= hi[c] >= p[c];
    minDist = (inter && (dist < minDist || minDist < 0)) ? dist : minDist;
  }
  return minDist;
}

float tri_intersect(V3 v0, V3 v1, V3 v2, const Ray& r) {
  V3 vs[3] = {v0, v1, v2}, lines[3], r2v, n;
  unsigned mask_pos = 0, mask_neg = 0;
  for (unsigned i = 0; i < 3; i++) {
    lines[i] = vs[(i + 1) % 3] - vs[i];
    r2v = vs[i] - r.o;
    n = cross(lines[i], r2v);
    mask_pos += (dot(r.d, n) >= 0) << i;
    mask_neg += (dot(r.d, n) <= 0) << i;
  }
  n = normalize(cross(lines[0], lines[1]));
  float orthoDist = dot(n, r2v);
  float d = dot(n, r.d);
  if (d == 0 || orthoDist / d < 0 || (mask_pos != 7 && mask_neg != 7)) return -1;
  return orthoDist / d;
}

V3 face_normal(V3 v0, V3 v1, V3 v2) { return normalize(cross(v1 - v0, v2 - v0)); }

bool legal(const Legality& c, const Ray& r, unsigned tid) {
  const int obj = c.tri_obj[tid];
  float d = dot(r.d, c.tri_normal[tid]);
  bool in = d <= 0, out = d >= 0;
  if (c.inside == c.contact) return !(in && c.contact == obj);
  return !(out && c.inside != obj);
}

static inline float tri_test(const Tree& t, unsigned tr, const Ray& r, const Legality* cond) {
  if ((int)tr == r.skip) return -1;
  float dist = tri_intersect(t.verts[t.tri_v[3 * tr]], t.verts[t.tri_v[3 * tr + 1]],
                             t.verts[t.tri_v[3 * tr + 2]], r);
  if (dist < 0) return -1;  // geometric miss returns -1 exactly
  if (cond && !legal(*cond, r, tr)) return -1;
  return dist;
}

static float exh_rec(const Tree& t, int ni, const Ray& r, unsigned& tid, const Legality* cond,
                     HitStats* st) {
  const Node& n = t.nodes[ni];
  if (st) st->boxes++;
  if (box_intersect(n.lo, n.hi, r) < 0) return -1;
  if (n.left < 0 && n.right < 0) {
    float minDist = -1;
    for (unsigned tr : n.tris) {
      if (st) st->tris++;
      float dist = tri_test(t, tr, r, cond);
      if (dist >= 0 && (dist < minDist || minDist < 0)) {
        minDist = dist;
        tid = tr;
      }
    }
    return minDist;
  }
  unsigned ltid = 0, rtid = 0;
  float ld = exh_rec(t, n.left, r, ltid, cond, st);
  float rd = exh_rec(t, n.right, r, rtid, cond, st);
  if (ld >= 0 && rd >= 0) {
    tid = ld < rd ? ltid : rtid;
    return ld < rd ? ld : rd;
  } else if (rd < 0) {
    tid = ltid;
    return ld;
  } else if (ld < 0) {
    tid = rtid;
    return rd;
  }
  return -1;
}

float intersect_exhaustive(const Tree& t, const Ray& r, unsigned& tid, const Legality* cond,
                           HitStats* st) {
  if (t.nodes.empty()) return -1;
  return exh_rec(t, 0, r, tid, cond, st);
}

// ---- canonical ordered traversal -----------------------------------------
// Slab entry distance with per-ray reciprocal; an ordering key only.  Same
// operations (one correctly rounded reciprocal per axis, oi = o * inv rounded,
// each slab end one fused multiply-add fma(plane, inv, -oi), compare-select)
// as the HIP traversal kernel so both produce the same key.
static inline float slab_tnear(V3 lo, V3 hi, const Ray& r, const float inv[3], const float oi[3]) {
  const float INF = std::numeric_limits<float>::infinity();
  float tn[3];
  for (int a = 0; a < 3; a++) {
    float t0, t1;
    if (r.d[a] == 0) {
      t0 = -INF;
      t1 = INF;
    } else {
      t0 = std::fmaf(lo[a], inv[a], -oi[a]);
      t1 = std::fmaf(hi[a], inv[a], -oi[a]);
    }
    tn[a] = t0 < t1 ? t0 : t1;
  }
  float t = tn[0];
  t = tn[1] > t ? tn[1] : t;
  t = tn[2] > t ? tn[2] : t;
  return t;
}
// The absolute term of every cull threshold: twice the rounding bound
// u |oi_a| of the keys' slab ends over the axes with d_a != 0 (INF when one
// is not finite), as the HIP traversal computes it.
static inline float cull_margin(const Ray& r, const float oi[3]) {
  const float INF = std::numeric_limits<float>::infinity();
  float m = 0.0f;
  for (int a = 0; a < 3; a++) {
    float x = r.d[a] != 0 ? std::fabs(oi[a]) : 0.0f;
    x = x <= INF ? x : INF;
    m = x > m ? x : m;
  }
  return m * (2.0f * 0x1p-24f);
}

float intersect_ordered(const Tree& t, const Ray& r, unsigned& tid, const Legality* cond,
                        HitStats* st) {
  if (t.nodes.empty()) return -1;
  // larger leaf rank wins distance ties (right subtree wins, src/KDTree.cpp:317-321);
  // first in leaf wins (:305)
  const std::vector<int>& rank = t.leaf_rank;
  const float INF = std::numeric_limits<float>::infinity();
  float inv[3], oi[3];
  for (int a = 0; a < 3; a++) inv[a] = 1.0f / r.d[a];
  for (int a = 0; a < 3; a++) oi[a] = r.o[a] * inv[a];
  const float cm = cull_margin(r, oi);
  bool has = false;
  float best = -1;
  int bestRank = -1;
  unsigned bestTid = 0;
  if (st) st->boxes++;
  if (box_intersect(t.nodes[0].lo, t.nodes[0].hi, r) < 0) return -1;
  struct E { int n; float tn; };
  std::vector<E> stack;
  int cur = 0;
  for (;;) {
    const Node& n = t.nodes[cur];
    bool descend = false;
    if (n.left < 0) {
      for (unsigned tr : n.tris) {
        if (st) st->tris++;
        float dist = tri_test(t, tr, r, cond);
        if (dist >= 0 && (!has || dist < best || (dist == best && rank[cur] > bestRank))) {
          has = true;
          best = dist;
          bestRank = rank[cur];
          bestTid = tr;
        }
      }
    } else {
      const Node& L = t.nodes[n.left];
      const Node& R = t.nodes[n.right];
      if (st) st->boxes += 2;
      bool hl = box_intersect(L.lo, L.hi, r) >= 0;
      bool hr = box_intersect(R.lo, R.hi, r) >= 0;
      float tl = slab_tnear(L.lo, L.hi, r, inv, oi), trr = slab_tnear(R.lo, R.hi, r, inv, oi);
      float thr = has ? std::fmaf(best, 1.001f, cm) : INF;
      bool vl = hl && tl <= thr, vr = hr && trr <= thr;
      if (vl && vr) {
        if (trr < tl) {
          stack.push_back({n.left, tl});
          cur = n.right;
        } else {
          stack.push_back({n.right, trr});
          cur = n.left;
        }
        descend = true;
      } else if (vl) {
        cur = n.left;
        descend = true;
      } else if (vr) {
        cur = n.right;
        descend = true;
      }
    }
    if (descend) continue;
    bool found = false;
    while (!stack.empty()) {
      E e = stack.back();
      stack.pop_back();
      float thr = has ? std::fmaf(best, 1.001f, cm) : INF;
      if (e.tn <= thr) {
        cur = e.n;
        found = true;
        break;
      }
    }
    if (!found) break;
  }
  if (!has) return -1;
  tid = bestTid;
  return best;
}

float intersect_bruteforce(const Tree& t, const Ray& r, unsigned& tid, const Legality* cond) {
  // Nearest legal hit over all triangles; ties broken by the tree's leaf order
  // exactly as the recursion does.
  const std::vector<int>& leafOf = t.tri_leaf;
  const std::vector<int>& posOf = t.tri_pos;
  bool has = false;
  float best = -1;
  unsigned bt = 0;
  for (unsigned tr = 0; tr < leafOf.size(); tr++) {
    float dist = tri_test(t, tr, r, cond);
    if (dist < 0 || !(dist >= 0)) continue;
    bool take = !has || dist < best ||
                (dist == best && (leafOf[tr] > leafOf[bt] ||
                                  (leafOf[tr] == leafOf[bt] && posOf[tr] < posOf[bt])));
    if (take) {
      has = true;
      best = dist;
      bt = tr;
    }
  }
  if (!has) return -1;
  tid = bt;
  return best;
}

}  // namespace orc
