// Host KD-tree builder: same splits as KDTree::splitBest (src/KDTree.cpp:129-228)
// but built in parallel with linear-time per-node work.
//
// Per node the reference builds an unordered_set of the node's vertex ids and
// calls nth_element on their coordinates per axis; only the multiset of
// unique-vertex coordinates matters for the chosen split, so a per-thread
// stamp array replaces the hash set.  Subtrees above a size threshold are
// built on separate threads into their own arenas and stitched in pre-order.
#include "rx_tree.h"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <thread>

namespace rx {
namespace {

struct BNode {
  float lo[3], hi[3];
  int32_t l = -1, r = -1;  // >= 0: node in the same arena; <= -2: root of child arena (-2 - k)
  uint32_t tb = 0, tc = 0; // leaf: range in arena.tris
};

struct Arena {
  std::vector<BNode> nodes;
  std::vector<uint32_t> tris;
  std::vector<std::unique_ptr<Arena>> kids;
  std::vector<int32_t> kid_root;
};

struct Ctx {
  const float* v;
  const uint32_t* t;
  uint32_t nv;
  bool abs_int;
  std::atomic<int> threads{1};
  int max_threads;
};

struct Scratch {
  std::vector<uint32_t> stamp;
  uint32_t id = 0;
  std::vector<uint32_t> uniq;
  std::vector<float> coords;
};

static inline double badness(float ratio, bool abs_int) {
  double v = (double)ratio - 0.5;
  return abs_int ? (double)std::abs((int)v) : std::fabs(v);
}

// Builds the subtree of `tris` into arena `A`; returns its node index there.
int32_t build_rec(Ctx& C, Arena& A, Scratch& S, std::vector<uint32_t>& tris, uint32_t depth,
                  uint32_t& maxdepth);

int32_t spawn_or_build(Ctx& C, Arena& A, Scratch& S, std::vector<uint32_t>& tris, uint32_t depth,
                       uint32_t& maxdepth, std::thread* th, std::unique_ptr<Arena>* kid, uint32_t* kid_depth,
                       int32_t* kid_root) {
  if (tris.size() > 32768 && C.threads.load() < C.max_threads) {
    C.threads++;
    kid->reset(new Arena);
    Arena* KA = kid->get();
    auto* moved = new std::vector<uint32_t>();
    moved->swap(tris);
    *th = std::thread([&C, KA, moved, depth, kid_depth, kid_root]() {
      Scratch S2;
      S2.stamp.assign(C.nv, 0);
      uint32_t md = depth;
      *kid_root = build_rec(C, *KA, S2, *moved, depth, md);
      *kid_depth = md;
      delete moved;
      C.threads--;
    });
    return INT32_MIN;  // placeholder, resolved after join
  }
  return build_rec(C, A, S, tris, depth, maxdepth);
}

int32_t build_rec(Ctx& C, Arena& A, Scratch& S, std::vector<uint32_t>& tris, uint32_t depth,
                  uint32_t& maxdepth) {
  if (depth > maxdepth) maxdepth = depth;
  int32_t me = (int32_t)A.nodes.size();
  A.nodes.emplace_back();
  const float* V = C.v;
  const uint32_t* T = C.t;
  // tight box with boundvec semantics (src/KDTree.cpp:4-20,137-148)
  float lo[3], hi[3];
  {
    const float* p0 = V + 3 * (size_t)T[3 * (size_t)tris[0]];
    for (int a = 0; a < 3; a++) lo[a] = hi[a] = p0[a];
  }
  S.id++;
  if (S.id == 0) {
    std::fill(S.stamp.begin(), S.stamp.end(), 0);
    S.id = 1;
  }
  S.uniq.clear();
  for (uint32_t tr : tris)
    for (int k = 0; k < 3; k++) {
      uint32_t vi = T[3 * (size_t)tr + k];
      const float* p = V + 3 * (size_t)vi;
      for (int a = 0; a < 3; a++) {
        lo[a] = p[a] < lo[a] ? p[a] : lo[a];
        hi[a] = p[a] > hi[a] ? p[a] : hi[a];
      }
      if (S.stamp[vi] != S.id) {
        S.stamp[vi] = S.id;
        S.uniq.push_back(vi);
      }
    }
  auto make_leaf = [&]() {
    BNode& n = A.nodes[me];
    memcpy(n.lo, lo, 12);
    memcpy(n.hi, hi, 12);
    n.tb = (uint32_t)A.tris.size();
    n.tc = (uint32_t)tris.size();
    A.tris.insert(A.tris.end(), tris.begin(), tris.end());
    return me;
  };
  if (tris.size() <= 1) return make_leaf();  // maxLeafTriangleNum = 1
  float bestRatio = 10.0f, bestSplit = 0.0f;
  int bestDim = 0;
  for (int dim = 0; dim < 3; dim++) {
    S.coords.resize(S.uniq.size());
    for (size_t i = 0; i < S.uniq.size(); i++) S.coords[i] = V[3 * (size_t)S.uniq[i] + dim];
    size_t mid = S.coords.size() >> 1;
    std::nth_element(S.coords.begin(), S.coords.begin() + mid, S.coords.end());
    float split = S.coords[mid];
    uint32_t leftCount = 0;
    for (uint32_t tr : tris) {
      const uint32_t* tv = T + 3 * (size_t)tr;
      leftCount += (V[3 * (size_t)tv[0] + dim] <= split) | (V[3 * (size_t)tv[1] + dim] <= split) |
                   (V[3 * (size_t)tv[2] + dim] <= split);
    }
    float ratio = leftCount / float(tris.size());
    if (badness(ratio, C.abs_int) < badness(bestRatio, C.abs_int)) {
      bestRatio = ratio;
      bestDim = dim;
      bestSplit = split;
    }
  }
  if (bestRatio == 0 || bestRatio == 1) return make_leaf();
  {
    BNode& n = A.nodes[me];
    memcpy(n.lo, lo, 12);
    memcpy(n.hi, hi, 12);
  }
  std::vector<uint32_t> L, R;
  const size_t nL = std::min(tris.size(), (size_t)(bestRatio * (double)tris.size() + 0.5));
  L.reserve(nL + 1);
  R.reserve(tris.size() - nL + 1);
  for (uint32_t tr : tris) {
    const uint32_t* tv = T + 3 * (size_t)tr;
    bool left = (V[3 * (size_t)tv[0] + bestDim] <= bestSplit) | (V[3 * (size_t)tv[1] + bestDim] <= bestSplit) |
                (V[3 * (size_t)tv[2] + bestDim] <= bestSplit);
    (left ? L : R).push_back(tr);
  }
  std::vector<uint32_t>().swap(tris);
  std::thread th;
  std::unique_ptr<Arena> kid;
  uint32_t kid_depth = 0;
  int32_t kid_root = -1;
  int32_t l = spawn_or_build(C, A, S, L, depth + 1, maxdepth, &th, &kid, &kid_depth, &kid_root);
  int32_t r = build_rec(C, A, S, R, depth + 1, maxdepth);
  if (l == INT32_MIN) {
    th.join();
    A.kids.push_back(std::move(kid));
    A.kid_root.push_back(kid_root);
    l = -2 - (int32_t)(A.kids.size() - 1);
    if (kid_depth > maxdepth) maxdepth = kid_depth;
  }
  A.nodes[me].l = l;
  A.nodes[me].r = r;
  return me;
}

void flatten(const Arena& A, int32_t ni, int32_t parent, HostTree& out) {
  if (ni <= -2) {
    const Arena& K = *A.kids[-2 - ni];
    flatten(K, A.kid_root[-2 - ni], parent, out);
    return;
  }
  const BNode& n = A.nodes[ni];
  int32_t me = (int32_t)out.nodes.size();
  out.nodes.emplace_back();
  PNode& p = out.nodes.back();
  memcpy(p.lo, n.lo, 12);
  memcpy(p.hi, n.hi, 12);
  p.parent = parent;
  if (n.l == -1 && n.r == -1) {
    p.right = -1;
    p.leaf_begin = (uint32_t)out.leaf_tris.size();
    out.leaf_tris.insert(out.leaf_tris.end(), A.tris.begin() + n.tb, A.tris.begin() + n.tb + n.tc);
    p.leaf_end = (uint32_t)out.leaf_tris.size();
    return;
  }
  p.leaf_begin = p.leaf_end = 0;
  flatten(A, n.l, me, out);
  out.nodes[me].right = (int32_t)out.nodes.size();
  flatten(A, n.r, me, out);
}

}  // namespace

void build_tree(const float* verts, uint32_t n_verts, const uint32_t* tris, uint32_t n_tris, bool abs_int,
                HostTree& out) {
  out = HostTree();
  if (n_tris == 0) return;
  Ctx C;
  C.v = verts;
  C.t = tris;
  C.nv = n_verts;
  C.abs_int = abs_int;
  unsigned hc = std::thread::hardware_concurrency();
  C.max_threads = (int)std::min(16u, std::max(1u, hc));
  Arena A;
  Scratch S;
  S.stamp.assign(n_verts, 0);
  std::vector<uint32_t> all(n_tris);
  for (uint32_t i = 0; i < n_tris; i++) all[i] = i;
  uint32_t md = 0;
  int32_t root = build_rec(C, A, S, all, 0, md);
  out.nodes.reserve(2 * (size_t)n_tris);
  out.leaf_tris.reserve(n_tris);
  flatten(A, root, -1, out);
  out.max_depth = md;
}

bool parse_serialized(uint32_t n_nodes, const float* nodes4, const float* min4, const float* max4,
                      uint32_t n_leaf, const float* v1, const float* v2, const float* v3, HostTree& out,
                      std::vector<float>& leaf_verts) {
  out = HostTree();
  out.nodes.resize(n_nodes);
  std::vector<uint32_t> depth(n_nodes, 0);
  for (uint32_t i = 0; i < n_nodes; i++) {
    PNode& p = out.nodes[i];
    for (int a = 0; a < 3; a++) {
      p.lo[a] = min4[4 * (size_t)i + a];
      p.hi[a] = max4[4 * (size_t)i + a];
    }
    p.parent = (int32_t)nodes4[4 * (size_t)i];
    p.right = (int32_t)nodes4[4 * (size_t)i + 1];
    if (p.right >= 0) {
      if ((uint32_t)p.right >= n_nodes || p.right <= (int32_t)i + 1) return false;
      p.leaf_begin = p.leaf_end = 0;
    } else {
      p.right = -1;
      p.leaf_begin = (uint32_t)nodes4[4 * (size_t)i + 2];
      p.leaf_end = (uint32_t)nodes4[4 * (size_t)i + 3];
      if (p.leaf_end < p.leaf_begin || p.leaf_end > n_leaf) return false;
    }
    if (p.parent >= 0) {
      if ((uint32_t)p.parent >= i) return false;
      depth[i] = depth[p.parent] + 1;
      if (depth[i] > out.max_depth) out.max_depth = depth[i];
    } else if (i != 0) {
      return false;  // only the root has no parent
    }
  }
  // the implicit links agree with the parent pointers (pre-order, left child
  // = i + 1): the depths above are then the traversal's stack bound
  for (uint32_t i = 0; i < n_nodes; i++) {
    const PNode& p = out.nodes[i];
    if (p.right < 0) continue;
    if (out.nodes[i + 1].parent != (int32_t)i || out.nodes[p.right].parent != (int32_t)i) return false;
  }
  out.leaf_tris.resize(n_leaf);
  leaf_verts.resize(9 * (size_t)n_leaf);
  for (uint32_t k = 0; k < n_leaf; k++) {
    out.leaf_tris[k] = (uint32_t)v1[4 * (size_t)k + 3];
    for (int a = 0; a < 3; a++) {
      leaf_verts[9 * (size_t)k + a] = v1[4 * (size_t)k + a];
      leaf_verts[9 * (size_t)k + 3 + a] = v2[4 * (size_t)k + a];
      leaf_verts[9 * (size_t)k + 6 + a] = v3[4 * (size_t)k + a];
    }
  }
  return true;
}

void face_normal(const float* a, const float* b, const float* c, float* n) {
  float b1[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
  float b2[3] = {c[0] - a[0], c[1] - a[1], c[2] - a[2]};
  float x = b1[1] * b2[2] - b1[2] * b2[1];
  float y = b1[2] * b2[0] - b1[0] * b2[2];
  float z = b1[0] * b2[1] - b1[1] * b2[0];
  float r = 0.0f;
  r += x * x;
  r += y * y;
  r += z * z;
  r = std::sqrt(r);
  if (r > 0) {
    x /= r;
    y /= r;
    z /= r;
  }
  n[0] = x;
  n[1] = y;
  n[2] = z;
}

void make_device_tree(const HostTree& t, const std::vector<float>& leaf_verts, const int32_t* tri_obj,
                      const uint32_t* tri_local, DeviceTreeHost& out) {
  out = DeviceTreeHost();
  const size_t N = t.nodes.size();
  if (N == 0) {
    out.root_ref = -1;  // a leaf with an empty range
    out.leaf_start = {0, 0};
    for (int a = 0; a < 3; a++) {
      out.root_lo[a] = 1.0f;
      out.root_hi[a] = -1.0f;
    }
    out.n_leaves = 1;
    return;
  }
  std::vector<int32_t> ref(N);
  uint32_t ni = 0, nl = 0;
  for (size_t i = 0; i < N; i++) ref[i] = (t.nodes[i].right >= 0) ? (int32_t)ni++ : ~(int32_t)nl++;
  out.n_interior = ni;
  out.n_leaves = nl;
  out.nodes.assign(16 * (size_t)ni, 0.0f);
  out.leaf_start.resize(nl + 1);
  for (size_t i = 0; i < N; i++) {
    const PNode& p = t.nodes[i];
    if (p.right < 0) {
      out.leaf_start[~ref[i]] = (int32_t)p.leaf_begin;
      out.leaf_start[~ref[i] + 1] = (int32_t)p.leaf_end;
      continue;
    }
    const PNode& L = t.nodes[i + 1];
    const PNode& R = t.nodes[p.right];
    float* f = &out.nodes[16 * (size_t)ref[i]];
    f[0] = L.lo[0]; f[1] = L.lo[1]; f[2] = L.lo[2]; f[3] = L.hi[0];
    f[4] = L.hi[1]; f[5] = L.hi[2]; f[6] = R.lo[0]; f[7] = R.lo[1];
    f[8] = R.lo[2]; f[9] = R.hi[0]; f[10] = R.hi[1]; f[11] = R.hi[2];
    int32_t* r = reinterpret_cast<int32_t*>(f + 12);
    r[0] = ref[i + 1];
    r[1] = ref[p.right];
    r[2] = r[3] = 0;
  }
  memcpy(out.root_lo, t.nodes[0].lo, 12);
  memcpy(out.root_hi, t.nodes[0].hi, 12);
  out.root_ref = ref[0];
  const size_t M = t.leaf_tris.size();
  out.tv0.resize(4 * M);
  out.tv1.resize(4 * M);
  out.tv2.resize(4 * M);
  out.tn.resize(4 * M);
  for (size_t k = 0; k < M; k++) {
    const float* v = &leaf_verts[9 * k];
    uint32_t tid = t.leaf_tris[k];
    int32_t obj = tri_obj ? tri_obj[tid] : -1;
    uint32_t loc = tri_local ? tri_local[tid] : tid;
    float n[3];
    face_normal(v, v + 3, v + 6, n);
    float* a = &out.tv0[4 * k];
    float* b = &out.tv1[4 * k];
    float* c = &out.tv2[4 * k];
    float* d = &out.tn[4 * k];
    a[0] = v[0]; a[1] = v[1]; a[2] = v[2]; memcpy(&a[3], &tid, 4);
    b[0] = v[3]; b[1] = v[4]; b[2] = v[5]; memcpy(&b[3], &obj, 4);
    c[0] = v[6]; c[1] = v[7]; c[2] = v[8]; c[3] = 0.0f;
    d[0] = n[0]; d[1] = n[1]; d[2] = n[2]; memcpy(&d[3], &loc, 4);
  }
}

}  // namespace rx
