"""Pins of the oracle's geometry restatement (src/KDTree.cpp, src/NoSelfIntersectionCondition.cpp)
against facts that do not depend on the restatement: brute force, an FP64
ray/plane solve, hand-computed box cases, structural tree invariants."""
import math

import numpy as np
import pytest

from renderx_amd import scenes


def _tree(oracle_mod, v, t, abs_mode=0):
    return oracle_mod.Tree(v, t, abs_mode)


def _random_rays(n, seed, lo=-0.2, hi=1.2):
    rng = np.random.default_rng(seed)
    o = rng.uniform(lo, hi, (n, 3)).astype(np.float32)
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    return o, d.astype(np.float32)


def test_mt19937_known_answer(oracle_mod):
    # C++ standard [rand.predef]: the 10000th output of default-seeded mt19937 is 4123659995
    assert oracle_mod.mt_nth(10000) == 4123659995


def test_single_triangle_closed_form(oracle_mod):
    v = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float32)
    t = np.array([[0, 1, 2]], np.uint32)
    tr = _tree(oracle_mod, v, t)
    o = np.array([[0.25, 0.25, 2.0], [0.25, 0.25, -3.0], [0.6, 0.6, 1.0], [0.2, 0.2, 1.0], [0.0, 0.0, 1.0]], np.float32)
    d = np.array([[0, 0, -1], [0, 0, 1], [0, 0, -1], [1, 0, 0], [0, 0, -1]], np.float32)
    dist, tid, _, _ = tr.intersect(o, d, 0)
    assert dist[0] == 2.0 and tid[0] == 0          # front face, exact distance
    assert dist[1] == 3.0 and tid[1] == 0          # double-sided test (both mask signs)
    assert tid[2] == -1                            # outside the hypotenuse
    assert tid[3] == -1                            # parallel to the plane: den == 0
    assert dist[4] == 1.0 and tid[4] == 0          # vertex hit counts (inclusive masks)


def test_triangle_distance_matches_fp64(oracle_mod):
    s = scenes.icosphere(8)
    tr = _tree(oracle_mod, s[0].astype(np.float32), s[1].astype(np.uint32))
    rng = np.random.default_rng(5)
    n = 4000
    o = (rng.normal(size=(n, 3)) * 3).astype(np.float32)
    tgt = rng.normal(size=(n, 3)) * 0.3
    d = tgt - o
    d = (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)
    dist, tid, _, _ = tr.intersect(o, d, 0)
    hit = tid >= 0
    assert hit.mean() > 0.5
    V = s[0].astype(np.float32).astype(np.float64)
    T = s[1]
    P = V[T[tid[hit]]]
    n_ = np.cross(P[:, 1] - P[:, 0], P[:, 2] - P[:, 0])
    od = o[hit].astype(np.float64)
    dd = d[hit].astype(np.float64)
    t64 = np.einsum("ij,ij->i", n_, P[:, 0] - od) / np.einsum("ij,ij->i", n_, dd)
    cosang = np.abs(np.einsum("ij,ij->i", n_ / np.linalg.norm(n_, axis=1, keepdims=True), dd))
    ok = cosang > 0.05
    rel = np.abs(dist[hit][ok] - t64[ok]) / t64[ok]
    assert rel.max() < 2e-6


def test_box_predicate_hand_cases(oracle_mod):
    # one triangle spanning the unit square in z=0: the root box is flat in z
    v = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0]], np.float32)
    t = np.array([[0, 1, 2], [0, 2, 3]], np.uint32)
    tr = _tree(oracle_mod, v, t)
    o = np.array([[0.5, 0.5, 1.0], [2.0, 0.5, 1.0], [0.5, 0.5, 1.0]], np.float32)
    d = np.array([[0, 0, -1], [0, 0, -1], [1, 0, 0]], np.float32)
    dist, tid, boxes, _ = tr.intersect(o, d, 0)
    assert tid[0] >= 0 and dist[0] == 1.0
    assert tid[1] == -1 and boxes[1] == 1     # misses the root box: one box test
    assert tid[2] == -1                       # parallel to the flat box: zero-direction faces skipped


@pytest.mark.parametrize("name", ["cornell", "sphere", "soup"])
def test_exhaustive_equals_bruteforce_and_ordered(oracle_mod, name):
    if name == "cornell":
        sc = scenes.cornell(8, 8)
        v, t = sc.verts, sc.tris
        o, d = _random_rays(20000, 1)
    elif name == "sphere":
        v, t = scenes.displaced_sphere(12, 9)
        v, t = v.astype(np.float32), t.astype(np.uint32)
        o, d = scenes.ray_batch(1.0, 20000, 2)
    else:
        sc = scenes.soup_scene(3, n_tris=20000)
        v, t = sc.verts, sc.tris
        o, d = _random_rays(20000, 3)
    tr = _tree(oracle_mod, v, t)
    e = tr.intersect(o, d, 0)
    b = tr.intersect(o, d, 2)
    od = tr.intersect(o, d, 1)
    # the ordered traversal returns the recursion's winner bit for bit
    assert np.array_equal(e[1], od[1])
    assert np.array_equal(e[0].view(np.uint32), od[0].view(np.uint32))
    # brute force differs only where the box predicate prunes a grazing hit
    agree = (e[1] == b[1]).mean()
    assert agree > 0.999
    assert (od[2] <= e[2]).all()   # culling never tests more boxes


def _leaves_and_boxes(ser):
    nodes, mn, mx, v1, v2, v3 = ser
    return nodes, mn, mx, v1, v2, v3


@pytest.mark.parametrize("abs_mode", [0, 1])
def test_tree_structure_invariants(oracle_mod, abs_mode):
    v, t = scenes.displaced_sphere(6, 3)
    v = v.astype(np.float32)
    t = t.astype(np.uint32)
    tr = _tree(oracle_mod, v, t, abs_mode)
    nodes, mn, mx, v1, v2, v3 = tr.serialize()
    n = len(nodes)
    # every triangle in exactly one leaf
    ids = np.sort(v1[:, 3].astype(np.int64))
    assert np.array_equal(ids, np.arange(len(t)))
    # pre-order: left child = i+1, parent links consistent, tight boxes
    def tris_of(i):
        if nodes[i, 1] < 0:
            return v1[int(nodes[i, 2]):int(nodes[i, 3]), 3].astype(np.int64)
        return np.concatenate([tris_of(i + 1), tris_of(int(nodes[i, 1]))])
    import sys
    sys.setrecursionlimit(10000)
    for i in range(n):
        ti = tris_of(i)
        pts = v[t[ti].reshape(-1)]
        assert np.array_equal(mn[i, :3], pts.min(0)) and np.array_equal(mx[i, :3], pts.max(0))
        if nodes[i, 1] >= 0:
            assert nodes[i + 1, 0] == i and nodes[int(nodes[i, 1]), 0] == i
            L, R = tris_of(i + 1), tris_of(int(nodes[i, 1]))
            # a median split on one axis separates the children: some axis/value puts
            # every left triangle's min vertex <= split < every right triangle's vertices
            ok = False
            for a in (range(1) if abs_mode == 1 else range(3)):
                lmax_of_min = v[t[L]][:, :, a].min(1).max()
                rmin = v[t[R]][:, :, a].min()
                ok |= bool(lmax_of_min < rmin)
            assert ok
        else:
            # a leaf holds one triangle unless no candidate axis separates it: at the
            # median of the node's unique vertices every triangle has a vertex <= median
            if len(ti) > 1:
                for a in (range(1) if abs_mode == 1 else range(3)):
                    vid = np.unique(t[ti].reshape(-1))
                    coords = np.sort(v[vid, a])
                    split = coords[len(coords) // 2]
                    assert (v[t[ti]][:, :, a] <= split).any(1).all()


def test_legality_condition_cases(oracle_mod):
    # two parallel quads facing +z at z=0 (object 0) and z=1 (object 1)
    v = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0], [0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1]], np.float32)
    t = np.array([[0, 1, 2], [0, 2, 3], [4, 5, 6], [4, 6, 7]], np.uint32)
    tr = _tree(oracle_mod, v, t)
    obj = np.array([0, 0, 1, 1], np.int32)
    nrm = oracle_mod.face_normals(v, t)
    assert np.allclose(nrm[:, 2], 1.0)
    o = np.array([[0.5, 0.5, 2.0], [0.5, 0.5, 0.5], [0.5, 0.5, 0.5], [0.5, 0.5, -1.0]], np.float32)
    d = np.array([[0, 0, -1], [0, 0, 1], [0, 0, 1], [0, 0, 1]], np.float32)
    # ray 0: from outside, contact none, inside none -> (inside == contact): accepts everything
    # ray 1: leaves object 0 upward (contact 0, inside -1): back face of object 1 (d.n>=0) is illegal
    # ray 2: inside object 1 (inside 1, contact 0): its own back face is legal
    # ray 3: contact 0, inside -1, going up through object 0's back face: illegal, then object 1 illegal
    inside = np.array([-1, -1, 1, -1], np.int32)
    contact = np.array([-1, 0, 0, 0], np.int32)
    dist, tid, _, _ = tr.intersect(o, d, 0, obj, nrm, inside, contact)
    assert tid[0] in (2, 3) and dist[0] == 1.0
    assert tid[1] == -1
    assert tid[2] in (2, 3) and dist[2] == 0.5
    assert tid[3] == -1


@pytest.mark.parametrize("scale,shift", [(1e3, 0.0), (1e3, 5e3), (1e4, 1e5)])
def test_ordered_cull_margin_at_large_coordinates(oracle_mod, scale, shift):
    """The ordered traversal culls a subtree only when its slab entry exceeds
    1.001 x the best distance (the GPU's rule, mirrored by the oracle's
    ordered mode).  Secondary rays start on a surface, where hit distances are
    short against the coordinates' magnitude; on a scene scaled and moved far
    from the origin, with half of the rays near-grazing their start triangle,
    the ordered winner must still be the reference recursion's, with the
    legality condition and with the shader's last_tid skip."""
    from renderx_amd import scenes

    sc = scenes.icosphere_scene(8, 8, freq=24)
    v = (sc.verts.astype(np.float64) * scale + shift).astype(np.float32)
    tr = oracle_mod.Tree(v, sc.tris, 0)
    rng = np.random.default_rng(1)
    n = 40000
    t = rng.integers(0, len(sc.tris), n)
    P = v.astype(np.float64)[sc.tris[t]]
    a, b = rng.random(n), rng.random(n)
    m = a + b > 1
    a[m], b[m] = 1 - a[m], 1 - b[m]
    o = (P[:, 0] + a[:, None] * (P[:, 1] - P[:, 0]) + b[:, None] * (P[:, 2] - P[:, 0])).astype(np.float32)
    nrm = oracle_mod.face_normals(v, sc.tris)
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    nn = nrm[t].astype(np.float64)
    dp = d - (d * nn).sum(1, keepdims=True) * nn
    dp /= np.linalg.norm(dp, axis=1, keepdims=True)
    dg = dp + (10 ** rng.uniform(-6, -2, n))[:, None] * nn * np.sign(rng.random(n) - 0.5)[:, None]
    g = rng.random(n) < 0.5
    d[g] = dg[g] / np.linalg.norm(dg[g], axis=1, keepdims=True)
    d = d.astype(np.float32)
    obj = sc.tri_obj()
    inside, contact = np.full(n, -1, np.int32), obj[t].astype(np.int32)
    e = tr.intersect(o, d, 0, obj, nrm, inside, contact)
    r = tr.intersect(o, d, 1, obj, nrm, inside, contact)
    assert np.array_equal(e[1], r[1]) and np.array_equal(e[0].view(np.uint32), r[0].view(np.uint32))
    assert (e[1] >= 0).mean() > 0.5
    e = tr.intersect(o, d, 0, skip=t.astype(np.int32))
    r = tr.intersect(o, d, 1, skip=t.astype(np.int32))
    assert np.array_equal(e[1], r[1]) and np.array_equal(e[0].view(np.uint32), r[0].view(np.uint32))
    assert not np.any(e[1] == t)  # the skipped triangle is never the hit


def test_skip_tid_matches_brute_force(oracle_mod):
    """IntersectionGPU::Ray::last_tid (Shader/Intersection.frag:63-64): the
    named triangle is passed over; the recursion equals brute force with the
    same skip, and rays re-cast from their hit point with the hit triangle
    skipped (Scene::testVisibility's re-test, src/Scene.cpp:474-505) find the
    next surface, never their own triangle at distance 0."""
    from renderx_amd import scenes

    sc = scenes.cornell(8, 8)
    tr = oracle_mod.Tree(sc.verts, sc.tris, 0)
    rng = np.random.default_rng(5)
    o = rng.uniform(0.05, 0.95, (20000, 3)).astype(np.float32)
    d = rng.normal(size=(20000, 3))
    d = (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)
    dist, tid, _, _ = tr.intersect(o, d, 0)
    h = tid >= 0
    o2 = (o[h] + d[h] * dist[h][:, None]).astype(np.float32)
    e = tr.intersect(o2, d[h], 0, skip=tid[h])
    b = tr.intersect(o2, d[h], 2, skip=tid[h])
    assert np.array_equal(e[1], b[1]) and np.array_equal(e[0].view(np.uint32), b[0].view(np.uint32))
    assert not np.any(e[1] == tid[h])
    # without the skip most re-cast rays hit their own triangle again
    z = tr.intersect(o2, d[h], 0)
    assert (z[1] == tid[h]).mean() > 0.3


def test_golden_icosphere_hits_reproduce(oracle_mod):
    """The committed 2^22-ray digests (tests/golden/icosphere_2p22_hits.json)
    match this checkout's scene and ray generator, and the oracle reproduces
    the digests of four of the chunks (the GPU test compares all 64)."""
    import json
    import os

    from tests.golden import make_icosphere_hits as golden

    with open(os.path.join(os.path.dirname(golden.__file__), "icosphere_2p22_hits.json")) as f:
        ref = json.load(f)
    sc, o, d = golden.inputs()
    assert golden.sha(o) == ref["input_sha256"]["org"] and golden.sha(d) == ref["input_sha256"]["dir"]
    assert golden.sha(sc.verts) == ref["input_sha256"]["verts"] and golden.sha(sc.tris) == ref["input_sha256"]["tris"]
    tree = oracle_mod.Tree(sc.verts, sc.tris, 0)
    C = ref["chunk"]
    for k in (0, 17, 40, 63):
        dist, tid, _, _ = tree.intersect(o[k * C:(k + 1) * C], d[k * C:(k + 1) * C], 0)
        assert golden.chunk_digests(tid, dist) == [ref["chunk_sha256"][k]]
