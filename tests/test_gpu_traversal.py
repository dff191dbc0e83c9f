"""GPU parity of the traversal kernel against the oracle's restatement of
KDTree::intersect (src/KDTree.cpp:267-339): hit triangle ids and distances bit
for bit, visit counts equal, barycentrics within 1e-6 of an FP64 solve."""
import numpy as np
import pytest

from renderx_amd import scenes

pytestmark = pytest.mark.gpu


def _rays(n, seed, lo, hi):
    rng = np.random.default_rng(seed)
    o = rng.uniform(lo, hi, (n, 3)).astype(np.float32)
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    return o + np.float32(0), d.astype(np.float32) + np.float32(0)


def _case(name):
    if name == "cornell":
        s = scenes.cornell(8, 8)
        o, d = _rays(50000, 1, -0.2, 1.2)
        return s.verts, s.tris, s.tri_obj(), o, d
    if name == "sphere":
        v, t = scenes.displaced_sphere(40, 4)
        o, d = scenes.ray_batch(1.0, 50000, 2)
        return v.astype(np.float32), t.astype(np.uint32), np.zeros(len(t), np.int32), o, d
    if name == "soup":
        s = scenes.soup_scene(2, n_tris=40000)
        o, d = _rays(50000, 3, -0.2, 1.2)
        return s.verts, s.tris, s.tri_obj(), o, d
    if name == "media_full":  # the participating-media workload's real geometry (504,448 triangles)
        s = scenes.media_scene(8, 8, freq=158, glass_freq=16)
        o, d = _rays(200000, 5, 0.0, 1.0)
        return s.verts, s.tris, s.tri_obj(), o, d
    s = scenes.media_scene(8, 8, freq=40, glass_freq=8)
    o, d = _rays(50000, 4, 0.0, 1.0)
    return s.verts, s.tris, s.tri_obj(), o, d


def _same_bits(a, b):
    return np.array_equal(np.asarray(a, np.float32).view(np.uint32), np.asarray(b, np.float32).view(np.uint32))


@pytest.mark.parametrize("name", ["cornell", "sphere", "soup", "media", "media_full"])
def test_hits_bit_exact_and_counts(rx, oracle_mod, gpu_ctx, name):
    print(f"[traversal parity] case={name}", flush=True)
    v, t, obj, o, d = _case(name)
    tree = rx.KDTree(v, t)
    ig = rx.IntersectionGPU(ctx=gpu_ctx)
    ig.loadKDTreeToShader(tree)
    ot = oracle_mod.Tree(v, t, 0)
    e_dist, e_tid, e_box, e_tri = ot.intersect(o, d, 0)
    r_dist, r_tid, r_box, r_tri = ot.intersect(o, d, 1)
    g = ig.query(o, d, bary=True, counts=True)
    assert np.array_equal(g["tid"], e_tid)
    assert _same_bits(g["dist"], e_dist)
    assert np.array_equal(g["counts"][:, 0], r_box) and np.array_equal(g["counts"][:, 1], r_tri)
    ge = ig.query(o, d, exhaustive=True, counts=True)
    assert np.array_equal(ge["tid"], e_tid) and _same_bits(ge["dist"], e_dist)
    assert np.array_equal(ge["counts"][:, 0], e_box) and np.array_equal(ge["counts"][:, 1], e_tri)
    hit = e_tid >= 0
    uv = oracle_mod.barycentrics(o, d, e_tid, v, t)
    err = np.abs(g["bary"][hit].astype(np.float64) - uv[hit])
    assert (err <= 1e-6 * np.maximum(1.0, np.abs(uv[hit]))).all()
    print(f"  hits={hit.mean():.3f} boxes/ray ordered={r_box.mean():.1f} exhaustive={e_box.mean():.1f}")


@pytest.mark.parametrize("name", ["cornell", "media", "media_full"])
def test_legality_condition_bit_exact(rx, oracle_mod, gpu_ctx, name):
    print(f"[legality parity] case={name}", flush=True)
    v, t, obj, o, d = _case(name)
    nobj = int(obj.max()) + 1
    rng = np.random.default_rng(9)
    inside = rng.integers(-1, nobj, len(o)).astype(np.int32)
    contact = np.where(rng.uniform(size=len(o)) < 0.3, inside, rng.integers(-1, nobj, len(o))).astype(np.int32)
    tree = rx.KDTree(v, t)
    ig = rx.IntersectionGPU(ctx=gpu_ctx)
    ig.loadKDTreeToShader(tree)
    ig.load_legality(obj)
    nrm = oracle_mod.face_normals(v, t)
    ot = oracle_mod.Tree(v, t, 0)
    e_dist, e_tid, _, _ = ot.intersect(o, d, 0, obj, nrm, inside, contact)
    g = ig.query(o, d, inside=inside, contact=contact)
    assert np.array_equal(g["tid"], e_tid)
    assert _same_bits(g["dist"], e_dist)
    ge = ig.query(o, d, inside=inside, contact=contact, exhaustive=True)
    assert np.array_equal(ge["tid"], e_tid)


@pytest.mark.parametrize("name", ["cornell", "soup", "media"])
def test_last_tid_skip_replays_testvisibility(rx, oracle_mod, gpu_ctx, name):
    """The seam's per-ray skip (IntersectionGPU::Ray::last_tid,
    Shader/Intersection.frag:63-64) on the pattern of Scene::testVisibility's
    re-test (src/Scene.cpp:474-505): each ray re-cast from its first hit point
    with last_tid = the hit triangle.  GPU ids and distances equal the oracle
    recursion and a brute force with the same skip, bit for bit; no ray
    returns its own triangle."""
    print(f"[last_tid skip] case={name}", flush=True)
    v, t, obj, o, d = _case(name)
    ig = rx.IntersectionGPU(ctx=gpu_ctx)
    ig.loadKDTreeToShader(rx.KDTree(v, t))
    ot = oracle_mod.Tree(v, t, 0)
    g0 = ig.query(o, d)
    h = g0["tid"] >= 0
    o2 = (o[h] + d[h] * g0["dist"][h][:, None]).astype(np.float32)
    d2, last = d[h], g0["tid"][h]
    e_dist, e_tid, _, _ = ot.intersect(o2, d2, 0, skip=last)
    b_dist, b_tid, _, _ = ot.intersect(o2, d2, 2, skip=last)
    assert np.array_equal(e_tid, b_tid) and _same_bits(e_dist, b_dist)
    g = ig.query(o2, d2, last_tid=last)
    assert np.array_equal(g["tid"], e_tid) and _same_bits(g["dist"], e_dist)
    assert not np.any(g["tid"] == last)
    ge = ig.query(o2, d2, last_tid=last, exhaustive=True)
    assert np.array_equal(ge["tid"], e_tid) and _same_bits(ge["dist"], e_dist)
    # -1 skips nothing: the same as no skip array
    gn = ig.query(o, d, last_tid=np.full(len(o), -1, np.int32))
    assert np.array_equal(gn["tid"], g0["tid"]) and _same_bits(gn["dist"], g0["dist"])
    print(f"  {h.sum()} re-cast rays, {(g['tid'] >= 0).mean():.3f} hit a further surface")


def test_load_tree_from_reference_serialization(rx, oracle_mod, gpu_ctx):
    """rx_load_tree takes KDTree::serializeForGPU's arrays unchanged (drop-in path)."""
    print("[drop-in serialized tree]", flush=True)
    v, t, obj, o, d = _case("soup")
    ot = oracle_mod.Tree(v, t, 0)
    ig = rx.IntersectionGPU(ctx=gpu_ctx)
    ig.loadKDTreeToShader(ot.serialize())
    e_dist, e_tid, _, _ = ot.intersect(o, d, 0)
    g = ig.query(o, d)
    assert np.array_equal(g["tid"], e_tid) and _same_bits(g["dist"], e_dist)


def test_empty_and_edge_batches(rx, gpu_ctx):
    print("[edge batches]", flush=True)
    s = scenes.cornell(8, 8)
    ig = rx.IntersectionGPU(ctx=gpu_ctx)
    ig.loadKDTreeToShader(rx.KDTree(s.verts, s.tris))
    g = ig.query(np.zeros((0, 3), np.float32), np.zeros((0, 3), np.float32))
    assert len(g["tid"]) == 0
    # zero direction: no face passes the box test, a clean miss
    g = ig.query(np.array([[0.5, 0.5, 0.5]], np.float32), np.zeros((1, 3), np.float32))
    assert g["tid"][0] == -1 and g["dist"][0] == -1


@pytest.mark.parametrize("n", [1, 63, 64, 65, 4095, 65537, 458751, 458753, 917505, 1000003])
def test_work_distribution_covers_every_ray(rx, oracle_mod, gpu_ctx, n):
    """Batch sizes around the work distribution's boundaries: queues short
    enough for static chunks alone (no counter), one chunk per wave exactly,
    and long ones whose counter part is cut into eight segments with ragged
    ends.  Every ray must get its own result: the exact tid and distance of
    the oracle recursion.  The same rays are first queried reversed, so a
    ray the second query skipped would keep the reversed ray's result."""
    print(f"[work distribution] n={n}", flush=True)
    sc = scenes.cornell(8, 8)
    tree = rx.KDTree(sc.verts, sc.tris)
    ig = rx.IntersectionGPU(ctx=gpu_ctx)
    ig.loadKDTreeToShader(tree)
    rng = np.random.default_rng(0x52454E44 + n)  # origins inside the box: forward and reversed results differ
    o = rng.uniform(0.05, 0.95, (n, 3)).astype(np.float32)
    d = rng.normal(size=(n, 3))
    d = (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)
    ig.query(o, -d)
    g = ig.query(o, d)
    ot = oracle_mod.Tree(sc.verts, sc.tris, 0)
    e_dist, e_tid, _, _ = ot.intersect(o, d, 0)
    assert np.array_equal(g["tid"], e_tid) and _same_bits(g["dist"], e_dist)


def test_full_size_icosphere_2p22_rays(rx, gpu_ctx):
    """The 1,003,526-triangle scene and all 2^22 rays of the hit-parity batch
    against the oracle recursion: the GPU's (tid, dist bits) stream hashed per
    2^16-ray chunk equals the committed digests of the oracle's stream
    (tests/golden/icosphere_2p22_hits.json, made by
    tests/golden/make_icosphere_hits.py), in ordered and exhaustive mode."""
    import json
    import os

    from tests.golden import make_icosphere_hits as golden

    print("[full size] icosphere 1M tris, 2^22 rays vs oracle digests", flush=True)
    with open(os.path.join(os.path.dirname(golden.__file__), "icosphere_2p22_hits.json")) as f:
        ref = json.load(f)
    sc, o, d = golden.inputs()
    assert golden.sha(o) == ref["input_sha256"]["org"] and golden.sha(d) == ref["input_sha256"]["dir"]
    assert golden.sha(sc.verts) == ref["input_sha256"]["verts"] and golden.sha(sc.tris) == ref["input_sha256"]["tris"]
    ig = rx.IntersectionGPU(ctx=gpu_ctx)
    ig.loadKDTreeToShader(rx.KDTree(sc.verts, sc.tris))
    g = ig.query(o, d)
    ms = ig.last_kernel_ms()
    print(f"  ordered traversal {ms:.2f} ms = {len(o) / ms / 1e6:.3f} Grays/s, hit rate {(g['tid'] >= 0).mean():.3f}")
    for mode, r in (("ordered", g), ("exhaustive", ig.query(o, d, exhaustive=True))):
        dig = golden.chunk_digests(r["tid"], r["dist"])
        bad = [k for k, (a, b) in enumerate(zip(dig, ref["chunk_sha256"])) if a != b]
        assert len(dig) == len(ref["chunk_sha256"]) and not bad, f"{mode}: chunks differing from the oracle: {bad}"
        assert int((r["tid"] >= 0).sum()) == ref["hits"]
