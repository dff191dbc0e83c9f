"""Writes icosphere_2p22_hits.json: the oracle recursion's nearest hits
(KDTree::intersect, src/KDTree.cpp:295-339) for the full hit-parity batch of
the 1,003,526-triangle displaced icosphere scene, 2^22 seeded rays (half from
a radius-3 sphere aimed at the unit sphere, half uniform in [-2,2]^3 with
uniform directions).  The 32 MB result stream is kept as SHA-256 digests of
(tid int32 LE || dist float32 bits LE) per 2^16-ray chunk, with digests of the
inputs so that a change of the scene or ray generator is reported as such.

Run from the repository root:  python tests/golden/make_icosphere_hits.py
"""
import hashlib
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "icosphere_2p22_hits.json")
CHUNK = 1 << 16
N_RAYS = 1 << 22
RAY_SEED = 0x52454E44 + 1


def inputs():
    from renderx_amd import scenes

    sc = scenes.icosphere_scene(8, 8)
    o, d = scenes.ray_batch(1.0, N_RAYS, RAY_SEED)
    return sc, o, d


def sha(*arrays):
    h = hashlib.sha256()
    for a in arrays:
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


def chunk_digests(tid, dist):
    tid = np.ascontiguousarray(tid, "<i4")
    bits = np.ascontiguousarray(dist, "<f4").view("<u4")
    return [sha(tid[k:k + CHUNK], bits[k:k + CHUNK]) for k in range(0, len(tid), CHUNK)]


def main():
    from oracle import oracle

    t = time.time()
    sc, o, d = inputs()
    tree = oracle.Tree(sc.verts, sc.tris, 0)
    dist, tid, _, _ = tree.intersect(o, d, 0)
    rec = {
        "what": "oracle KDTree::intersect recursion (exhaustive, no condition) on the icosphere hit-parity batch",
        "n_tris": int(sc.n_tris), "n_rays": N_RAYS, "ray_seed": RAY_SEED, "chunk": CHUNK,
        "input_sha256": {"verts": sha(sc.verts), "tris": sha(sc.tris), "org": sha(o), "dir": sha(d)},
        "hits": int((tid >= 0).sum()),
        "chunk_sha256": chunk_digests(tid, dist),
    }
    with open(OUT, "w") as f:
        json.dump(rec, f, indent=1)
    print(f"wrote {OUT}: {rec['hits']} hits of {N_RAYS} rays in {time.time() - t:.1f} s")


if __name__ == "__main__":
    main()
