"""CPU checks of the C ABI: the library loads, exports every function
include/rxhip.h declares, and its host tree builder reproduces the oracle's
KDTree::build + serializeForGPU bit for bit."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from renderx_amd import scenes

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    src = open(os.path.join(ROOT, "include", "rxhip.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(rx_[a-z_]+)\s*\(", src)))


def test_library_exports_every_declared_symbol(rx):
    from renderx_amd import _lib

    L = C.CDLL(_lib.LIB_PATH)
    names = _declared()
    assert len(names) >= 15
    for n in names:
        assert hasattr(L, n), n
    assert set(names) == set(_lib.SIGNATURES), "binding table out of sync with include/rxhip.h"


def test_no_device_is_a_clean_error(rx):
    import torch

    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    from renderx_amd import _lib

    h = C.c_void_p()
    assert _lib.lib().rx_create(0, C.byref(h)) == -5  # RX_ERR_NO_DEVICE


def _cases():
    c = scenes.cornell(8, 8)
    yield "cornell", c.verts, c.tris
    v, t = scenes.displaced_sphere(20, 11)
    yield "sphere", v.astype(np.float32), t.astype(np.uint32)
    s = scenes.soup_scene(1, n_tris=30000)
    yield "soup", s.verts, s.tris
    m = scenes.media_scene(8, 8, freq=24, glass_freq=6)
    yield "media", m.verts, m.tris


@pytest.mark.parametrize("abs_int", [False, True])
def test_tree_build_matches_oracle(rx, oracle_mod, abs_int):
    for name, v, t in _cases():
        a = rx.KDTree(v, t, abs_int=abs_int).serializeForGPU()
        b = oracle_mod.Tree(v, t, 1 if abs_int else 0).serialize()
        for x, y in zip(a, b):
            assert x.shape == y.shape, name
            assert np.array_equal(x.view(np.uint32), y.view(np.uint32)), name


def test_tree_build_large_parallel_matches_oracle(rx, oracle_mod):
    # big enough for the builder's threaded subtrees
    v, t = scenes.displaced_sphere(64, 5)
    v, t = v.astype(np.float32), t.astype(np.uint32)
    a = rx.KDTree(v, t).serializeForGPU()
    b = oracle_mod.Tree(v, t, 0).serialize()
    for x, y in zip(a, b):
        assert np.array_equal(x.view(np.uint32), y.view(np.uint32))


def test_tree_build_rejects_bad_indices(rx):
    from renderx_amd import _lib

    v = np.zeros((3, 3), np.float32)
    t = np.array([[0, 1, 5]], np.uint32)
    with pytest.raises(_lib.RxError):
        rx.KDTree(v, t)
