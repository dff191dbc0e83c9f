"""Film I/O and comparison (renderx_amd/film.py over librxhip.so, and the
rx_cmp tool): the PFM bytes MCRenderer::saveImagePFM writes
(src/MCRenderer.cpp:556-574), the reader (include/visualizeUtils.h:23-48,
channel clamp [0, 4]), snapshot names (src/MCRenderer.cpp:465-490), and the
cmp tool's RMSE (cmp/compare.cpp:85-124, 242-331) pinned by closed forms and
by a float restatement of its operation order.  No GPU needed."""
import os
import struct

import numpy as np
import pytest

film = pytest.importorskip("renderx_amd.film")


def _img(h, w, seed=0, lo=0.0, hi=3.0):
    return np.random.default_rng(seed).uniform(lo, hi, (h, w, 3)).astype(np.float32)


def test_pfm_bytes_match_the_reference_writer(tmp_path):
    img = np.arange(2 * 3 * 3, dtype=np.float32).reshape(3, 2, 3)  # H=3, W=2
    p = str(tmp_path / "a.pfm")
    film.save_pfm(p, img)
    raw = open(p, "rb").read()
    head = b"PF\n2 3\n-1.000000\n"
    assert raw[:len(head)] == head
    body = raw[len(head):]
    assert len(body) == 2 * 3 * 3 * 4
    # rows bottom to top, R G B little endian
    expect = b"".join(struct.pack("<3f", *img[y, x]) for y in (2, 1, 0) for x in (0, 1))
    assert body == expect


def test_pfm_round_trip_and_clamp(tmp_path):
    img = _img(7, 5, lo=-1.0, hi=6.0)
    p = str(tmp_path / "b.pfm")
    film.save_pfm(p, img)
    assert np.array_equal(film.read_pfm(p), img)
    assert np.array_equal(film.read_pfm(p, clamp=True), np.clip(img, 0, 4))


def test_big_endian_pfm(tmp_path):
    img = _img(2, 3)
    p = str(tmp_path / "be.pfm")
    with open(p, "wb") as f:
        f.write(b"PF\n3 2\n1.0\n")
        for y in (1, 0):
            for x in range(3):
                f.write(struct.pack(">3f", *img[y, x]))
    assert np.array_equal(film.read_pfm(p), img)


def _rmse_restated(img, ref, mask=None):
    """CalculateRMSE's operations: float difference and length (B, G, R order),
    double square and sum, x-major pixel loop."""
    h, w, _ = img.shape
    tot, cnt = 0.0, 0
    for x in range(w):
        for y in range(h):
            if mask is not None and not mask[y, x]:
                continue
            d = np.float32(0)
            for c in (2, 1, 0):
                dc = np.float32(ref[y, x, c] - img[y, x, c])
                d = np.float32(d + np.float32(dc * dc))
            ln = float(np.sqrt(d, dtype=np.float32))
            tot += ln * ln
            cnt += 1
    return (tot / cnt) ** 0.5


def test_rmse_closed_form_and_operation_order():
    ref = _img(9, 11, seed=1, lo=0.5, hi=3.0)
    delta = np.float32(0.25)
    img = ref + delta
    r = film.rmse(img, ref)
    assert abs(r - np.sqrt(3) * 0.25) < 1e-6
    noisy = _img(9, 11, seed=2)
    assert film.rmse(noisy, ref) == _rmse_restated(noisy, ref)
    mask = np.zeros((9, 11), np.uint8)
    mask[::2, 1::3] = 1
    assert film.rmse(noisy, ref, mask) == _rmse_restated(noisy, ref, mask)
    assert film.rmse(ref, ref) == 0.0


def test_relmse_closed_form():
    ref = np.full((4, 5, 3), 2.0, np.float32)
    img = ref * 1.5
    # (1.0)^2 / (4 + eps) on every channel, eps passed as float32
    assert abs(film.relmse(img, ref) - 1.0 / (4.0 + float(np.float32(0.01)))) < 1e-12


def test_snapshot_names_and_cmp_tool(tmp_path):
    if not os.path.exists(film.CMP):
        pytest.skip("rx_cmp not built")
    d = tmp_path / "results"
    d.mkdir()
    ref = _img(6, 8, seed=3, lo=0.5, hi=2.0)
    refp = str(tmp_path / "ref.pfm")
    film.save_pfm(refp, ref)
    expect = {}
    for t, it, delta in [(30, 3, 0.3), (10, 1, 0.1), (20, 2, 0.2)]:
        name = film.save_snapshot(str(d / "cb_PT.pfm"), ref + np.float32(delta), t, it)
        assert os.path.basename(name) == f"cb_PT_time_{t}_iter_{it}.pfm" and os.path.exists(name)
        expect[t] = np.sqrt(3) * delta
    (d / "no_stamp.pfm").write_bytes(open(refp, "rb").read())  # skipped: no time / iter tokens
    rows = film.cmp(str(d), refp, str(tmp_path / "rmse.txt"), with_relmse=True)
    assert [r[0] for r in rows] == [10, 20, 30]
    for t, e, rel in rows:
        assert abs(e - expect[t]) < 1e-6
        assert rel > 0
    assert open(tmp_path / "rmse.txt").read().splitlines()[0].startswith("10 ")
    # mask listed from the bottom row up: only the bottom row counts
    m = tmp_path / "mask.txt"
    m.write_text("\n".join(" ".join("1" if y == 0 else "0" for _ in range(8)) for y in range(6)) + "\n")
    img = ref.copy()
    img[5] += np.float32(1.0)  # the bottom row (y = H-1) differs by 1
    film.save_snapshot(str(d / "m.pfm"), img, 99, 9)
    rows = film.cmp(str(d), refp, str(tmp_path / "rmse2.txt"), mask_file=str(m))
    assert abs(rows[-1][1] - np.sqrt(3)) < 1e-6
