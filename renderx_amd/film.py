"""Film I/O and image comparison (host side of librxhip.so): the reference's
PFM writer / reader (src/MCRenderer.cpp:556-574, include/visualizeUtils.h:23-48),
time-stamped snapshots (src/MCRenderer.cpp:465-490), the cmp tool's RMSE
(cmp/compare.cpp:85-124) and relMSE.  Images are (H, W, 3) float32, row 0 at
the top, like PathTracerHIP.renderPixels' "mean"."""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

from ._lib import RxError, lib, ptr

CMP = os.path.join(os.path.dirname(os.path.abspath(__file__)), "rx_cmp")


def _img(a):
    a = np.ascontiguousarray(a, np.float32)
    if a.ndim != 3 or a.shape[2] != 3:
        raise ValueError("image must be (H, W, 3)")
    return a


def save_pfm(path: str, img) -> None:
    a = _img(img)
    if lib().rx_save_pfm(os.fsencode(path), a.shape[1], a.shape[0], ptr(a)) != 0:
        raise RxError(f"rx_save_pfm: cannot write {path}")


def save_snapshot(save_path: str, img, time: int, it: int) -> str:
    """Writes <save_path without extension>_time_<time>_iter_<it>.pfm; returns that name."""
    a = _img(img)
    if lib().rx_save_pfm_snapshot(os.fsencode(save_path), time, it, a.shape[1], a.shape[0], ptr(a)) != 0:
        raise RxError(f"rx_save_pfm_snapshot: cannot write for {save_path}")
    stem = save_path
    i = stem.rfind(".")
    stem = stem[:i] if i >= 0 else ""
    return f"{stem}_time_{time}_iter_{it}.pfm"


def read_pfm(path: str, clamp: bool = False) -> np.ndarray:
    w, h = C.c_int32(), C.c_int32()
    if lib().rx_read_pfm(os.fsencode(path), C.byref(w), C.byref(h), None, 0) != 0:
        raise RxError(f"rx_read_pfm: cannot read {path}")
    a = np.empty((h.value, w.value, 3), np.float32)
    if lib().rx_read_pfm(os.fsencode(path), C.byref(w), C.byref(h), ptr(a), 1 if clamp else 0) != 0:
        raise RxError(f"rx_read_pfm: cannot read {path}")
    return a


def rmse(img, ref, mask=None) -> float:
    a, r = _img(img), _img(ref)
    if a.shape != r.shape:
        raise ValueError("image sizes differ")
    m = None if mask is None else np.ascontiguousarray(mask, np.uint8)
    out = C.c_double()
    if lib().rx_image_rmse(a.shape[1], a.shape[0], ptr(a), ptr(r), ptr(m), C.byref(out)) != 0:
        raise RxError("rx_image_rmse")
    return out.value


def relmse(img, ref, eps: float = 1e-2) -> float:
    a, r = _img(img), _img(ref)
    if a.shape != r.shape:
        raise ValueError("image sizes differ")
    out = C.c_double()
    if lib().rx_image_relmse(a.shape[1], a.shape[0], ptr(a), ptr(r), eps, C.byref(out)) != 0:
        raise RxError("rx_image_relmse")
    return out.value


def cmp(input_dir: str, ref_img: str, output_file: str, mask_file: str | None = None, with_relmse=False):
    """Runs the cmp tool; returns its rows [(time, rmse[, relmse])]."""
    args = [CMP] + (["--relmse"] if with_relmse else []) + [input_dir, ref_img, output_file]
    if mask_file:
        args.append(mask_file)
    subprocess.run(args, check=True, capture_output=True)
    rows = []
    with open(output_file) as f:
        for line in f:
            v = line.split()
            rows.append((int(v[0]), *[float(x) for x in v[1:]]))
    return rows
