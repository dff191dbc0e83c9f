"""A C program compiled against include/rxhip.h and linked to librxhip.so:
compile-time checks of the struct layouts the adapters and the Python
bindings share, then calls through the C ABI (host entry points here; the
device entry points, the last_tid skip among them, on the GPU)."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIBDIR = os.path.join(ROOT, "renderx_amd")


def _build(tmp_path):
    exe = str(tmp_path / "capi_check")
    subprocess.run(["gcc", "-std=c11", "-O1", "-Wall", "-Werror", "-o", exe,
                    os.path.join(HERE, "native", "capi_check.c"), "-L" + LIBDIR, "-lrxhip",
                    "-Wl,-rpath," + LIBDIR], check=True, capture_output=True, text=True)
    return exe


def _run(exe, mode):
    r = subprocess.run([exe, mode], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0 and "ALL OK" in r.stdout, r.stdout + r.stderr


@pytest.mark.skipif(shutil.which("gcc") is None, reason="gcc not available")
def test_c_caller_host(rx, tmp_path):
    import torch

    if torch.cuda.is_available():
        pytest.skip("the no-device branch needs a machine without a GPU")
    _run(_build(tmp_path), "cpu")


@pytest.mark.gpu
def test_c_caller_device(rx, tmp_path):
    print("[C caller] rx_create / rx_load_tree / rx_intersect", flush=True)
    _run(_build(tmp_path), "gpu")
