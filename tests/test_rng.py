"""The renderer's per-path random numbers (Philox4x32-10 in
renderx_amd/csrc/rx_dev.h, no counterpart in the reference, which shares one
MT19937 across threads, src/RandGenerator.cpp:9): the published known-answer
vectors (Salmon et al., SC'11; Random123) and the independence of the streams
keyed by (pixel, seed) with counter (sample, bounce, stream).  The product's
own header is compiled for the host and checked there."""
import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")


@pytest.mark.skipif(not os.path.exists(HIPCC), reason="hipcc not available")
def test_philox_known_answers_and_streams(tmp_path):
    exe = str(tmp_path / "philox_kat")
    subprocess.run([HIPCC, "-O2", "-std=c++17", "-o", exe, os.path.join(HERE, "native", "philox_kat.cpp")],
                   check=True, capture_output=True)
    r = subprocess.run([exe], capture_output=True, text=True)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.count(" ok") == 5
