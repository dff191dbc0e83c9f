"""The traversal's slab filter (slab_classify in renderx_amd/csrc/rx_dev.h)
never rules out a box that the reference box predicate (src/KDTree.cpp:34-52)
accepts, and never certifies one it rejects: checked on the CPU over random
and adversarial rays (aimed at faces, edges and corners with ulp nudges,
origins on face planes, near axis-parallel directions, flat boxes, origins
far outside the scene box); its
ordering key equals slab_tnear's."""
import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")


@pytest.mark.skipif(not os.path.exists(HIPCC), reason="hipcc not available")
def test_certain_miss_filter_is_conservative(tmp_path):
    exe = str(tmp_path / "box_filter_check")
    subprocess.run([HIPCC, "-O2", "-std=c++17", "-ffp-contract=off", "-fhip-fp32-correctly-rounded-divide-sqrt",
                    "-o", exe, os.path.join(HERE, "native", "box_filter_check.cpp")], check=True,
                   capture_output=True)
    r = subprocess.run([exe, "400000", "11"], capture_output=True, text=True)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    lines = [l for l in r.stdout.splitlines() if l.startswith("family")]
    assert len(lines) == 8
    # the filter must also be useful: random pairs are almost all ruled out
    fam0 = lines[0].split()
    vals = {fam0[i]: int(fam0[i + 1]) for i in range(2, len(fam0) - 1, 2)}
    assert vals["certain_miss"] >= 0.99 * vals["exact_miss"]
    fam6 = lines[6].split()
    v6 = {fam6[i]: int(fam6[i + 1]) for i in range(2, len(fam6) - 1, 2)}
    assert v6["certain_hit"] >= 0.3 * v6["exact_hit"]
