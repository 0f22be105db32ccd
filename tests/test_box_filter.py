"""The dataset-box filter's fast classification equals its full triangle test
(box_filter_pass vs box_filter_full in csrc/traverse.h), checked on the host: the header is
compiled as host code by hipcc with the library's floating-point flags and run on
adversarial points (tests/box_filter_check.cpp)."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")


@pytest.mark.skipif(not os.path.exists(HIPCC), reason="hipcc not present")
def test_box_filter_fast_path_equals_full_test(tmp_path):
    exe = str(tmp_path / "box_filter_check")
    subprocess.run([HIPCC, "-O2", "-std=c++17", "-ffp-contract=off", "-Xclang", "-target-feature", "-Xclang",
                    "-packed-fp32-ops", "--offload-arch=gfx950", "-I", os.path.join(ROOT, "include"),
                    "-I", os.path.join(ROOT, "rayspace3d_amd", "csrc"), "-o", exe,
                    os.path.join(HERE, "box_filter_check.cpp")], check=True, capture_output=True, timeout=300)
    r = subprocess.run([exe, "1500"], capture_output=True, text=True, timeout=300)
    print(r.stdout)
    assert r.returncode == 0, r.stdout
    line = r.stdout.strip().splitlines()[-1].split()
    # points <n> fast-path decided <k> mismatches 0
    assert int(line[1]) == 1500 * 3 * 3000 and int(line[4]) > 1_000_000 and int(line[6]) == 0
