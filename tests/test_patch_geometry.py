"""The one-launch first sweep's patch geometry, checked on the host: the same
inline functions the kernel uses (orbit_internal.h) compiled into a small
host program (tests/native/patch_geometry_check.cpp) that enumerates every
patch of 17 x 13 grid sizes (even and odd nx, periodic and open x, patches
cut by the grid's edge in x and y, nx = 2) and five patch shapes.  A wrong
ring position, a column owned twice or a neighbour looked up in the wrong
place fails it."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"


@pytest.mark.skipif(not os.path.exists(HIPCC), reason="hipcc is not installed")
def test_patch_geometry_on_the_host(tmp_path):
    exe = tmp_path / "patch_geometry_check"
    src = os.path.join(ROOT, "tests", "native", "patch_geometry_check.cpp")
    subprocess.run([HIPCC, "-O1", "-std=c++17", "-I", os.path.join(ROOT, "orbit_amd", "csrc"), "-I",
                    os.path.join(ROOT, "include"), "-o", str(exe), src], check=True, capture_output=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stdout[-2000:]
    assert out.stdout.strip().endswith("0 problems"), out.stdout[-2000:]
