"""cfl_above (flow_amd/csrc/kernels/hydro_math.h), the test that lets the
Lagrangian marches leave out a CFL candidate that cannot lower a lane's
running minimum, compiled for the host against the candidates themselves
(cell_timestep in FP64, cell_timestep_sq in the FP32-state form): over
random cells, limits within a few ulps of each candidate and at the test's
margin, and special values, no cell it lets through may have a NaN
candidate or one below the limit, so the minimum and the count of
undefined candidates are bitwise the same with the cells left out."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "hostcheck", "cfl_skip_host.cpp")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")


def test_cfl_above_never_drops_a_lower_or_undefined_candidate(tmp_path):
    if not shutil.which(HIPCC) and not os.path.exists(HIPCC):
        pytest.skip("hipcc not found")
    exe = str(tmp_path / "cfl_skip_host")
    subprocess.run([HIPCC, "-O2", "-std=c++17", SRC, "-o", exe], check=True,
                   capture_output=True, timeout=600)
    out = subprocess.run([exe], capture_output=True, text=True, check=True, timeout=120)
    n64, pass64, bad64, n32, pass32, bad32 = map(int, out.stdout.split())
    print(f"FP64: {pass64} of {n64} cases left out; FP32 state: {pass32} of {n32}")
    assert n64 > 40_000_000 and n32 > 40_000_000
    # the test must let a real share through, or it saves nothing
    assert pass64 > n64 // 10 and pass32 > n32 // 10
    assert bad64 == 0 and bad32 == 0
