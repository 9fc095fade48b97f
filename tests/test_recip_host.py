"""recip_pow2 (flow_amd/csrc/kernels/common.h), the exponent-arithmetic
reciprocal the march kernels use for cell heights on power-of-two meshes,
compiled for the host and compared bit for bit with the IEEE quotient 1/x on
every input that path admits: 2^k for |k| <= 300 in FP64 and |k| <= 40 in
FP32 (the exponent bounds of the host's power-of-two test)."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "hostcheck", "recip_host.cpp")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")


def test_recip_pow2_is_the_ieee_quotient(tmp_path):
    if not shutil.which(HIPCC) and not os.path.exists(HIPCC):
        pytest.skip("hipcc not found")
    exe = str(tmp_path / "recip_host")
    subprocess.run([HIPCC, "-O2", "-std=c++17", SRC, "-o", exe], check=True,
                   capture_output=True, timeout=600)
    out = subprocess.run([exe], capture_output=True, text=True, check=True, timeout=60)
    n64, bad64, n32, bad32 = map(int, out.stdout.split())
    assert (n64, n32) == (601, 81)
    assert bad64 == 0 and bad32 == 0
