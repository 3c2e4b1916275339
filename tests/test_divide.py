"""Host check of the kernels' division by a loop-constant divisor
(tacuda_amd/csrc/divide.hpp): Markstein's correction of x * RN(1/d) equals
IEEE division bit for bit for every integer period 1..100000 and every WMA
weight total p(p+1)/2, over random, float32-sum and signed-zero / Inf / NaN
numerators (tests/div_check.cpp, built for the host with hipcc and run, no
GPU)."""
import os
import shutil
import subprocess
import tempfile

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")


@pytest.mark.skipif(not (os.path.exists(HIPCC) or shutil.which("hipcc")), reason="no hipcc")
def test_div_by_equals_ieee_division():
    with tempfile.TemporaryDirectory() as d:
        exe = os.path.join(d, "div_check")
        subprocess.run([HIPCC, "-x", "hip", "--offload-host-only", "-O2", "-std=c++17", "-o", exe,
                        os.path.join(HERE, "div_check.cpp")], check=True)
        out = subprocess.run([exe], capture_output=True, text=True, timeout=300)
        assert out.returncode == 0, out.stdout + out.stderr
        assert "div_by ok" in out.stdout
