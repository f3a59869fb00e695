"""Host scoring (h1, psnr of an AB map at one t) as the library computes it,
against the straightforward form of get_h_psnr.tpp:5-31, bit for bit, on
random maps whose quantizer bins fall inside and outside the flat count
array (including keys whose bins overflow int64).  Built with the compiler
and flags the library's host code uses."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "..", "asz_amd", "csrc")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
HIPCC = os.path.join(ROCM, "bin", "hipcc")


@pytest.mark.skipif(not os.path.exists(HIPCC), reason="hipcc is not installed")
def test_flat_bin_scoring_equals_direct_map(tmp_path):
    exe = str(tmp_path / "scoring_harness")
    cmd = [HIPCC, "-x", "c++", "-std=c++17", "-O2", "-ffp-contract=off", "-I", CSRC,
           os.path.join(HERE, "scoring_harness.cpp"), os.path.join(CSRC, "preprocess.cpp"), "-o", exe]
    subprocess.run(cmd, check=True, capture_output=True, timeout=300)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=300)
    print(r.stdout)
    assert r.returncode == 0, r.stdout
    cases, far = int(r.stdout.split()[1]), int(r.stdout.split()[3])
    assert cases > 1000 and far > 100
