"""The binary32 exp/log of the microphysics and turbulence kernels
(cress_amd/csrc/xmath.h, float-float internals, no binary64) against
glibc's expf/logf -- what the reference's FP32 build calls -- and against
the binary64 functions rounded once, on every 97th binary32 argument of the
range (exp: [-104, 88.8]; log: every positive finite float).  The host
build of the same header is tested (the device build compiles the same
source with the same flags: FMA contraction off, IEEE division)."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HIPCC = "/opt/rocm/bin/hipcc"


@pytest.fixture(scope="module")
def checker(tmp_path_factory):
    if not os.path.exists(HIPCC):
        pytest.skip("hipcc not found")
    exe = str(tmp_path_factory.mktemp("xm") / "xmath_check")
    r = subprocess.run([HIPCC, "-O2", "-ffp-contract=off", "-x", "hip", "--offload-arch=gfx950",
                        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
                        "-I", os.path.join(ROOT, "cress_amd", "csrc"),
                        os.path.join(ROOT, "tests", "xmath_check.cpp"), "-o", exe],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    return exe


@pytest.mark.parametrize("fn", ["exp", "log"])
def test_xmath_against_libm(checker, fn):
    r = subprocess.run([checker, fn, "97"], capture_output=True, text=True, timeout=300)
    n, vs_glibc, vs_double, maxulp = map(int, r.stdout.split())
    print("%s: %d arguments, %d differ from glibc, %d from binary64 rounded, max %d ulp"
          % (fn, n, vs_glibc, vs_double, maxulp))
    assert n > 1e7
    assert maxulp <= 1
    assert vs_glibc <= 3e-4 * n
    assert vs_double <= 5e-5 * n
