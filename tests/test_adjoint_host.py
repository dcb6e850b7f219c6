"""Host build of the general-relativistic vector-Jacobian products
(cosmos_amd/csrc/fv_gr_adjoint.hpp): every product checked against central
finite differences of the forward functions (tests/adjoint_host/gr_vjp_check.hip,
compiled for the host with hipcc; no GPU needed)."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
SRC = os.path.join(HERE, "adjoint_host", "gr_vjp_check.hip")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")


def test_gr_vjps_match_finite_differences(tmp_path):
    if not os.path.exists(HIPCC) and not shutil.which("hipcc"):
        pytest.skip("hipcc not available")
    exe = str(tmp_path / "gr_vjp_check")
    build = subprocess.run([HIPCC, "--offload-arch=gfx950", "-O2", "-std=c++17", "-ffp-contract=off",
                            "-I" + os.path.join(ROOT, "cosmos_amd", "csrc"), "-o", exe, SRC],
                           capture_output=True, text=True)
    assert build.returncode == 0, build.stderr[-2000:]
    run = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    assert run.returncode == 0, run.stdout[-4000:]
    assert "OK: 0 failure(s)" in run.stdout
    assert run.stdout.count(" ok") >= 6 * 21
