"""The hand-written VJPs of disco2_amd/csrc/adjoint.h against central
differences of the forward device functions they differentiate, compiled for
the host (tests/adjoint/vjp_check.cpp): prim<->cons with floors and caps,
HLL/HLLC in all four wave states for r, phi and z faces, viscous r and phi
faces with their cell mappings, gravity and geometric sources with the torque
term, the CFL time step, the phi limiter."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")


@pytest.fixture(scope="module")
def vjp_results(tmp_path_factory):
    if not os.path.exists(HIPCC) and not shutil.which("hipcc"):
        pytest.skip("hipcc not available")
    exe = str(tmp_path_factory.mktemp("vjp") / "vjp_check")
    src = os.path.join(ROOT, "tests", "adjoint", "vjp_check.cpp")
    r = subprocess.run([HIPCC, "-x", "hip", "--offload-host-only", "-O2", "-ffp-contract=off", "-std=c++17",
                        src, "-o", exe], capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stderr[-3000:]
    out = subprocess.run([exe], capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stderr
    res = {}
    for line in out.stdout.split("\n"):
        if line.strip():
            name, n, fail, worst = line.split()
            res[name] = (int(n), int(fail), float(worst))
    return res


EXPECTED = ["prim2cons", "cons2prim", "cons2prim_pfloor", "cons2prim_vcap", "hllc_dir0", "hllc_dir1", "hllc_dir2",
            "hll_dir0", "hll_dir1", "hll_dir2", "rface_visc", "phiface_visc", "sources_2d", "sources_3d", "cell_dt",
            "cell_dt_nu_const", "cell_dt_alpha", "plm_p"]


@pytest.mark.parametrize("name", EXPECTED)
def test_vjp_matches_central_differences(vjp_results, name):
    n, fail, worst = vjp_results[name]
    print(f"{name}: {n} random states, worst relative error {worst:.2e}")
    assert n >= 100
    assert fail == 0, f"{fail} of {n} states beyond 1e-6 (worst {worst:.2e})"
