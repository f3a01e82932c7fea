"""The reference's shipped parfiles end to end on the GPU.

* Parity: each shipped parfile (parfiles/code_tests/{shear,vortex,SStest}.par,
  parfiles/torus.par, read from tests/golden/code_*), its initial data from
  Batch.initial_data(), a few RK2 steps against the oracle on the same
  steps (the oracle reproduces the reference's digests of these runs byte
  for byte, tests/test_code_tests.py): bit for bit for the vortex, within
  the per-zone tolerance of the other transcendental cases.
* Physics, to the parfiles' own Time_Max, against what the reference's
  plotting scripts compare with:
  - shear.par: the viscous spreading of the Gaussian shear layer
    (plotting_scripts/plot_shear.py:44-52): v_y(x) on the slice |y - 3| < 0.1
    against 1/sqrt(2 pi nu t) exp(-(x-3)^2 / (4 nu t)), t = t0 + T;
  - vortex.par: the vortex is a steady state; v_phi, P and rho after T = 20.
"""
import numpy as np
import pytest

from compare import bitwise_equal, describe, zone_rel
from golden_util import Golden, code_names

pytestmark = pytest.mark.gpu

RTOL = 1e-11
# SStest.par is stiff (its cooling term goes as (P/rho)^4 / PoRho_r1^3 with
# PoRho_r1 = 0.01) and its A_KEPLER frame leaves omega a small difference of
# large terms: on the oracle alone, one ulp more in the initial rho and P
# moves the state by 8.5e-11 per zone after its 40 steps (1.9e-11 after
# one); the GPU's ulp-level libm differences move it by 3.4e-10
TOLERANCE = {"code_SStest": 1e-8}


def start(name):
    from disco2_amd.api import Batch, Setup
    g = Golden(name)
    s = Setup(g.params)
    b = Batch([s])
    b.initial_data()
    b.init()
    return g, s, b


@pytest.mark.parametrize("name", code_names())
def test_shipped_parfile_matches_oracle(name):
    from oracle.oracle import Oracle
    g, s, b = start(name)
    steps = int(g.get("digest/steps")[0])
    print(f"{name}: {s.nzones} zones, {steps} steps", flush=True)
    o = Oracle(g.params, s.N_p, s.r_faces, s.z_faces)
    o.set_state(tiph=s.tiph0, prim=np.zeros((s.NQ, s.nzones)))
    o.initial_data()
    o.init()
    assert bitwise_equal(b.download("prim"), o.get("prim"))
    b.run(steps)
    o.step(steps)
    mine, ref = b.download("prim"), o.get("prim")
    if name == "code_vortex":
        assert bitwise_equal(mine, ref), describe(mine, ref)
        assert bitwise_equal(b.download("tiph"), o.get("tiph"))
    else:
        tol = TOLERANCE.get(name, RTOL)
        assert zone_rel(mine, ref) < tol, f"rel {zone_rel(mine, ref)}"
    assert abs(b.scalars()[0] - o.scalars()[0]) <= TOLERANCE.get(name, RTOL) * o.scalars()[0]
    b.close()


def run_to_end(b, T):
    """steps until t reaches T (the step that would pass Time_Max is cut to
    end on it, timestep_set.c:16-21; steps after it have dt = 0)"""
    n = 0
    while b.scalars()[0] < T * (1. - 1e-14):
        b.run(200)
        n += 200
        assert n < 100000
    assert abs(b.scalars()[0] - T) <= 1e-14 * T
    return n


def cell_centres(s, b):
    rc = 0.5 * (s.r_faces[:-1] + s.r_faces[1:])
    r = np.tile(np.repeat(rc, s.N_p), s.NZ)
    return r, b.download("tiph") - 0.5 * b.download("dphi")


def test_shear_layer_spreads_as_the_diffusion_solution():
    g, s, b = start("code_shear")
    T = g.params["Time_Max"]
    n = run_to_end(b, T)
    print(f"shear.par: t = {T} after <= {n} steps", flush=True)
    p = b.download("prim")
    r, phi = cell_centres(s, b)
    x, y = r * np.cos(phi), r * np.sin(phi)
    vy = p[2] * np.sin(phi) + p[3] * r * np.cos(phi)  # w_analytic_type 0: v_phi = r omega
    m = np.abs(y - 3.0) < 0.1
    nu, t0 = g.params["Explicit_Viscosity"], 0.25
    exact = 1. / np.sqrt(2 * np.pi * nu * (t0 + T)) * np.exp(-(x - 3.) ** 2 / (4. * nu * (t0 + T)))
    at_t0 = 1. / np.sqrt(2 * np.pi * nu * t0) * np.exp(-(x - 3.) ** 2 / (4. * nu * t0))
    peak = exact[m].max()
    err = np.max(np.abs(vy[m] - exact[m])) / peak
    print(f"slice: {m.sum()} cells, peak {vy[m].max():.4f} vs {peak:.4f}, max error {err:.4f} of the peak")
    assert m.sum() > 100
    assert err < 0.06  # measured 0.035 at NumR 128; the initial profile is off by 2.0
    assert np.max(np.abs(at_t0[m] - exact[m])) / peak > 1.0
    b.close()


def test_vortex_stays_steady():
    g, s, b = start("code_vortex")
    p0 = b.download("prim")
    r0, _ = cell_centres(s, b)
    T = g.params["Time_Max"]
    n = run_to_end(b, T)
    print(f"vortex.par: t = {T} after <= {n} steps", flush=True)
    p = b.download("prim")
    r, _ = cell_centres(s, b)
    assert np.array_equal(r, r0)
    vphi0, vphi = p0[3] * r, p[3] * r
    dv = np.max(np.abs(vphi - vphi0)) / np.max(vphi0)
    dP = np.max(np.abs(p[1] - p0[1]) / p0[1])
    drho = np.max(np.abs(p[0] - p0[0]))
    print(f"max |d v_phi| / max v_phi {dv:.2e}, max |dP|/P {dP:.2e}, max |d rho| {drho:.2e}")
    assert dv < 1e-3 and dP < 1e-3 and drho < 1e-3  # measured 2.9e-5, 9.0e-5, 5.3e-5
    assert p[5].min() >= 0.0 and p[5].max() <= 1.0 + 1e-12  # the passive marker stays bounded
    b.close()
