"""Initial-data plugins, fixed r boundaries and boundary damping on the GPU,
against the reference's plugin fixtures (tests/golden/plugin_*, written by
the reference build running its own plugins, fixed boundaries and damping:
tests/golden/make_golden.py).

* Batch.initial_data() + init(): the initial state bit for bit (the plugins
  run on the host with the reference's C library), then whole RK2 steps:
  the fixed-boundary ghosts take the single-cell initial data at their new
  edges and the damped annuli relax towards it on the device
  (k_bc_init).  Bit for bit for the vortex (no transcendental enters: its
  boundary values are tabulated on the host, the passive marker is a sign);
  else within the per-zone tolerance of the other transcendental cases
  (device and glibc sin/cos/exp differ by ulps).
* several plugin disks of different types and boundaries in one batch;
* the class API: the C driver runs cell_init_ptr as src/main.c:103-104;
* the combinations the reference cannot run are rejected with an error.
"""
import subprocess

import numpy as np
import pytest

from compare import bitwise_equal, describe, zone_rel
from golden_util import Golden, golden_names

pytestmark = pytest.mark.gpu

PLUGINS = [n for n in golden_names() if n.startswith("plugin_")]
BITWISE = {"plugin_vortex_fixed"}
RTOL = 1e-11


def check_step(name, g, st, prim):
    ref = g.get(f"run/step{st}/prim")
    if name in BITWISE:
        assert bitwise_equal(prim, ref), f"{name} step {st}: {describe(prim, ref)}"
    else:
        assert zone_rel(prim, ref) < RTOL, f"{name} step {st}: rel {zone_rel(prim, ref)}"


@pytest.mark.parametrize("name", PLUGINS)
def test_batch_initial_data_and_steps(name):
    from disco2_amd.api import Batch, Setup
    print(f"plugin case {name}", flush=True)
    g = Golden(name)
    b = Batch([Setup(g.params)])
    b.initial_data()
    b.init()
    p0 = b.download("prim")
    assert bitwise_equal(p0, g.prim0), describe(p0, g.prim0)
    last = 0
    for st in g.steps():
        b.run(st - last)
        last = st
        check_step(name, g, st, b.download("prim"))
    b.close()


def test_plugin_disks_of_several_kinds_in_one_batch():
    from disco2_amd.api import Batch, Setup
    names = ["plugin_sstest_fixed_damp", "plugin_rad_dom_fixed", "plugin_middle_outflow_damp", "plugin_shear_fixed"]
    gs = [Golden(n) for n in names]
    print("one batch:", names, flush=True)
    b = Batch([Setup(g.params) for g in gs])
    for d, g in enumerate(gs):
        b.upload("prim", g.prim0, disk=d)
    b.init()
    steps = gs[0].steps()
    assert all(g.steps() == steps for g in gs)
    last = 0
    for st in steps:
        b.run(st - last)
        last = st
        for d, (n, g) in enumerate(zip(names, gs)):
            check_step(n, g, st, b.download("prim", disk=d))
    b.close()


def test_c_driver_runs_the_plugin(tmp_path):
    from disco2_amd.params import write_parfile
    from test_c_driver import build_driver
    name = "plugin_shear_fixed"
    g = Golden(name)
    write_parfile(str(tmp_path / "run.par"), g.params)
    exe = build_driver(str(tmp_path))
    st = g.steps()[-1]
    print(f"C driver: {name}, {st} steps from cell_init_ptr", flush=True)
    r = subprocess.run([exe, "run.par", "plugin", str(st), "out.bin"], capture_output=True, text=True, timeout=120,
                       cwd=str(tmp_path))
    assert r.returncode == 0, r.stderr
    out = np.fromfile(str(tmp_path / "out.bin"))
    prim = out[:g.NQ * g.Nz].reshape(g.NQ, g.Nz)
    check_step(name, g, st, prim)


def rejected(params):
    from disco2_amd._lib import DiscoError
    from disco2_amd.api import Batch, Setup
    with pytest.raises(DiscoError) as e:
        Batch([Setup(params)])
    return str(e.value)


def test_combinations_the_reference_cannot_run_are_rejected():
    from disco2_amd import configs
    p, _, _ = configs.plugin("torus", 10, BoundTypeR=0)
    assert "torus" in rejected(p)
    p, _, _ = configs.plugin("torus", 10, BoundTypeZ=0)
    assert "torus" in rejected(p)
    p, _, _ = configs.plugin("shear", 12, DAMP_TIME=1.0, RDAMP_INNER=2.5, RDAMP_OUTER=3.5)
    assert "shear" in rejected(p)
    p, _, _ = configs.plugin("sstest", 16, NUM_N=1)
    assert "passive" in rejected(p)
    p, _, _ = configs.plugin("vortex", 12, NUM_N=0, R_Max=1.5, DAMP_TIME=1.0, RDAMP_INNER=0.3, RDAMP_OUTER=0.9)
    assert "vortex" in rejected(p)
    p, _, _ = configs.plugin("sstest", 16, InitialDataType=9)
    assert "initial data" in rejected(p)


def test_adjoint_rejects_damping():
    from disco2_amd import configs
    from disco2_amd._lib import DiscoError
    from disco2_amd.api import Batch, Setup
    p, _, _ = configs.plugin("middle", 16, w_analytic_type=0)
    b = Batch([Setup(p)])
    b.initial_data()
    b.init()
    with pytest.raises(DiscoError) as e:
        b.adjoint(2, 2)
    assert "damping" in str(e.value)
    b.close()
