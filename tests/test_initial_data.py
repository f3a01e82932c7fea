"""Initial-data plugins (src/Cell/initial_data/cell_init_*.c) on the host.

The plugin fixtures under tests/golden/ (plugin_*) hold the state the
reference's own plugin leaves before the first step (src/main.c:103-119 in
the reference build: gravMass init, the plugin, clean_pi, the z sync), for
all seven InitialDataTypes; their run/ records then cover the fixed r
boundaries and the boundary damping through whole RK2 steps
(test_oracle_golden.py on the CPU, test_gpu_parity.py on the GPU).

Here, bit for bit against those states:
* the oracle's restatement (oracle_initial_data) followed by its init;
* the product's host plugins (disco_initial_prim, compiled like the
  reference with g++): the raw plugin output, so the z ghost layers, which
  the reference's z sync overwrites, are left out in 3D.
And, where the fixtures say nothing, the two restatements against each
other: the vortex writes nothing at r >= 1, the torus's one draw per cell.
"""
import numpy as np
import pytest

from compare import bitwise_equal, describe
from golden_util import Golden, golden_names

PLUGINS = [n for n in golden_names() if n.startswith("plugin_")]


def interior_layers(g):
    """zone mask of the layers the z sync leaves alone"""
    m = np.ones(g.Nz, bool)
    if g.NZ > 1:
        ng = g.params["ng"]
        NL = int(g.N_p.sum())
        m[:ng * NL] = False
        m[(g.NZ - ng) * NL:] = False
    return m


def test_every_initial_data_type_has_a_fixture():
    types = {Golden(n).params["InitialDataType"] for n in PLUGINS}
    assert types == set(range(7))


@pytest.mark.parametrize("name", PLUGINS)
def test_oracle_plugin_matches_reference(name):
    from oracle.oracle import Oracle
    g = Golden(name)
    o = Oracle(g.params, g.N_p, g.r_faces, g.z_faces)
    o.set_state(tiph=g.tiph0, prim=np.zeros_like(g.prim0))
    o.initial_data()
    o.init()
    assert bitwise_equal(o.get("prim"), g.prim0), describe(o.get("prim"), g.prim0)


@pytest.mark.parametrize("name", PLUGINS)
def test_host_plugin_matches_reference(name):
    from disco2_amd.api import Setup
    g = Golden(name)
    s = Setup(g.params)
    assert np.array_equal(s.N_p, g.N_p) and bitwise_equal(s.tiph0, g.tiph0)
    p = s.initial_prim()
    m = interior_layers(g)
    assert bitwise_equal(p[:, m], g.prim0[:, m]), describe(p[:, m], g.prim0[:, m])


def test_vortex_writes_nothing_outside_its_core():
    """cell_init_vortex.c:173-187: annuli with r >= 1 keep what they held"""
    from disco2_amd import configs
    from disco2_amd.api import Setup
    from oracle.oracle import Oracle
    p, _, _ = configs.plugin("vortex", 12, R_Min=0.3, R_Max=1.6)
    s = Setup(p)
    before = np.full((s.NQ, s.nzones), 7.0)
    mine = s.initial_prim(before)
    o = Oracle(p, s.N_p, s.r_faces, s.z_faces)
    o.set_state(tiph=s.tiph0, prim=before)
    o.initial_data()
    assert bitwise_equal(mine, o.get("prim"))
    rc = 0.5 * (s.r_faces[:-1] + s.r_faces[1:])
    r = np.repeat(rc, s.N_p)
    assert np.all(mine[:, r >= 1.0] == 7.0)
    assert np.all(mine[0, r < 1.0] == 1.0)
    assert set(np.unique(mine[5, r < 1.0])) <= {0.0, 1.0}


@pytest.mark.parametrize("kind", ["shear", "sstest", "middle", "milos", "rad_dom", "torus"])
def test_host_and_oracle_plugins_agree_on_other_meshes(kind):
    """a mesh of another size and N_phi rule than the fixtures'"""
    from disco2_amd import configs
    from disco2_amd.api import Setup
    from oracle.oracle import Oracle
    p, _, _ = configs.plugin(kind, 20, N_p_round=0, NPCAP=40)
    s = Setup(p)
    mine = s.initial_prim()
    o = Oracle(p, s.N_p, s.r_faces, s.z_faces)
    o.set_state(tiph=s.tiph0, prim=np.zeros((s.NQ, s.nzones)))
    o.initial_data()
    assert bitwise_equal(mine, o.get("prim")), describe(mine, o.get("prim"))
    assert np.all(np.isfinite(mine))
