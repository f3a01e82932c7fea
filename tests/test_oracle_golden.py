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


def io_names():
    return [n for n in golden_names() if Golden(n).has("io/buffer")]


@pytest.mark.parametrize("name", io_names())
def test_checkpoint_rows_are_the_reference_buffer(name):
    """io_setbuf's rows (io_buffer.c:26-55) hold, per cell in cell order,
    tiph, the annulus and layer centres and the primitives of the state, and
    GravMassBuffer the masses' r, phi, M and omega"""
    from ckpt_util import centres
    g = Golden(name)
    buf = g.get("io/buffer").reshape(g.Nz, g.NQ + 3)
    assert bitwise_equal(buf[:, 0], g.get("io/state/tiph"))
    r, z = centres(g.N_p, g.r_faces, g.z_faces)
    assert bitwise_equal(buf[:, 1], r) and bitwise_equal(buf[:, 2], z)
    assert bitwise_equal(buf[:, 3:].T, g.get("io/state/prim"))
    gm = g.get("io/state/gravmass").reshape(2, 12)[:, :4]
    assert bitwise_equal(g.get("io/gravmass_buffer"), gm.ravel())


@pytest.mark.parametrize("name", io_names())
def test_restart_rules_match_reference(name):
    """the restart the checkpoint tests run on the oracle (tests/ckpt_util.py)
    is the reference's: io_readbuf into fresh cells and masses, then the
    set-up and RK2 steps, bit for bit"""
    from ckpt_util import restart_oracle
    g = Golden(name)
    buf = g.get("io/buffer").reshape(g.Nz, g.NQ + 3)
    T = g.get("io/state/t_dt")[0]
    o = restart_oracle(g.params, g.N_p, g.r_faces, g.z_faces, g.dphi0, buf, T,
                       g.get("io/gravmass_buffer").reshape(2, 4))
    for pfx in ("io/restart0", "io/restart"):
        if pfx == "io/restart":
            o.step(int(g.get("io/steps")[1]))
        for f in ("prim", "tiph", "dphi"):
            assert bitwise_equal(o.get(f), g.get(f"{pfx}/{f}")), f"{pfx} {f}: {describe(o.get(f), g.get(f'{pfx}/{f}'))}"
        sc = o.scalars()
        nt = 1 if pfx == "io/restart0" else 2  # (no dt before the first step)
        assert bitwise_equal(sc[:nt], g.get(f"{pfx}/t_dt")[:nt]), pfx
        gm = g.get(f"{pfx}/gravmass").reshape(2, 12)
        for m in range(2):
            assert bitwise_equal(sc[3 + 10 * m: 3 + 10 * m + 4], gm[m, :4]), f"{pfx} mass {m}"


@pytest.mark.parametrize("name", io_names())
def test_io_setup_schedule_matches_reference(name):
    """io_setup (io_setup.c:10-19) through the C ABI, host only: the next
    checkpoint time and the file number at the checkpointed time and after
    the restart"""
    from disco2_amd._lib import lib
    from disco2_amd.api import Setup
    g = Golden(name)
    L = lib()
    s = Setup(g.params)
    ts = L.timestep_create(s.sim)
    io = L.io_create(s.sim)
    L.timestep_set_t(ts, float(g.get("io/state/t_dt")[0]))
    L.io_setup(io, s.sim, ts)
    assert L.io_tcheck(io) == g.get("io/tcheck")[0]
    assert L.io_nfile(io) == g.get("io/nfile")[0]
    assert L.io_tcheck(io) == g.get("io/restart_tcheck")[0] and L.io_nfile(io) == g.get("io/restart_nfile")[0]
    L.io_destroy(io)
    L.timestep_destroy(ts)
