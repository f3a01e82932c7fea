"""The diagnostics of the CPU oracle (oracle/disco_oracle.c oracle_diag_measure
and oracle.oracle.Diagnostics) against the reference's own diagnostics
(src/Diagnostics, compiled into oracle/_ref; tests/golden/diag_*.npz).

At every measurement the fixture holds the reference's state and its
diagnostics' arrays at full precision: the oracle, given that state, must
reproduce the time-averaged radial profiles (VectorDiag), the scalars
(ScalarDiag) and the equatorial cells (EquatDiag) bit for bit, the
measure/dump schedule exactly, and the text files the reference wrote
(DiagMdot.dat, DiagScalar.dat, DiagVector_*.dat) character for character.
Exception, in the reference itself: with NUM_Q = 5 its passive-scalar column
(EquatDiag 9, VectorDiag 8, ScalarDiag 7) reads past the 5-entry primitive
array; those columns are not compared there (the restatement writes 0).
"""
import os
import tempfile

import numpy as np
import pytest

from golden_util import Golden, golden_names

DIAG = [n for n in golden_names() if n.startswith("diag_")]


def state_at(g, o, m):
    pre = f"diag/meas{m}/state"
    o.put("prim", g.get(pre + "/prim"))
    o.put("tiph", g.get(pre + "/tiph"))
    o.put("dphi", g.get(pre + "/dphi"))
    gm = g.get(pre + "/gravmass").reshape(2, 12)
    t, dt = g.get(pre + "/t_dt")
    sc = np.concatenate([[t, dt, 0.0], gm[0, :10], gm[1, :10]])
    o.set_scalars(sc)
    return t


def measures(g):
    n, _ = (int(x) for x in g.get("diag/counts"))
    return [(m, *g.get(f"diag/meas{m}/step_t")) for m in range(n)]


def dumps(g):
    _, n = (int(x) for x in g.get("diag/counts"))
    return [tuple(g.get(f"diag/dump{d}/step_t")) for d in range(n)]


def make(g):
    from oracle.oracle import Oracle
    o = Oracle(g.params, g.N_p, g.r_faces, g.z_faces)
    o.set_state(tiph=g.tiph0, prim=g.prim0)
    o.init()
    return o


@pytest.mark.parametrize("name", DIAG)
def test_measurements_bitwise(name):
    from oracle.oracle import Diagnostics
    g = Golden(name)
    o = make(g)
    d = Diagnostics(o, g.params, 0.0)
    mask_v = np.ones(17, bool)
    mask_e = np.ones(18, bool)
    mask_s = np.ones(16, bool)
    if g.NQ == 5:
        mask_v[8] = mask_e[9] = mask_s[7] = False
    dump_steps = {int(s) for s, _ in dumps(g)}
    assert int(g.get("diag/N_eq")[0]) == make(g).diag_measure(1.0, d.rz)[2].shape[0]
    with tempfile.TemporaryDirectory() as tmp:
        for m, step, t in measures(g):
            state_at(g, o, m)
            assert d.set(t, tmp), f"measurement {m} at step {int(step)} not scheduled"
            v = g.get(f"diag/meas{m}/vector").reshape(-1, 17)
            assert np.array_equal(d.vec[:, mask_v], v[:, mask_v]), f"meas {m} vector"
            assert np.array_equal(d.scal[mask_s], g.get(f"diag/meas{m}/scalar")[mask_s]), f"meas {m} scalar"
            e = g.get(f"diag/meas{m}/equat").reshape(-1, 18)
            assert np.array_equal(d.equat[:, mask_e], e[:, mask_e]), f"meas {m} equat"
            assert d.print(t, tmp) == (int(step) in dump_steps), f"dump schedule at step {int(step)}"
        files = {k[5:]: str(g.get(k)) for k in g.d if k.startswith("file/")}
        assert sorted(files) == sorted(os.listdir(tmp))
        for fn, ref in files.items():
            mine = open(os.path.join(tmp, fn)).read()
            if g.NQ > 5:
                assert mine == ref, fn
                continue
            # NUM_Q = 5: every column but the passive scalar's
            skip = {"DiagScalar.dat": 1 + 7, "DiagMdot.dat": None}.get(fn, 8)
            for lm, lr in zip(mine.splitlines(), ref.splitlines(), strict=True):
                a, b = lm.split(), lr.split()
                assert len(a) == len(b)
                for c, (x, y) in enumerate(zip(a, b)):
                    if c != skip:
                        assert x == y, f"{fn} column {c}: {x} != {y}"
