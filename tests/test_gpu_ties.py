"""Tied edges: cell edges of neighbouring annuli (or layers) that coincide
exactly, the collisions of the face walk (build_jloop,
src/Face/face_create_destroy.c:63-128, compares edges strictly and creates
its faces by which side of a tie it lands on).

The face builder's per-cell walks start from a guessed outer cell; at a tie
the guess and the reference's strict comparisons can disagree, and the
builder then falls back to the reference's sequential walk for that pair.
None of the seeded states has ties (the reference draws every annulus's
offset at random), so these cases make them:

* vortex: every annulus starts its cells at the previous annulus's first
  edge, so each r pair has at least one tie in the first substep (the
  annuli then move apart at their own angular speeds);
* 3D disk: every layer takes layer 0's edges, so every edge of every z pair
  ties, in every substep (all layers of an annulus move alike).

Each runs a few RK2 steps on the GPU and on the oracle from the same state.
"""
import numpy as np
import pytest

from compare import bitwise_equal, describe, zone_rel

pytestmark = pytest.mark.gpu


def ties_between(t_in, t_out):
    """exactly coinciding edges (mod 2 pi) of two annuli's cells"""
    a = np.mod(np.asarray(t_in), 2 * np.pi)
    b = np.mod(np.asarray(t_out), 2 * np.pi)
    return len(np.intersect1d(a, b))


def run_both(params, kind, kw, tiph, steps):
    from disco2_amd import ics
    from disco2_amd.api import Batch, Setup
    from oracle.oracle import Oracle
    s = Setup(params)
    prim0 = ics.make_ic(kind, s.geometry(), params, **kw)
    b = Batch([s])
    b.upload("tiph", tiph)
    b.upload("prim", prim0)
    b.init()
    b.run(steps)
    gpu = {f: b.download(f) for f in ("prim", "cons", "tiph", "dphi")}
    b.close()
    o = Oracle(params, s.N_p, s.r_faces, s.z_faces)
    o.set_state(tiph=tiph, prim=prim0)
    o.init()
    o.step(steps)
    ref = {f: o.get(f) for f in ("prim", "cons", "tiph", "dphi")}
    return gpu, ref


def test_r_pair_ties_bitwise():
    from disco2_amd import configs
    from disco2_amd.api import Setup
    p, kind, kw = configs.vortex(24)
    s = Setup(p)
    N = np.asarray(s.N_p)
    off = np.concatenate([[0], np.cumsum(N)])
    t = np.array(s.tiph0, dtype=np.float64)
    t0 = t[0]
    for i in range(len(N)):  # the reference's tiph_0 + j dphi, with one tiph_0 for all
        dphi = 2 * np.pi / N[i]
        t[off[i]:off[i + 1]] = t0 + np.arange(N[i]) * dphi
    ties = [ties_between(t[off[i]:off[i + 1]], t[off[i + 1]:off[i + 2]]) for i in range(len(N) - 1)]
    assert min(ties) >= 1, ties
    print(f"vortex NumR 24: {sum(ties)} tied edges over {len(ties)} r pairs, 3 steps", flush=True)
    gpu, ref = run_both(p, kind, kw, t, 3)
    for f in ("prim", "cons", "tiph", "dphi"):
        assert bitwise_equal(gpu[f], ref[f]), f"{f}: {describe(gpu[f], ref[f])}"


def test_z_pair_ties():
    from disco2_amd import configs
    from disco2_amd.api import Setup
    p, kind, kw = configs.disk3d(16, 8, 48, R_Min=0.5, R_Max=3.0)
    s = Setup(p)
    NL = int(np.sum(s.N_p))
    t = np.array(s.tiph0, dtype=np.float64).reshape(s.NZ, NL)
    t[1:] = t[0]
    print(f"disk3d NumR 16 NumZ 8: every z-pair edge tied ({(s.NZ - 1) * NL} per substep), 5 steps", flush=True)
    gpu, ref = run_both(p, kind, kw, t.reshape(-1), 5)
    err = zone_rel(gpu["cons"], ref["cons"])
    print(f"max per-zone relative difference of cons: {err:.3e}")
    assert err < 1e-12
    # the Keplerian mesh speed goes through sqrt and pow, whose device and
    # glibc results differ by ulps: the angles agree to round-off
    assert np.max(np.abs(gpu["tiph"] - ref["tiph"])) < 1e-13
