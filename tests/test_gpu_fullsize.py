"""The measured workloads at their full sizes.

* binary2d_r512 (881,230 zones): ten RK2 steps against the oracle on the
  same seeded state (the oracle is the reference's arithmetic bit for bit;
  the device's sin/cos/pow differ from glibc's by ulps), per zone to 1e-11;
  t and dt to 1e-13;
* the same workload run twice gives identical bits (no atomics touch the
  state), and inside a batch with another disk it gives the bits it gives
  alone;
* disk3d_r256_z64 (13.4M zones with ghosts, 64 layers, HLL, periodic z):
  two RK2 steps against the oracle (about 35 s per step on one core), per
  zone to 1e-11, the cell edges to 1e-13 rad, t and dt to 1e-13; a run of
  three steps twice gives
  identical bits, and the disk beside a small 2D disk in a batch gives the
  bits it gives alone; the state stays finite with positive density and
  pressure;
* binary2d_r256_adjoint (221,930 interior zones, 200 steps, checkpoints every
  50): dL/d alpha and dL/d q against central differences of the GPU loss.
  Over 200 steps the loss is piecewise smooth and strongly curved in q (the
  masses' positions move which cells set the CFL minimum and where the
  limiters switch): the central difference converges to the adjoint only as
  the step shrinks (relative difference 1.4e-4 at h = 1e-6, 9.8e-6 at 1e-7,
  3.1e-6 at 1e-8, 1.1e-6 at 3e-9, the one-sided differences closing in
  linearly in h), so q is differenced with h = 3e-9 and alpha with 1e-7
  (1.5e-7);
* the same workload, its seeded state untied (P, v_r and omega perturbed by
  1e-3 relative, as the per-zone checks of test_gpu_adjoint.py), at the
  benchmark's checkpoint interval (every 20 steps): dL/d q and dL/d alpha
  against one-sided and central differences at h = 1e-6 ... 1e-8.  At
  h = 1e-6 the central difference agrees to 1e-5 (alpha: 1.4e-6) unless a
  kink of the loss falls inside the step (the CFL minimum and the limiters
  switch): then the two one-sided differences disagree by more than 1e-3
  while one of them is within 1e-3 of the adjoint (q: 1.2e-4 and 1.3e-2, the
  central difference 6.3e-3).  Some step of the sweep must agree to 1e-6
  (q: the minus side at 1e-7, 3.9e-7; alpha: 7.8e-8 at 3e-7);
* dL/d(initial primitives) of 48 zones spread over the full-size disk (six
  annuli from the inner edge to the outer, two cells each, rho, P, v_r,
  omega) against central differences at h = 1e-6, 1e-7, 1e-8 relative:
  each to 1e-4 of |fd| + 1e-3 max|g| at one of them (a kink can fall inside
  the wider steps: annulus 1's omega misses by 30% at h = 1e-6, where its
  minus-side difference jumps, and agrees to 1e-7 at h = 1e-7).
"""
import numpy as np
import pytest

from compare import bitwise_equal, describe, zone_rel

pytestmark = pytest.mark.gpu


def start(spec, extra=()):
    from disco2_amd import ics
    from disco2_amd.api import Batch, Setup
    p, kind, kw = spec
    s = Setup(p)
    prim0 = ics.make_ic(kind, s.geometry(), p, **kw)
    setups, prims = [s], [prim0]
    for q, k2, kw2 in extra:
        s2 = Setup(q)
        setups.append(s2)
        prims.append(ics.make_ic(k2, s2.geometry(), q, **kw2))
    b = Batch(setups)
    for d, pr in enumerate(prims):
        b.upload("prim", pr, disk=d)
    b.init()
    return p, s, prim0, b


def test_binary_full_size_matches_oracle():
    from disco2_amd import configs
    from oracle.oracle import Oracle
    p, s, prim0, b = start(configs.c2_binary())
    steps = 10
    print(f"binary2d_r512: {s.nzones} zones, {steps} steps against the oracle", flush=True)
    b.run(steps)
    o = Oracle(p, s.N_p, s.r_faces, s.z_faces)
    o.set_state(tiph=s.tiph0, prim=prim0)
    o.init()
    o.step(steps)
    rel = zone_rel(b.download("prim"), o.get("prim"))
    print(f"max per-zone relative difference {rel:.2e}")
    assert rel < 1e-11
    assert zone_rel(b.download("tiph")[None], o.get("tiph")[None]) < 1e-13
    sg, so = b.scalars(), o.scalars()
    assert abs(sg[0] - so[0]) <= 1e-13 * so[0] and abs(sg[1] - so[1]) <= 1e-13 * so[1]
    b.close()


def test_binary_full_size_is_deterministic_and_batch_invariant():
    from disco2_amd import configs
    steps = 20
    print(f"binary2d_r512: {steps} steps twice, then beside a vortex disk", flush=True)
    out = []
    for extra in ((), (configs.vortex(32),)):
        for _ in range(2 if not extra else 1):
            _, _, _, b = start(configs.c2_binary(), extra)
            b.run(steps)
            out.append((b.download("prim"), b.download("tiph"), b.scalars()))
            b.close()
    for k in (1, 2):
        assert bitwise_equal(out[k][0], out[0][0]), describe(out[k][0], out[0][0])
        assert bitwise_equal(out[k][1], out[0][1])
        assert bitwise_equal(out[k][2], out[0][2])


def test_disk3d_full_size_matches_oracle():
    from disco2_amd import configs
    from oracle.oracle import Oracle
    p, s, prim0, b = start(configs.c3_disk3d())
    steps = 2
    print(f"disk3d_r256_z64: {s.nzones} zones, {steps} steps against the oracle", flush=True)
    b.run(steps)
    prim, tiph, sg = b.download("prim"), b.download("tiph"), b.scalars()
    b.close()
    o = Oracle(p, s.N_p, s.r_faces, s.z_faces)
    o.set_state(tiph=s.tiph0, prim=prim0)
    o.init()
    o.step(steps)
    rel = zone_rel(prim, o.get("prim"))
    # the cell edges are angles of order one, some of them near zero: their
    # difference is measured in radians
    dt_abs = float(np.max(np.abs(tiph - o.get("tiph"))))
    print(f"max per-zone relative difference {rel:.2e}, cell edges {dt_abs:.2e} rad")
    assert rel < 1e-11
    assert dt_abs < 1e-13
    so = o.scalars()
    assert abs(sg[0] - so[0]) <= 1e-13 * so[0] and abs(sg[1] - so[1]) <= 1e-13 * so[1]


def test_disk3d_full_size_is_deterministic_and_batch_invariant():
    from disco2_amd import configs
    steps = 3
    out = []
    for extra in ((), (), (configs.vortex(32),)):
        _, s, _, b = start(configs.c3_disk3d(), extra)
        if not out:
            print(f"disk3d_r256_z64: {s.nzones} zones, {steps} steps twice, then beside a vortex disk", flush=True)
        b.run(steps)
        out.append((b.download("prim"), b.download("tiph"), b.scalars()))
        b.close()
    prim = out[0][0]
    assert np.all(np.isfinite(prim)) and prim[0].min() > 0.0 and prim[1].min() > 0.0
    assert out[0][2][0] > 0.0  # time advanced
    for k in (1, 2):
        assert bitwise_equal(out[k][0], prim), describe(out[k][0], prim)
        assert bitwise_equal(out[k][1], out[0][1])
        assert bitwise_equal(out[k][2], out[0][2])


def test_adjoint_full_size_matches_differences():
    from disco2_amd import configs, ics
    from disco2_amd.api import Batch, Setup
    p, kind, kw = configs.c5_adjoint()
    steps, every, ha, hq = 200, 50, 1e-7, 3e-9
    s = Setup(p)
    prim0 = ics.make_ic(kind, s.geometry(), p, **kw)

    def loss(**over):
        q = dict(p, **over)
        b = Batch([Setup(q)])
        b.upload("prim", prim0)
        b.init()
        try:
            return b.torque_loss(steps)
        finally:
            b.close()

    b = Batch([s])
    b.upload("prim", prim0)
    b.init()
    print(f"binary2d_r256_adjoint: {steps} steps, checkpoints every {every}", flush=True)
    L, ga, gq, _ = b.adjoint(steps, checkpoint_every=every)
    b.close()
    a0, q0 = p["Explicit_Viscosity"], p["MassRatio"]
    fa = (loss(Explicit_Viscosity=a0 * (1 + ha)) - loss(Explicit_Viscosity=a0 * (1 - ha))) / (2 * ha * a0)
    fq = (loss(MassRatio=q0 * (1 + hq)) - loss(MassRatio=q0 * (1 - hq))) / (2 * hq * q0)
    print(f"L {L:.12e}; dL/dalpha {ga:.10e} vs {fa:.10e}; dL/dq {gq:.10e} vs {fq:.10e}")
    assert abs(L - loss()) <= 1e-12 * abs(L)
    assert abs(ga - fa) <= 1e-6 * abs(fa)
    assert abs(gq - fq) <= 5e-6 * abs(fq)


@pytest.fixture(scope="module")
def untied_adjoint():
    from disco2_amd import configs, ics
    from disco2_amd.api import Batch, Setup
    p, kind, kw = configs.c5_adjoint()
    s = Setup(p)
    pr = ics.make_ic(kind, s.geometry(), p, **kw)
    rs = np.random.RandomState(7)
    n = pr.shape[1]
    cs = np.sqrt(pr[1] / pr[0])
    pr[1] *= 1 + 1e-3 * rs.uniform(-1, 1, n)
    pr[2] += 1e-3 * cs * rs.uniform(-1, 1, n)
    pr[3] *= 1 + 1e-3 * rs.uniform(-1, 1, n)
    steps, every = 200, 20
    print(f"binary2d_r256_adjoint untied: {steps} steps, checkpoints every {every}", flush=True)
    b = Batch([s])
    b.upload("prim", pr)
    b.init()
    L, ga, gq, gp = b.adjoint(steps, checkpoint_every=every)
    b.close()

    def loss(prim=None, **over):
        bb = Batch([Setup(dict(p, **over))])
        bb.upload("prim", pr if prim is None else prim)
        bb.init()
        try:
            return bb.torque_loss(steps)
        finally:
            bb.close()

    return dict(p=p, s=s, prim0=pr, L=L, ga=ga, gq=gq, gp=gp, loss=loss)


@pytest.mark.parametrize("name", ["q", "alpha"])
def test_adjoint_full_size_parameters_untied(untied_adjoint, name):
    u = untied_adjoint
    key = "MassRatio" if name == "q" else "Explicit_Viscosity"
    x0, g = u["p"][key], (u["gq"] if name == "q" else u["ga"])
    L0 = u["loss"]()
    assert abs(u["L"] - L0) <= 1e-12 * abs(L0)
    rows = {}
    for h in (1e-6, 3e-7, 1e-7, 3e-8, 1e-8):
        Lp, Lm = u["loss"](**{key: x0 * (1 + h)}), u["loss"](**{key: x0 * (1 - h)})
        fp, fm, fc = (Lp - L0) / (h * x0), (L0 - Lm) / (h * x0), (Lp - Lm) / (2 * h * x0)
        rows[h] = [abs(v - g) / abs(g) for v in (fc, fp, fm)]
        print(f"dL/d{name} h {h:.0e}: adjoint {g:.10e} central {fc:.10e} one-sided + {fp:.10e} - {fm:.10e}; "
              f"relative differences {rows[h][0]:.1e} {rows[h][1]:.1e} {rows[h][2]:.1e}")
    c6, p6, m6 = rows[1e-6]
    if c6 > 1e-5:  # a kink inside [x0 - h, x0 + h]: the two sides disagree, one is near the adjoint
        assert abs(p6 - m6) > 1e-3 and min(p6, m6) < 1e-3, rows[1e-6]
    assert min(min(r) for r in rows.values()) <= 1e-6, rows


def test_adjoint_full_size_initial_state_untied(untied_adjoint):
    u = untied_adjoint
    s, prim0, gp = u["s"], u["prim0"], u["gp"]
    aoff = np.concatenate([[0], np.cumsum(s.N_p)])
    gmax = np.max(np.abs(gp))
    worst, n = 0.0, 0
    for i in (1, 2, s.NR // 4, s.NR // 2, 3 * s.NR // 4, s.NR - 3):
        for j in (3, int(s.N_p[i]) // 2):
            z = int(aoff[i] + j)
            for q in (0, 1, 2, 3):  # rho, P, v_r, omega
                x0 = prim0[q, z]
                errs = []
                for hr in (1e-6, 1e-7, 1e-8):
                    h = hr * max(abs(x0), 1e-4)
                    a, c = prim0.copy(), prim0.copy()
                    a[q, z] = x0 + h
                    c[q, z] = x0 - h
                    fd = (u["loss"](a) - u["loss"](c)) / (2 * h)
                    errs.append(abs(gp[q, z] - fd) / (abs(fd) + 1e-3 * gmax))
                worst = max(worst, min(errs))
                n += 1
                assert min(errs) <= 1e-4, f"zone {z} (annulus {i}) q {q}: adjoint {gp[q, z]:.6e}, errors {errs}"
    print(f"{n} initial-state derivatives at full size, worst {worst:.2e} of |fd| + 1e-3 max|g|")
