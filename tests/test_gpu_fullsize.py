"""The measured workloads at their full sizes.

* binary2d_r512 (881,230 zones): two RK2 steps against the oracle on the
  same seeded state (the oracle is the reference's arithmetic bit for bit;
  the device's sin/cos/pow differ from glibc's by ulps), per zone to 1e-11;
  t and dt to 1e-13;
* the same workload run twice gives identical bits (no atomics touch the
  state), and inside a batch with another disk it gives the bits it gives
  alone;
* disk3d_r256_z64 (12.5M zones, 64 layers, HLL, periodic z): the oracle is
  too slow at this size (about a minute per step on one core), so the test
  checks what does not depend on it: a run twice gives identical bits, and
  the disk beside a small 2D disk in a batch gives the bits it gives alone;
  the state stays finite with positive density and pressure;
* binary2d_r256_adjoint (221,930 interior zones, 200 steps, checkpoints every
  50): dL/d alpha and dL/d q against central differences of the GPU loss.
  Over 200 steps the loss is piecewise smooth and strongly curved in q (the
  masses' positions move which cells set the CFL minimum and where the
  limiters switch): the central difference converges to the adjoint only as
  the step shrinks (relative difference 1.4e-4 at h = 1e-6, 9.8e-6 at 1e-7,
  3.1e-6 at 1e-8, 1.1e-6 at 3e-9, the one-sided differences closing in
  linearly in h), so q is differenced with h = 3e-9 and alpha with 1e-7
  (1.5e-7).
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
    print(f"binary2d_r512: {s.nzones} zones, 2 steps against the oracle", flush=True)
    b.run(2)
    o = Oracle(p, s.N_p, s.r_faces, s.z_faces)
    o.set_state(tiph=s.tiph0, prim=prim0)
    o.init()
    o.step(2)
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
