"""The batched workload at its full size: 64 independent binary disks (NumR
uniform in [64, 384] on r in [0.3, 10], N_phi = round(2 pi r / dr), q
log-uniform in [0.01, 1], alpha log-uniform in [1e-3, 1e-1], Mach uniform in
[10, 30], each with its own 1% seeded density noise; about 10.3M interior
zones) in one batch:

* after 2 RK2 steps, the two smallest disks agree with the oracle (the CPU
  restatement, bitwise the reference) to 1e-11 per zone;
* after 5 steps every disk is finite with positive density and pressure,
  and four sampled disks hold exactly the bits they reach when run alone."""
import numpy as np
import pytest

from compare import bitwise_equal, describe, zone_rel

pytestmark = pytest.mark.gpu

SAMPLED = (0, 21, 42, 63)


@pytest.fixture(scope="module")
def c4():
    from disco2_amd import configs, ics
    from disco2_amd.api import Setup
    specs = configs.c4_batch(64)
    setups = [Setup(p) for p, _, _ in specs]
    prims = [ics.make_ic(k, s.geometry(), p, **kw) for (p, k, kw), s in zip(specs, setups)]
    return specs, setups, prims


def test_c4_batch_full_size(c4):
    from disco2_amd.api import Batch
    from oracle.oracle import Oracle
    specs, setups, prims = c4
    n_int = sum(s.interior_zones() for s in setups)
    print(f"64 disks, {n_int} interior zones", flush=True)
    assert n_int > 10_000_000
    big = Batch(setups)
    for d, pr in enumerate(prims):
        big.upload("prim", pr, disk=d)
    big.init()
    big.run(2)
    small = sorted(range(64), key=lambda d: setups[d].nzones)[:2]
    at2 = {d: big.download("cons", disk=d) for d in small}
    print("2 steps done; oracle on disks", small, flush=True)
    for d in small:
        (p, _, _), s = specs[d], setups[d]
        o = Oracle(p, s.N_p, s.r_faces, s.z_faces)
        o.set_state(tiph=s.tiph0, prim=prims[d])
        o.init()
        o.step(2)
        err = zone_rel(at2[d], o.get("cons"))
        print(f"disk {d} (NumR {p['NumR']}): max per-zone relative difference {err:.3e}", flush=True)
        assert err < 1e-11
    big.run(3)
    for d in range(64):
        pr = big.download("prim", disk=d)
        assert np.all(np.isfinite(pr)), f"disk {d}"
        assert np.all(pr[0] > 0.0) and np.all(pr[1] > 0.0), f"disk {d}"
    got = {d: (big.download("prim", disk=d), big.download("tiph", disk=d), big.scalars(d)) for d in SAMPLED}
    big.close()
    for d in SAMPLED:
        print(f"disk {d} alone", flush=True)
        b = Batch([setups[d]])
        b.upload("prim", prims[d])
        b.init()
        b.run(5)
        for a, r in zip(got[d], (b.download("prim"), b.download("tiph"), b.scalars())):
            assert bitwise_equal(a, r), f"disk {d}: {describe(a, r)}"
        b.close()
