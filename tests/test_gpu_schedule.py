"""Launch schedules and the autotuner (disco_batch_autotune).

A schedule changes only which workgroup runs which tile (XCD bands or launch
order) and in which order the face builder takes the annulus pairs (table
order or largest first), whether it stages the outer annulus in LDS, and
whether the update evaluates the r / z faces' fluxes itself (each face by
both of its cells) instead of reading them stored per face, and how many
threads the face builder runs per annulus pair; never the arithmetic or its
order.  So every schedule must give the same
bits, and the autotuner, which times the candidates on a scratch copy of the
batch, must leave the batch's state as it was.
"""
import numpy as np
import pytest

from compare import bitwise_equal, describe

pytestmark = pytest.mark.gpu


def batch(with3d=True):
    from disco2_amd import configs, ics
    from disco2_amd.api import Batch, Setup
    specs = [configs.binary(48, r_min=0.4, r_max=4.0, q=0.3), configs.vortex(40)]
    if with3d:
        specs.append(configs.disk3d(10, 6, 40, R_Min=0.5, R_Max=2.5))
    setups = [Setup(p) for p, _, _ in specs]
    b = Batch(setups)
    for d, ((p, kind, kw), s) in enumerate(zip(specs, setups)):
        b.upload("prim", ics.make_ic(kind, s.geometry(), p, **kw), disk=d)
    b.init()
    return b, len(specs)


def state(b, n):
    return [(b.download("prim", disk=d), b.download("tiph", disk=d), b.scalars(d)) for d in range(n)]


@pytest.mark.parametrize("with3d,fused", [(True, 2), (False, 1)])
def test_every_schedule_gives_the_same_bits(with3d, fused):
    ref = None
    extra = [5 | 1 << 5, 6 | 1 << 5 | fused << 3] if with3d else []
    widths = [5 | 1 << 8, 4 | 2 << 8, 7 | 2 << 8]  # face builder at 512 / 256 threads
    for k in list(range(8)) + [5 | fused << 3, 2 | fused << 3] + extra + widths:
        print(f"schedule {k} (3D disk: {with3d})", flush=True)
        b, n = batch(with3d)
        b.set_schedule(k)
        assert b.schedule() == k
        b.run(6)
        st = state(b, n)
        b.close()
        if ref is None:
            ref = st
            continue
        for d in range(n):
            for a, r in zip(st[d], ref[d]):
                assert bitwise_equal(a, r), f"schedule {k} disk {d}: {describe(a, r)}"


def test_autotune_keeps_the_batch_state_and_picks_the_fastest():
    b, n = batch()
    b.run(2)
    before = state(b, n)
    k, ms = b.autotune(4)
    print(f"chosen {k}: {ms}", flush=True)
    assert k in ms and ms[k] == min(ms.values())
    assert b.schedule() == k
    after = state(b, n)
    for d in range(n):
        for a, r in zip(after[d], before[d]):
            assert bitwise_equal(a, r)
    b.run(3)  # the batch runs on under the chosen schedule
    b.close()
