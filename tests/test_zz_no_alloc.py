"""The allocation-free torque loss (disco_batch_torque_loss_noalloc) against
disco_batch_torque_loss through the C ABI: the same loss bit for bit, on two
calls in a row, the second on a smaller disk (fewer update tiles than the
first left in the shared scratch), each from the same initial state as the
allocating call it is compared with."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _batch(numr, r_max):
    from disco2_amd import configs, ics
    from disco2_amd.api import Batch, Setup
    p, kind, kw = configs.binary(numr, r_min=0.5, r_max=r_max, q=0.5, alpha=0.01)
    s = Setup(p)
    b = Batch([s])
    b.upload("prim", ics.make_ic(kind, s.geometry(), p, **kw))
    b.init()
    return b


def _loss(numr, r_max, steps, noalloc):
    from disco2_amd._lib import check, dptr, lib
    b = _batch(numr, r_max)
    try:
        out = np.zeros(1)
        fn = lib().disco_batch_torque_loss_noalloc if noalloc else lib().disco_batch_torque_loss
        rc = fn(b.h, steps, dptr(out))
        check()
        assert rc == 0
        return float(out[0]), b.download("prim")
    finally:
        b.close()


def test_noalloc_matches_allocating_twice():
    cases = [(48, 6.0, 6), (16, 2.0, 4)]  # the second is smaller
    for numr, r_max, steps in cases:
        print(f"no-alloc torque loss: NumR {numr}, r_max {r_max}, {steps} steps", flush=True)
        ref, prim_ref = _loss(numr, r_max, steps, noalloc=False)
        got, prim_got = _loss(numr, r_max, steps, noalloc=True)
        assert np.isfinite(ref) and ref != 0.0
        assert got == ref, (numr, got, ref)
        assert np.array_equal(prim_got, prim_ref)
