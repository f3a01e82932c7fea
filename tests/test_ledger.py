"""Conservation ledger of the oracle (the CPU restatement, bitwise the
reference): over 5 RK2 steps of outflow disks, per substep, the interior
mass and angular momentum change by exactly what the boundary r faces carry
plus the gravitational torque term (tests/ledger_util.py), to 1e-13 of the
total.  This pins the torque term (a dropped cell, a wrong sign, a wrong
substep weight breaks the angular-momentum identity) and hence the adjoint's
loss accumulator oracle_torque, which sums the same terms."""
import numpy as np
import pytest

from ledger_util import DDD, LLL, interior_mask, predicted

CASES = {
    "binary_inviscid": dict(alpha=-1.0),
    "binary_viscous": dict(alpha=0.01),
}
STEPS = 5


def setup(alpha):
    from disco2_amd import configs, ics
    from disco2_amd.api import Setup
    # the binary (separation 1) orbits inside the inner boundary: no floors
    p, kind, kw = configs.binary(24, r_min=1.5, r_max=4.0, q=0.5, alpha=alpha)
    s = Setup(p)
    return p, s, ics.make_ic(kind, s.geometry(), p, **kw)


@pytest.mark.parametrize("case", sorted(CASES))
def test_oracle_conserves_mass_and_angular_momentum(case):
    from oracle.oracle import Oracle
    p, s, prim0 = setup(**CASES[case])
    o = Oracle(p, s.N_p, s.r_faces, s.z_faces)
    o.set_state(tiph=s.tiph0, prim=prim0)
    o.init()
    mask = interior_mask(s.N_p)
    nq = s.NQ
    led_torque = 0.0
    for _ in range(STEPS):
        o.step_begin()
        for fac, RK in ((1.0, 0.0), (0.5, 0.5)):
            u0 = o.get("cons")[:, mask]
            rk = o.get("RKcons")[:, mask]
            o.substep(fac)
            led = o.ledger()
            u1 = o.get("cons")[:, mask]
            for q in (DDD, LLL):
                pred = predicted(u0[q].sum(), rk[q].sum(), RK, led, nq, q)
                scale = np.abs(u0[q]).sum()
                assert abs(u1[q].sum() - pred) <= 1e-13 * scale, (case, q, u1[q].sum(), pred)
            assert abs(led[2 * (nq + 3)]) > 0.0  # the binary exerts a torque
            led_torque += led[2 * (nq + 3)]
        o.step_end()
    # the adjoint's loss accumulator sums the same per-cell terms
    assert abs(o.torque() - led_torque) <= 1e-13 * abs(led_torque)


def test_ledger_catches_a_wrong_torque():
    """the identity fails for the angular momentum if the torque term is
    dropped (the check has teeth)"""
    from oracle.oracle import Oracle
    p, s, prim0 = setup(-1.0)
    o = Oracle(p, s.N_p, s.r_faces, s.z_faces)
    o.set_state(tiph=s.tiph0, prim=prim0)
    o.init()
    mask = interior_mask(s.N_p)
    o.step_begin()
    u0 = o.get("cons")[:, mask]
    rk = o.get("RKcons")[:, mask]
    o.substep(1.0)
    led = o.ledger()
    led[2 * (s.NQ + 3)] = 0.0
    u1 = o.get("cons")[:, mask]
    pred = predicted(u0[LLL].sum(), rk[LLL].sum(), 0.0, led, s.NQ, LLL)
    assert abs(u1[LLL].sum() - pred) > 1e-9 * np.abs(u0[LLL]).sum()
