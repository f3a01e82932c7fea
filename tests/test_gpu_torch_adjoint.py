"""The torch.autograd front end of the adjoint: its gradients are the C ABI's,
scaled through a composed loss, and agree with central differences of the
loss (torch.autograd.gradcheck on alpha and q, and a 1e-6 relative check).
Binary disk NumR 24 on r in [0.5, 3], q = 0.5, alpha = 0.01, 6 RK2 steps."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

STEPS = 6


def problem():
    from disco2_amd import configs, ics
    from disco2_amd.api import Setup
    p, kind, kw = configs.binary(24, r_min=0.5, r_max=3.0, q=0.5, alpha=0.01)
    s = Setup(p)
    prim0 = ics.make_ic(kind, s.geometry(), p, **kw)
    # away from the min/minmod kinks of the smooth disk (v_r = 0, omega and
    # the sound speed independent of phi): perturb P, v_r and omega by 1e-3
    rs = np.random.RandomState(3)
    n = prim0.shape[1]
    cs = np.sqrt(prim0[1] / prim0[0])
    prim0[1] *= 1 + 1e-3 * rs.uniform(-1, 1, n)
    prim0[2] += 1e-3 * cs * rs.uniform(-1, 1, n)
    prim0[3] *= 1 + 1e-3 * rs.uniform(-1, 1, n)
    s.close()
    return p, prim0


def test_gradients_are_the_c_abi_ones():
    import torch
    from disco2_amd.api import Batch, Setup
    from disco2_amd.torch_adjoint import torque_loss
    p, prim0 = problem()
    alpha = torch.tensor(0.01, dtype=torch.float64, requires_grad=True)
    q = torch.tensor(0.5, dtype=torch.float64, requires_grad=True)
    pr = torch.tensor(prim0, requires_grad=True)
    loss = torque_loss(alpha, q, pr, p, STEPS, checkpoint_every=4)
    (3.0 * loss).backward()
    b = Batch([Setup(p)])
    b.upload("prim", prim0)
    b.init()
    l0, ga, gq, gp = b.adjoint(STEPS, checkpoint_every=4)
    b.close()
    print(f"loss {loss.item():.15e}  dL/dalpha {alpha.grad.item():.10e}  dL/dq {q.grad.item():.10e}")
    # the outflow remap's backward adds with atomics: runs agree to rounding
    assert loss.item() == l0
    assert abs(alpha.grad.item() - 3.0 * ga) <= 1e-12 * abs(3.0 * ga)
    assert abs(q.grad.item() - 3.0 * gq) <= 1e-12 * abs(3.0 * gq)
    assert np.max(np.abs(pr.grad.numpy() - 3.0 * gp)) <= 1e-12 * np.max(np.abs(3.0 * gp))


def test_gradcheck_alpha_q():
    import torch
    from disco2_amd.torch_adjoint import torque_loss
    p, prim0 = problem()
    pr = torch.tensor(prim0)

    def f(a, q):
        return torque_loss(a, q, pr, p, STEPS)

    alpha = torch.tensor(0.01, dtype=torch.float64, requires_grad=True)
    q = torch.tensor(0.5, dtype=torch.float64, requires_grad=True)
    assert torch.autograd.gradcheck(f, (alpha, q), eps=1e-7, atol=0.0, rtol=1e-5)
    # and to 1e-6 relative with a step of 1e-6 relative
    f(alpha, q).backward()
    for x, x0, g in ((0, 0.01, alpha.grad.item()), (1, 0.5, q.grad.item())):
        h = 1e-6 * x0
        args = lambda v: (torch.tensor(v if x == 0 else 0.01, dtype=torch.float64),
                          torch.tensor(v if x == 1 else 0.5, dtype=torch.float64))
        with torch.no_grad():
            fd = (f(*args(x0 + h)).item() - f(*args(x0 - h)).item()) / (2 * h)
        print(f"input {x}: adjoint {g:.10e} differences {fd:.10e}")
        assert abs(g - fd) <= 1e-6 * abs(fd)
