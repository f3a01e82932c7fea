"""The hand-written adjoint of the RK2 step on the GPU against central
differences of the loss, on a small binary disk (NumR 24, r in [0.5, 3],
q = 0.5, alpha = 0.01, 10 RK2 steps, checkpoints every 4 steps):

* dL/d alpha and dL/d q against differences (h = 1e-6 relative) of the GPU
  loss and of the oracle's loss (the CPU restatement that is bit for bit the
  reference), to 1e-6;
* dL/d(initial primitives) of chosen zones against differences of the GPU
  loss;
* the loss itself against the oracle's;
* the checkpoint interval does not change the gradients, and a second call
  gives the same bits (every sum of the backward runs in a fixed order).

"Away from kinks": the seeded disk of the benchmark has v_r = 0 and a
phi-independent omega and sound speed, so every phi limiter of v_r and omega
and the CFL minimum over a whole annulus sit exactly at a kink of min / minmod
(the one-sided derivatives differ and a central difference averages them).
The per-zone checks therefore use that state with a 1e-3 relative
perturbation of P, v_r and omega; the parameter checks use both states.

Loss: the time-averaged gravitational torque on the gas,
L = (1/T) sum_steps sum_substeps sum_interior dt_s dV rho F_phi r.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

STEPS = 10
REL = 1e-6


def case(alpha=0.01, q=0.5):
    from disco2_amd import configs
    return configs.binary(24, r_min=0.5, r_max=3.0, q=q, alpha=alpha)


def initial_state(p, kind, kw, untie):
    from disco2_amd import ics
    from disco2_amd.api import Setup
    s = Setup(p)
    pr = ics.make_ic(kind, s.geometry(), p, **kw)
    if untie:
        rs = np.random.RandomState(7)
        n = pr.shape[1]
        cs = np.sqrt(pr[1] / pr[0])
        pr[1] *= 1 + 1e-3 * rs.uniform(-1, 1, n)
        pr[2] += 1e-3 * cs * rs.uniform(-1, 1, n)
        pr[3] *= 1 + 1e-3 * rs.uniform(-1, 1, n)
    return pr, s


def gpu(p, prim0, adjoint=False, every=4):
    from disco2_amd.api import Batch, Setup
    b = Batch([Setup(p)])
    b.upload("prim", prim0)
    b.init()
    try:
        return b.adjoint(STEPS, checkpoint_every=every) if adjoint else b.torque_loss(STEPS)
    finally:
        b.close()


def oracle_loss(p, prim0):
    from disco2_amd.api import Setup
    from oracle.oracle import Oracle
    s = Setup(p)
    o = Oracle(p, s.N_p, s.r_faces, s.z_faces)
    o.set_state(tiph=s.tiph0, prim=prim0)
    o.init()
    t0 = o.scalars()[0]
    o.step(STEPS)
    return o.torque() / (o.scalars()[0] - t0)


@pytest.fixture(scope="module", params=[False, True], ids=["seeded", "untied"])
def adj(request):
    p, kind, kw = case()
    prim0, s = initial_state(p, kind, kw, request.param)
    print(f"adjoint: binary NumR 24, {STEPS} steps, checkpoints every 4, {request.param and 'untied' or 'seeded'}",
          flush=True)
    loss, ga, gq, gp = gpu(p, prim0, adjoint=True)
    return dict(loss=loss, ga=ga, gq=gq, gp=gp, prim0=prim0, setup=s, untied=request.param)


def central(f, x, h):
    return (f(x + h) - f(x - h)) / (2 * h)


def test_loss_matches_oracle(adj):
    p, _, _ = case()
    lo = oracle_loss(p, adj["prim0"])
    print(f"loss GPU {adj['loss']:.15e} oracle {lo:.15e}")
    assert abs(adj["loss"] - lo) <= 1e-11 * abs(lo)


@pytest.mark.parametrize("name", ["alpha", "q"])
def test_parameter_gradient_matches_differences(adj, name):
    x0 = 0.01 if name == "alpha" else 0.5
    h = 1e-6 * x0

    def params(x):
        return case(alpha=x)[0] if name == "alpha" else case(q=x)[0]

    fd = central(lambda x: gpu(params(x), adj["prim0"]), x0, h)
    fdo = central(lambda x: oracle_loss(params(x), adj["prim0"]), x0, h)
    g = adj["ga"] if name == "alpha" else adj["gq"]
    print(f"dL/d{name}: adjoint {g:.10e}  GPU differences {fd:.10e}  oracle differences {fdo:.10e}")
    assert abs(g - fd) <= REL * abs(fd)
    assert abs(g - fdo) <= REL * abs(fdo)


def test_initial_state_gradient_matches_differences(adj):
    if not adj["untied"]:
        pytest.skip("per-zone checks need the state away from the min/minmod kinks")
    p, _, _ = case()
    s, prim0, gp = adj["setup"], adj["prim0"], adj["gp"]
    aoff = np.concatenate([[0], np.cumsum(s.N_p)])
    gmax = np.max(np.abs(gp))
    worst = 0.0
    n = 0
    for i in (0, 1, 2, s.NR // 2, s.NR - 3, s.NR - 2):
        for j in (3, int(s.N_p[i]) // 2):
            z = int(aoff[i] + j)
            for q in (0, 1, 2, 3):  # rho, P, v_r, omega
                x0 = prim0[q, z]
                h = 1e-5 * max(abs(x0), 1e-4)

                def f(x):
                    pr = prim0.copy()
                    pr[q, z] = x
                    return gpu(p, pr)

                fd = central(f, x0, h)
                err = abs(gp[q, z] - fd) / (abs(fd) + 1e-2 * gmax)
                worst = max(worst, err)
                n += 1
                assert err <= 1e-4, f"zone {z} q {q}: adjoint {gp[q, z]:.6e} differences {fd:.6e}"
    print(f"{n} initial-state derivatives, worst difference {worst:.2e} of |fd| + 1e-2 max|g|")


def test_checkpoint_interval_does_not_change_gradients(adj):
    """The whole run recorded and checkpoints every 4 steps (recomputed
    segments, the recompute beside the reverse sweep on a second stream)
    give the same bits: the recompute repeats the forward exactly and the
    backward's sums run in a fixed order."""
    p, _, _ = case()
    loss, ga, gq, gp = gpu(p, adj["prim0"], adjoint=True, every=STEPS)
    assert loss == adj["loss"]
    assert ga == adj["ga"]
    assert gq == adj["gq"]
    assert np.array_equal(gp, adj["gp"])


def test_adjoint_is_reproducible(adj):
    p, _, _ = case()
    loss, ga, gq, gp = gpu(p, adj["prim0"], adjoint=True)
    assert (loss, ga, gq) == (adj["loss"], adj["ga"], adj["gq"])
    assert np.array_equal(gp, adj["gp"])


def test_inviscid_adjoint_matches_differences():
    """Explicit_Viscosity 0: the step takes no viscous branch and neither
    does the backward (no viscous face kernels).  dL/dq and chosen initial
    primitives against central differences of the GPU loss on the untied
    state; dL/d alpha is the derivative of the inviscid branch, 0 (the loss
    has a kink at alpha = 0: from the right it grows with the viscous flux)."""
    p, kind, kw = case(alpha=0.0)
    prim0, s = initial_state(p, kind, kw, True)
    print(f"inviscid adjoint: binary NumR 24, {STEPS} steps", flush=True)
    loss, ga, gq, gp = gpu(p, prim0, adjoint=True)
    assert ga == 0.0
    assert abs(loss - oracle_loss(p, prim0)) <= 1e-11 * abs(loss)
    fd = central(lambda x: gpu(case(alpha=0.0, q=x)[0], prim0), 0.5, 0.5e-6)
    print(f"dL/dq: adjoint {gq:.10e}  GPU differences {fd:.10e}")
    assert abs(gq - fd) <= REL * abs(fd)
    aoff = np.concatenate([[0], np.cumsum(s.N_p)])
    gmax = np.max(np.abs(gp))
    for i in (1, s.NR // 2, s.NR - 3):
        z = int(aoff[i] + 3)
        for q in (0, 1, 2, 3):
            x0 = prim0[q, z]
            h = 1e-5 * max(abs(x0), 1e-4)

            def f(x):
                pr = prim0.copy()
                pr[q, z] = x
                return gpu(p, pr)

            d = central(f, x0, h)
            assert abs(gp[q, z] - d) / (abs(d) + 1e-2 * gmax) <= 1e-4, f"zone {z} q {q}: {gp[q, z]:.6e} vs {d:.6e}"


@pytest.mark.parametrize("variant", ["hll", "fixed_mesh"])
def test_adjoint_variants_match_differences(variant):
    """The HLL solver's backward, and a mesh at rest (w_numeric_type 1, no
    annulus rotation to differentiate), on the untied state: the loss against
    the oracle's, dL/d alpha and dL/d q against central differences of the
    GPU loss and of the oracle's loss.  Steps: 1e-6 relative in q (the loss
    is curved in q: at 1e-5 the truncation error is 3e-6 with HLL) and 1e-5
    in alpha (the mesh at rest has a small dL/d alpha, 0.085 against 1.5
    with HLL, and at 1e-6 the round-off of the losses' difference reaches
    1e-6 of it: GPU and oracle differences 9e-7 apart)."""
    from disco2_amd.params import C_FIXED, HLL
    over = {"hll": {"RiemannSolver": HLL}, "fixed_mesh": {"w_numeric_type": C_FIXED}}[variant]

    def params(alpha=0.01, q=0.5):
        p = case(alpha=alpha, q=q)[0]
        p.update(over)
        return p

    p = params()
    _, kind, kw = case()
    prim0, _ = initial_state(p, kind, kw, True)
    print(f"adjoint {variant}: binary NumR 24, {STEPS} steps", flush=True)
    loss, ga, gq, _ = gpu(p, prim0, adjoint=True)
    assert abs(loss - oracle_loss(p, prim0)) <= 1e-11 * abs(loss)
    for name, g, x0 in (("alpha", ga, 0.01), ("q", gq, 0.5)):
        h = (1e-5 if name == "alpha" else 1e-6) * x0
        mk = (lambda x: params(alpha=x)) if name == "alpha" else (lambda x: params(q=x))
        fd = central(lambda x: gpu(mk(x), prim0), x0, h)
        fdo = central(lambda x: oracle_loss(mk(x), prim0), x0, h)
        print(f"dL/d{name}: adjoint {g:.10e}  GPU differences {fd:.10e}  oracle differences {fdo:.10e}")
        assert abs(g - fd) <= REL * abs(fd)
        assert abs(g - fdo) <= REL * abs(fdo)


def test_out_of_scope_is_rejected():
    """3D, another frame, state-dependent mesh motion: the adjoint refuses."""
    from disco2_amd import configs, ics
    from disco2_amd._lib import DiscoError
    from disco2_amd.api import Batch, Setup
    p, kind, kw = configs.vortex(16)  # C_RIGID: the mesh velocity depends on the state
    s = Setup(p)
    b = Batch([s])
    b.upload("prim", ics.make_ic(kind, s.geometry(), p, **kw))
    b.init()
    with pytest.raises((DiscoError, RuntimeError), match="adjoint"):
        b.adjoint(2)
    b.close()
