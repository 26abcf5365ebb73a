"""GPU checks of the isothermal partial-bounce-back step and its
hand-written adjoint: against the torch FP64 restatement (forward values and
its autograd gradient), against central finite differences of the engine's
own forward (h = 1e-6), and through the torch.autograd.Function."""
import numpy as np
import pytest
import torch

import iso_reference as R

pytestmark = pytest.mark.gpu

SHAPE = (8, 8, 16)   # nz, ny, nx
T0 = 1.0 / 3.0
U_IN = 0.04
TAU = 0.6
NSTEPS = 12


def _inputs():
    ns = torch.tensor(R.blobs(SHAPE, 5, mean=0.05, nblob=3, width=2.0))
    u0 = torch.zeros((3,) + SHAPE, dtype=torch.float64)
    u0[0] = U_IN
    u0[1] = 0.01 * torch.sin(torch.arange(SHAPE[2], dtype=torch.float64) * 2 * np.pi / SHAPE[2])
    return ns, u0


@pytest.fixture(scope="module")
def solver(engine):
    from marbles_amd.adjoint import IsoSolver

    S = IsoSolver((SHAPE[2], SHAPE[1], SHAPE[0]), T0=T0, u_in=U_IN)
    yield S
    S.close()


@pytest.fixture(scope="module")
def reference():
    ns, u0 = _inputs()
    ns_r = ns.clone().requires_grad_(True)
    u0_r = u0.clone().requires_grad_(True)
    tau_r = torch.tensor(TAU, dtype=torch.float64, requires_grad=True)
    J, f = R.run(SHAPE, ns_r, tau_r, u0_r, U_IN, T0, NSTEPS)
    J.backward()
    return {"J": J.item(), "f": f.detach().numpy(), "dtau": tau_r.grad.item(), "dns": ns_r.grad.numpy(),
            "du0": u0_r.grad.numpy()}


def test_forward_matches_torch_restatement(solver, reference):
    ns, u0 = _inputs()
    J = solver.forward(ns.cuda(), TAU, u0.cuda(), NSTEPS)
    f = solver.state()
    assert abs(J - reference["J"]) <= 1e-12 * abs(reference["J"])
    assert np.abs(f - reference["f"]).max() <= 1e-12 * np.abs(reference["f"]).max()


def test_adjoint_matches_torch_autograd(solver, reference):
    ns, u0 = _inputs()
    J, dtau, dns, du0 = solver.gradient(ns.cuda(), TAU, u0.cuda(), NSTEPS)
    assert abs(J - reference["J"]) <= 1e-12 * abs(reference["J"])
    assert abs(dtau - reference["dtau"]) <= 1e-9 * abs(reference["dtau"])
    for got, ref in ((dns.cpu().numpy(), reference["dns"]), (du0.cpu().numpy(), reference["du0"])):
        assert np.abs(got - ref).max() <= 1e-9 * np.abs(ref).max()


@pytest.mark.parametrize("every", [1, 3, 5, NSTEPS, -1])
def test_checkpoint_interval_does_not_change_the_gradient(solver, every):
    """sqrt(N) checkpoints against other intervals: every state kept (no
    recomputation) and the memory-aware choice, which keeps every state of
    this small run."""
    ns, u0 = _inputs()
    a = solver.gradient(ns.cuda(), TAU, u0.cuda(), NSTEPS, every=0)
    assert solver.record_interval() == 4
    b = solver.gradient(ns.cuda(), TAU, u0.cuda(), NSTEPS, every=every)
    assert solver.record_interval() == (NSTEPS if every < 0 else every)
    assert a[0] == b[0] and a[1] == b[1]
    assert torch.equal(a[2], b[2]) and torch.equal(a[3], b[3])


def test_adjoint_matches_central_differences(solver):
    """h = 1e-6 in FP64 on tau, on n_s of 32 seeded cells and on u0 of 8
    seeded (cell, component) pairs."""
    ns, u0 = _inputs()
    ns, u0 = ns.cuda(), u0.cuda()
    J, dtau, dns, du0 = solver.gradient(ns, TAU, u0, NSTEPS)
    h = 1e-6

    def fd(fun):
        return (fun(+h) - fun(-h)) / (2 * h)

    checks = [(dtau, fd(lambda e: solver.forward(ns, TAU + e, u0, NSTEPS)))]
    rng = np.random.default_rng(7)
    N = ns.numel()
    for c in rng.choice(N, 32, replace=False):
        def fun(e, c=int(c)):
            p = ns.clone().view(-1)
            p[c] += e
            return solver.forward(p.view(ns.shape), TAU, u0, NSTEPS)
        checks.append((dns.view(-1)[int(c)].item(), fd(fun)))
    for c in rng.choice(3 * N, 8, replace=False):
        def fun(e, c=int(c)):
            p = u0.clone().view(-1)
            p[c] += e
            return solver.forward(ns, TAU, p.view(u0.shape), NSTEPS)
        checks.append((du0.view(-1)[int(c)].item(), fd(fun)))
    scale = max(abs(g) for g, _ in checks)
    for g, d in checks:
        assert abs(g - d) <= 1e-6 * scale, (g, d)


def test_autograd_function_returns_the_adjoint_gradients(solver):
    from marbles_amd.adjoint import flow_loss

    ns, u0 = _inputs()
    ns = ns.cuda().requires_grad_(True)
    u0 = u0.cuda().requires_grad_(True)
    tau = torch.tensor(TAU, dtype=torch.float64, device="cuda", requires_grad=True)
    J = flow_loss(solver, ns, tau, u0, NSTEPS)
    (2.0 * J).backward()
    _, dtau, dns, du0 = solver.gradient(ns.detach(), TAU, u0.detach(), NSTEPS)
    assert tau.grad.item() == 2.0 * dtau
    assert torch.equal(ns.grad, 2.0 * dns) and torch.equal(u0.grad, 2.0 * du0)


def test_record_then_reverse_equals_the_gradient(solver):
    ns, u0 = _inputs()
    ns, u0 = ns.cuda(), u0.cuda()
    J, dtau, dns, du0 = solver.gradient(ns, TAU, u0, NSTEPS, every=4)
    J2, token = solver.record(ns, TAU, u0, NSTEPS, every=4)
    dtau2, dns2, du02 = solver.reverse(token)
    assert J == J2 and dtau == dtau2
    assert torch.equal(dns, dns2) and torch.equal(du0, du02)
    # the recording is used up, and a forward run replaces a new one
    assert solver.reverse(token) is None
    _, token = solver.record(ns, TAU, u0, NSTEPS)
    solver.forward(ns, TAU, u0, NSTEPS)
    assert solver.reverse(token) is None


def test_autograd_with_two_evaluations_before_backward(solver):
    """The second evaluation replaces the first one's recording; the first
    one's backward then recomputes its gradient from scratch."""
    from marbles_amd.adjoint import flow_loss

    ns, u0 = _inputs()
    ns = ns.cuda().requires_grad_(True)
    u0 = u0.cuda().requires_grad_(True)
    t1 = torch.tensor(TAU, dtype=torch.float64, device="cuda", requires_grad=True)
    t2 = torch.tensor(0.7, dtype=torch.float64, device="cuda", requires_grad=True)
    (flow_loss(solver, ns, t1, u0, NSTEPS) + flow_loss(solver, ns, t2, u0, NSTEPS)).backward()
    _, d1, n1, v1 = solver.gradient(ns.detach(), TAU, u0.detach(), NSTEPS)
    _, d2, n2, v2 = solver.gradient(ns.detach(), 0.7, u0.detach(), NSTEPS)
    assert t1.grad.item() == d1 and t2.grad.item() == d2
    assert torch.allclose(ns.grad, n1 + n2, rtol=1e-14, atol=0) and torch.allclose(u0.grad, v1 + v2, rtol=1e-14, atol=0)
