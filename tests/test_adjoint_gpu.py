"""The HIP fixed-contact model and adjoint (tfs_fixed_* / tfs_adjoint_*) against the float64 model of
oracle/fixed_contacts.py on the throw workload: forward centres of mass, colours, and gradients with
respect to initial velocities, spins and per-world friction, directly and through central
differences of the float64 loss; plus the torch.autograd.Function the harness wraps it in.

Tolerances: the HIP path computes in float32 and the model in float64. Forward: centres of mass
within 1e-4 of the displacement's norm. Gradients over 2 steps: within 1e-4 relative, norm-wise per
parameter kind (float32 rounding of the reverse sweep itself). Over the workload's 32 steps the loss is
piecewise smooth (every clamp is a kink) and float32 rounding moves a few marginal clamps to their other
side: the float64 model's own gradient moves by 1e-4 when v0 is perturbed by 1e-6 relative over 16
steps. There the norm-wise comparison with the float64 gradient is held to 1e-3 for velocities and
friction and 5e-3 for spins, whose gradient is the smallest and most sensitive, and so are directional
derivatives against central differences of the float64 loss, along directions with a real projection on
the gradient."""
import numpy as np
import pytest

from twofold_amd.solver import Solver, SolverError
from adjoint_cases import Case

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


@pytest.fixture(scope="module")
def case():
    return Case([0x61, 0x62], steps=32)


def run_gpu(case, v0=None, w0=None, mu=None, steps=None):
    steps = case.steps if steps is None else steps
    b = case.bodies.copy()
    c = case.contacts.copy()
    if v0 is not None:
        b["v"] = v0
    if w0 is not None:
        b["w"] = w0
    if mu is not None:
        c["friction"] = mu
    s = Solver(len(b), len(c), settings=case.settings)
    s.upload_worlds(b, case.body_start, c)
    s.fixed_begin(case.dt, steps)
    for _ in range(steps):
        s.fixed_step()
    return s, s.bodies()["com"].astype(np.float64)


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / max(np.linalg.norm(b), 1e-30))


@pytest.fixture(scope="module")
def gpu_run(case):
    s, com = run_gpu(case)
    colors = s.lists()["colors"]
    s.adjoint_seed(case.dl_dcom(com))
    for _ in range(case.steps):
        s.adjoint_step()
    gv, gw, gmu = s.adjoint_read()
    s.close()
    return dict(com=com, colors=colors, gv=gv, gw=gw, gmu=gmu)


@pytest.fixture(scope="module")
def model_run(case):
    m = case.model
    com = m.forward(steps=case.steps)
    gv, gw, gmu = m.adjoint(case.target, steps=case.steps)
    return dict(com=com, gv=gv, gw=gw, gmu=gmu)


def test_colours_are_the_oracles(case, gpu_run):
    assert np.array_equal(np.asarray(gpu_run["colors"])[:len(case.contacts)], case.colors)


def test_forward_matches_model(case, gpu_run, model_run):
    print(f"{len(case.bodies)} bodies, {len(case.contacts)} contacts, {case.steps} steps", flush=True)
    moved = model_run["com"] - case.model.com0
    err = np.linalg.norm(gpu_run["com"] - model_run["com"]) / np.linalg.norm(moved)
    assert err <= 1e-4, err


def gradient_errors(case, gpu, model):
    mv = case.model.moving
    return dict(v=rel(gpu["gv"][mv], model["gv"][mv]), w=rel(gpu["gw"][mv], model["gw"][mv]),
                friction=rel(case.per_world(gpu["gmu"]), case.per_world(model["gmu"])))


def test_gradients_match_model_over_two_steps():
    short = Case([0x61, 0x62], steps=2)
    s, com = run_gpu(short)
    s.adjoint_seed(short.dl_dcom(com))
    for _ in range(short.steps):
        s.adjoint_step()
    gv, gw, gmu = s.adjoint_read()
    s.close()
    mv, mw, mmu = short.model.adjoint(short.target, steps=short.steps)
    errs = gradient_errors(short, dict(gv=gv, gw=gw, gmu=gmu), dict(gv=mv, gw=mw, gmu=mmu))
    print(errs, flush=True)
    assert max(errs.values()) <= 1e-4, errs


def test_gradients_match_model(case, gpu_run, model_run):
    errs = gradient_errors(case, gpu_run, model_run)
    print(errs, flush=True)
    assert errs["v"] <= 1e-3 and errs["friction"] <= 1e-3 and errs["w"] <= 5e-3, errs


def _direction(rng, g, mask=None):
    """The unit direction of g plus half a seeded random unit direction: the derivative along it is at
    least |g| / 2, so float32 rounding of the gradient (held norm-wise to the tolerances above) is not
    magnified by cancellation, as along a random direction, which among hundreds of components is
    nearly orthogonal to any given gradient."""
    r = rng.standard_normal(g.shape)
    if mask is not None:
        r[~mask] = 0.0
        g = np.where(mask[:, None], g, 0.0) if g.ndim == 2 else np.where(mask, g, 0.0)
    return 0.5 * r / np.linalg.norm(r) + g / np.linalg.norm(g)


def test_gradients_match_central_differences(case, gpu_run, model_run):
    """Directional derivatives of the float64 loss (central, h = 1e-5) against the HIP gradient, to the
    norm-wise tolerances of the comparison with the model: 1e-3 for velocities and friction, 5e-3 for
    spins."""
    m = case.model
    rng = np.random.default_rng(11)
    h = 1e-5
    u = _direction(rng, model_run["gv"], m.moving)
    fd = (m.loss(case.target, v0=m.v0 + h * u, steps=case.steps)
          - m.loss(case.target, v0=m.v0 - h * u, steps=case.steps)) / (2 * h)
    an = float(np.sum(gpu_run["gv"].astype(np.float64) * u))
    assert abs(fd - an) <= 1e-3 * abs(fd), (fd, an)
    uw0 = _direction(rng, model_run["gw"], m.moving)
    fd = (m.loss(case.target, w0=m.w0 + h * uw0, steps=case.steps)
          - m.loss(case.target, w0=m.w0 - h * uw0, steps=case.steps)) / (2 * h)
    an = float(np.sum(gpu_run["gw"].astype(np.float64) * uw0))
    assert abs(fd - an) <= 5e-3 * abs(fd), (fd, an)
    uw = _direction(rng, case.per_world(model_run["gmu"]))
    du = uw[case.world_of_contact]
    fd = (m.loss(case.target, mu=m.mu0 + h * du, steps=case.steps)
          - m.loss(case.target, mu=m.mu0 - h * du, steps=case.steps)) / (2 * h)
    an = float(np.sum(case.per_world(gpu_run["gmu"].astype(np.float64)) * uw))
    assert abs(fd - an) <= 1e-3 * abs(fd), (fd, an)


class Rollout(torch.autograd.Function):
    """Centres of mass after the case's steps as a function of initial velocities, spins and
    per-world friction; backward is the HIP adjoint."""

    @staticmethod
    def forward(ctx, v0, w0, mu, case):
        friction = mu.detach().numpy()[case.world_of_contact].astype(np.float32)
        s, com = run_gpu(case, v0.detach().numpy(), w0.detach().numpy(), friction)
        ctx.solver, ctx.case = s, case
        return torch.from_numpy(com)

    @staticmethod
    def backward(ctx, grad_com):
        s, case = ctx.solver, ctx.case
        s.adjoint_seed(grad_com.numpy())
        for _ in range(case.steps):
            s.adjoint_step()
        gv, gw, gmu = s.adjoint_read()
        s.close()
        return (torch.from_numpy(gv.astype(np.float64)), torch.from_numpy(gw.astype(np.float64)),
                torch.from_numpy(case.per_world(gmu.astype(np.float64))), None)


def test_torch_autograd_function(case, gpu_run):
    v0 = torch.tensor(case.bodies["v"], dtype=torch.float32, requires_grad=True)
    w0 = torch.tensor(case.bodies["w"], dtype=torch.float32, requires_grad=True)
    mu = torch.tensor([float(sc.shapes[0]["friction"]) for sc in case.scenes], dtype=torch.float32,
                      requires_grad=True)
    com = Rollout.apply(v0, w0, mu, case)
    target = torch.from_numpy(case.target)
    moving = torch.from_numpy(case.model.moving)
    loss = (((com - target) ** 2).sum(1) * moving).sum()
    loss.backward()
    assert rel(v0.grad.numpy(), gpu_run["gv"]) <= 1e-5
    assert rel(mu.grad.numpy(), case.per_world(gpu_run["gmu"])) <= 1e-5


def test_state_errors(case):
    s = Solver(len(case.bodies), len(case.contacts), settings=case.settings)
    try:
        s.upload_worlds(case.bodies, case.body_start, case.contacts)
        with pytest.raises(SolverError):
            s.fixed_step()
        s.fixed_begin(case.dt, 1)
        s.fixed_step()
        with pytest.raises(SolverError):
            s.fixed_step()
        s.adjoint_seed(np.zeros((len(case.bodies), 3), np.float32))
        s.adjoint_step()
        with pytest.raises(SolverError):
            s.adjoint_step()
        s.upload_contacts(case.contacts)
        with pytest.raises(SolverError):
            s.adjoint_read()
    finally:
        s.close()
