"""The float64 fixed-contact model's adjoint (oracle/fixed_contacts.py) against central differences
of its loss. These pin the reverse sweep's mathematics; the HIP adjoint is then checked against this
model (tests/test_adjoint_gpu.py)."""
import numpy as np
import pytest

import refworld
from adjoint_cases import Case

needs_oracle = pytest.mark.skipif(not __import__("os").path.exists(refworld.ORACLE_SO),
                                  reason="oracle not built (make -C oracle)")


@pytest.fixture(scope="module")
def case():
    return Case([0x53, 0x54], steps=4)


def directional(case, name, u, h):
    m = case.model

    def loss(e):
        kw = {"v0": m.v0 + e * u} if name == "v" else {"w0": m.w0 + e * u} if name == "w" else {"mu": m.mu0 + e * u}
        return m.loss(case.target, steps=case.steps, **kw)

    return (loss(h) - loss(-h)) / (2 * h)


@needs_oracle
@pytest.mark.parametrize("name", ["v", "w", "mu"])
def test_adjoint_matches_central_differences(case, name):
    m = case.model
    gv, gw, gmu = m.adjoint(case.target, steps=case.steps)
    rng = np.random.default_rng(7)
    shape = m.v0.shape if name != "mu" else m.mu0.shape
    u = rng.standard_normal(shape)
    if name != "mu":
        u[~m.moving] = 0.0
    g = {"v": gv, "w": gw, "mu": gmu}[name]
    fd = directional(case, name, u, 1e-6)
    an = float(np.sum(g * u))
    # float64 central differences at h = 1e-6 cross no clamp of this case; the two agree to rounding
    assert abs(fd - an) <= 1e-6 * max(abs(fd), 1e-3), (fd, an)


@needs_oracle
def test_adjoint_checker_sees_a_planted_mistake(case, monkeypatch):
    """A wrong derivative planted in the model's reverse sweep (d bias / d separation doubled; the
    forward ignores that value) must fail the same comparison, so the check is not vacuous."""
    m = case.model
    rng = np.random.default_rng(7)
    u = rng.standard_normal(m.v0.shape)
    u[~m.moving] = 0.0
    fd = directional(case, "v", u, 1e-6)
    good = float(np.sum(m.adjoint(case.target, steps=case.steps)[0] * u))
    assert abs(fd - good) <= 1e-6 * max(abs(fd), 1e-3)
    right = m._bias

    def wrong(sep, soft, biased):
        bias, ms, isc, dbias = right(sep, soft, biased)
        return bias, ms, isc, 2.0 * dbias

    monkeypatch.setattr(m, "_bias", wrong)
    bad = float(np.sum(m.adjoint(case.target, steps=case.steps)[0] * u))
    assert abs(fd - bad) > 1e-4 * abs(fd), (fd, bad)


@needs_oracle
def test_every_contact_has_a_colour(case):
    assert np.all(case.colors < 32)
    assert case.contacts["count"].min() >= 1


def _columns(seed):
    """Free spheres in vertical columns, touching along y, with vertical velocities only: every
    contact normal is exactly vertical and no tangential velocity ever arises, so no body spins and
    the reference's rotations (turned() with dq, rebaseContact) are exactly the identity that the
    fixed-contact model assumes."""
    from twofold_amd import scenes
    s = scenes.Scene(f"columns-{seed}", substeps=4, steps=2)
    rng = np.random.default_rng(seed)
    for cx in range(3):
        for cz in range(2):
            for k in range(4):
                b = s.body(scenes.DYNAMIC, pos=(2.0 * cx, 0.5 * k - 0.002 * k, 2.0 * cz),
                           v=(0.0, float(rng.normal(0.0, 0.15)), 0.0))
                s.shape(b, scenes.SPHERE, (0.25,), density=1000.0, friction=0.6)
    return s


@needs_oracle
@pytest.mark.parametrize("seed", [3, 4])
def test_model_forward_matches_the_oracle(seed):
    """One step of the float64 model against the C oracle's full step (tfo_step, itself bit-exact with
    the reference) on a scene where the model's simplifications are exact: no refresh, no impact
    pass, no restitution, no rest list and no rotation. The second step carries stored impulses, so
    the warm start is compared too. Agreement to float32 rounding pins the model's forward: contact
    softness, bias, the clamp, the warm start, the substep order and the position update."""
    from twofold_amd.world import HostWorld
    import sys, os
    sys.path.insert(0, os.path.join(refworld.ROOT, "oracle"))
    from fixed_contacts import FixedContacts
    world = HostWorld(_columns(seed))
    settings, dt = world.settings, float(world.dt)
    bodies = world.bodies()
    contacts, _ = world.collide(bodies, world.dt)
    for step in range(2):
        assert len(contacts) > 0
        lib = refworld._load_oracle()
        b = bodies.copy()
        for i in range(len(b)):
            lib.tfo_update_inverse_mass(b[i:i + 1].ctypes.data)
        colors = np.full(len(contacts), 255, np.uint8)
        lst = np.arange(len(contacts), dtype=np.uint32)
        lib.tfo_color(b.ctypes.data, len(b), contacts.ctypes.data, lst.ctypes.data, len(lst), colors.ctypes.data)
        model = FixedContacts(bodies, contacts, colors, settings, dt)
        com = model.forward(steps=1)
        v, w = model.last[0], model.last[1]
        ob, oc, lists = refworld.oracle_step(bodies, contacts, settings, dt)
        assert len(lists["fast"]) == 0 and len(lists["impact"]) == 0 and len(lists["rest"]) == 0
        assert len(lists["active"]) == len(contacts)
        assert np.all(ob["w"] == 0.0)
        if step == 1:
            assert np.any(contacts["points"]["normal_impulse"] > 0.0)
        scale = max(float(np.abs(ob["v"]).max()), 1.0)
        np.testing.assert_allclose(v, ob["v"], rtol=0, atol=2e-6 * scale)
        np.testing.assert_allclose(w, 0.0, atol=0)
        np.testing.assert_allclose(com, ob["com"], rtol=0, atol=2e-6)
        lam = model.last[2]
        ref = oc["points"]["normal_impulse"][:, :1]
        np.testing.assert_allclose(lam[:, :1], ref, rtol=0, atol=1e-5 * max(float(np.abs(ref).max()), 1.0))
        bodies = ob
        contacts, _ = world.collide(bodies, world.dt, oc)
    world.close()


def test_full_size_sample_answers_are_the_models():
    """The kept answers of the float64 model for the sampled full-size adjoint worlds
    (tests/golden/adjoint-throw1024-sample.npz, which the GPU test reads instead of running the model
    for 45 s a world) are still what the model gives: one of the compared worlds, recomputed."""
    import os
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "adjoint-throw1024-sample.npz")
    fx = np.load(path)
    k = int(fx["sample"][1])
    case = Case([int(fx["seed0"]) + k], steps=int(fx["steps"]))
    m = case.model
    assert int(fx[f"w{k}_n_contacts"]) == len(case.contacts)
    np.testing.assert_array_equal(m.forward(steps=int(fx["steps"])), fx[f"w{k}_com"])
    gv, gw, gmu = m.adjoint(case.target, steps=int(fx["steps"]))
    np.testing.assert_array_equal(gv, fx[f"w{k}_gv"])
    np.testing.assert_array_equal(gw, fx[f"w{k}_gw"])
    assert np.sum(gmu) == fx[f"w{k}_gmu_sum"]
