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
    return Case([0x51, 0x52], steps=4)


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
