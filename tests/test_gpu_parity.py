"""The HIP path against the reference, bit for bit: golden fixtures recorded from the reference,
whole runs of the live reference where its checker is built, and the oracle on a larger pile.

Each step feeds the GPU the exact inputs the reference's solve block saw (bodies and the contact
manifolds from the reference's narrowphase) and compares lists, colours, row order, island ids and
sleep flags exactly, and every velocity, pose and impulse bit for bit."""
import numpy as np
import pytest

from twofold_amd import scenes
from twofold_amd.solver import Solver
import golden_io
import parity
import refworld

pytestmark = pytest.mark.gpu


def recorded(refreshes):
    by = {r["substep"]: r for r in refreshes}

    def fresh(substep, remaining, again, bodies, contacts):
        r = by[substep]
        assert np.array_equal(again, r["again"]), "the GPU asked to refresh other contacts"
        return r["fresh"]

    return fresh


def gpu_step(solver, settings, case):
    solver.set_settings(settings)
    solver.upload(case["pre_bodies"], case["pre_contacts"])
    solver.step(case["dt"], recorded(case["refreshes"]) if case["refreshes"] else None)
    return solver.bodies(), solver.contacts(), solver.lists()


@pytest.fixture(scope="module")
def solver():
    s = Solver(64, 64)
    yield s
    s.close()


@pytest.mark.parametrize("case", golden_io.all_cases(), ids=lambda c: f"{c['scene']}-{c['step']}")
def test_gpu_matches_golden(solver, case):
    print(f"golden {case['scene']} step {case['step']}", flush=True)
    b, c, lists = gpu_step(solver, case["settings"], case)
    assert parity.compare(case, b, c, lists) == []


needs_ref = pytest.mark.skipif(not refworld.ref_available(), reason="reference checker not built (make -C oracle ref)")


def run_against_reference(solver, scene, steps=None):
    w = refworld.RefWorld(scene)
    checked = 0
    for k in range(steps or scene.steps):
        rec = w.step()
        if rec is None:
            continue
        case = parity.record_dict(rec)
        b, c, lists = gpu_step(solver, w.settings, case)
        bad = parity.compare(case, b, c, lists)
        assert bad == [], f"{scene.name} step {k}: {bad[:5]}"
        checked += 1
    return checked


@needs_ref
@pytest.mark.parametrize("make", [scenes.pyramid, scenes.stack, scenes.bounce, scenes.rolling, scenes.planar_mix,
                                  scenes.tumble], ids=lambda f: f.__name__)
def test_gpu_matches_reference_every_step(solver, make):
    scene = make()
    print(f"reference run {scene.name}", flush=True)
    assert run_against_reference(solver, scene) > 0


@needs_ref
def test_gpu_matches_reference_pile():
    """An 8x8x8 drop-and-settle pile of boxes and spheres inside walls: fast pairs refreshed mid-step,
    impact passes, up to a dozen colours."""
    scene = scenes.pile(side=8, half_room=6.0, steps=200)
    print(f"reference run {scene.name}", flush=True)
    s = Solver(1024, 4096)
    try:
        assert run_against_reference(s, scene) > 0
    finally:
        s.close()
