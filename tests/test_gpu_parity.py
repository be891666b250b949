"""The HIP path against the reference, bit for bit: golden fixtures recorded from the reference,
whole runs of the live reference where its checker is built, and the oracle on a larger pile.

Each step feeds the GPU the exact inputs the reference's solve block saw (bodies and the contact
manifolds from the reference's narrowphase) and compares lists, colours, row order, island ids and
sleep flags exactly, and every velocity, pose and impulse bit for bit."""
import numpy as np
import pytest

from twofold_amd import _abi, scenes
from twofold_amd.solver import Solver
from twofold_amd.world import HostWorld, Simulation
import golden_io
import parity
import refworld

pytestmark = pytest.mark.gpu


def recorded(refreshes):
    by = {r["substep"]: r for r in refreshes}

    def fresh(substep, remaining, again, contacts, bodies):
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
        bad += parity.compare_touches(rec, c, *solver.touches())
        assert bad == [], f"{scene.name} step {k}: {bad[:5]}"
        checked += 1
    return checked


@needs_ref
@pytest.mark.parametrize("make", [scenes.pyramid, scenes.stack, scenes.bounce, scenes.rolling, scenes.planar_mix,
                                  scenes.tumble, scenes.hub, scenes.planar_pile_small, scenes.edges],
                         ids=lambda f: f.__name__)
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


@needs_ref
def test_gpu_matches_reference_full_pile():
    """The 32,768-body pile at full size: 240 drop-and-settle steps with the host world and the GPU
    solver, then two steps of the reference from that body state (its own narrowphase and refresh
    manifolds) replayed on the GPU, bit for bit, and the results' size-independent properties."""
    scene = scenes.pile(side=32)
    print("full pile: settling 240 steps", flush=True)
    sim = Simulation(scene)
    for _ in range(240):
        sim.step()
    bodies = sim.solver.bodies()
    sim.solver.close()
    w = refworld.RefWorld(scene)
    w.set_bodies(bodies)
    s = Solver(len(bodies), 4 * len(bodies))
    try:
        for k in range(2):
            rec = w.step()
            assert rec is not None
            case = parity.record_dict(rec)
            print(f"full pile: step {k}, {len(case['pre_contacts'])} contacts", flush=True)
            b, c, lists = gpu_step(s, w.settings, case)
            assert parity.compare(case, b, c, lists) == []
            assert parity.invariants(case["pre_bodies"], case["pre_contacts"], c, b, lists) == []
            assert len(lists["active"]) > 50000 and len(set(lists["colors"][lists["active"]])) > 8
    finally:
        s.close()


@needs_ref
def test_gpu_matches_reference_full_pile_drop():
    """The 32,768-body pile's drop at full size, the reference stepping from the start: every 40th of
    its first 240 steps (while the bodies fall, pile up and settle, with hundreds of fast pairs
    refreshed mid-step) is replayed on the GPU from the inputs the reference's solve block saw (its
    own narrowphase and refresh manifolds), bit for bit."""
    scene = scenes.pile(side=32)
    w = refworld.RefWorld(scene)
    s = Solver(len(scene.bodies), 4 * len(scene.bodies))
    checked = refreshed = 0
    try:
        for k in range(240):
            if k % 40 != 39:
                w.step(record=False)
                continue
            rec = w.step()
            assert rec is not None
            case = parity.record_dict(rec)
            b, c, lists = gpu_step(s, w.settings, case)
            bad = parity.compare(case, b, c, lists)
            assert bad == [], f"full pile drop step {k}: {bad[:5]}"
            checked += 1
            refreshed += sum(len(r["again"]) for r in case["refreshes"])
            print(f"full pile drop: step {k}, {len(case['pre_contacts'])} contacts, "
                  f"{len(case['refreshes'])} refreshes, bit-exact", flush=True)
    finally:
        s.close()
    assert checked == 6 and refreshed > 0


def same_bits(a, b):
    """Fields of two structured arrays that differ in any bit."""
    out = []
    for f in a.dtype.names:
        x, y = np.ascontiguousarray(a[f]), np.ascontiguousarray(b[f])
        if x.view(np.uint8).tobytes() != y.view(np.uint8).tobytes():
            out.append(f)
    return out


def test_gpu_empty_world():
    print("empty world", flush=True)
    s = Solver(1, 1)
    try:
        s.upload(np.zeros(0, _abi.BODY_DTYPE), np.zeros(0, _abi.CONTACT_DTYPE))
        s.step(np.float32(1 / 60))
        assert len(s.bodies()) == 0 and s.counts()["active"] == 0
    finally:
        s.close()


def test_gpu_bodies_without_contacts_match_oracle():
    """Free flight: bodies of every motion type and no contact at all."""
    print("bodies without contacts", flush=True)
    sc = scenes.Scene("free")
    rng = scenes.Rng(11)
    for i in range(40):
        kind = (_abi.DYNAMIC, _abi.DYNAMIC, _abi.KINEMATIC, _abi.STATIC)[i % 4]
        b = sc.body(kind, pos=(rng.range(-5, 5), rng.range(0, 9), rng.range(-5, 5)),
                    v=(rng.range(-3, 3), rng.range(-3, 3), rng.range(-3, 3)),
                    w=(rng.range(-4, 4), rng.range(-4, 4), rng.range(-4, 4)))
        sc.shape(b, scenes.BOX if i % 2 else scenes.SPHERE, (0.3, 0.2, 0.4) if i % 2 else (0.3,), density=800.0)
    bodies = HostWorld(sc).bodies()
    contacts = np.zeros(0, _abi.CONTACT_DTYPE)
    settings = _abi.default_settings(substeps=4)
    s = Solver(len(bodies), 1, settings=settings)
    try:
        for step in range(3):
            want, _, _ = refworld.oracle_step(bodies, contacts, settings, np.float32(1 / 60))
            s.upload(bodies, contacts)
            s.step(np.float32(1 / 60))
            got = s.bodies()
            assert same_bits(want, got) == [], f"step {step}"
            bodies = got
    finally:
        s.close()


@needs_ref
@pytest.mark.parametrize("ahead", [1, 2])
def test_gpu_island_rounds_checked_by_the_host(monkeypatch, ahead):
    """The island pass launches a few hook-and-jump rounds before the host first checks whether they
    settled; a pile whose islands need more rounds goes on under the host's control. With only one or
    two rounds launched ahead (TWOFOLD_ISLAND_AHEAD) every step takes that path, and the island ids,
    sleep flags and velocities must still match the reference bit for bit (src/world.cpp:1174-1247)."""
    monkeypatch.setenv("TWOFOLD_ISLAND_AHEAD", str(ahead))
    s = Solver(1024, 4096)
    try:
        for scene in (scenes.pyramid(rows=12, steps=120), scenes.pile(side=6, half_room=4.0, steps=120)):
            print(f"island rounds ahead {ahead}: {scene.name}", flush=True)
            assert run_against_reference(s, scene) > 0
    finally:
        s.close()
