"""The oracle (oracle/liboracle.so) against the reference: golden fixtures recorded from the
reference, whole runs of the live reference where it is built, and closed forms for the pieces."""
import math

import numpy as np
import pytest

from twofold_amd import _abi, scenes
import golden_io
import parity
import refworld


@pytest.mark.parametrize("case", golden_io.all_cases(), ids=lambda c: f"{c['scene']}-{c['step']}")
def test_oracle_matches_golden(case):
    b, c, lists = refworld.oracle_step(case["pre_bodies"], case["pre_contacts"], case["settings"], case["dt"],
                                       refworld.recorded_refresh(case["refreshes"]))
    assert parity.compare(case, b, c, lists) == []


needs_ref = pytest.mark.skipif(not refworld.ref_available(), reason="reference checker not built (make -C oracle ref)")


@needs_ref
@pytest.mark.parametrize("make", [scenes.pyramid, scenes.stack, scenes.bounce, scenes.rolling, scenes.planar_mix,
                                  scenes.tumble, scenes.hub, scenes.planar_pile_small, scenes.edges],
                         ids=lambda f: f.__name__)
def test_oracle_matches_reference_every_step(make):
    scene = make()
    w = refworld.RefWorld(scene)
    checked = 0
    for k in range(scene.steps):
        rec = w.step()
        if rec is None:
            continue
        b, c, lists = refworld.oracle_step(rec.pre_bodies, rec.pre_contacts, w.settings, scene.dt,
                                           refworld.recorded_refresh(rec.refreshes))
        bad = parity.compare(parity.record_dict(rec), b, c, lists)
        bad += parity.compare_touches(rec, c, lists["touch_events"], lists["hits"], lists["touching"])
        assert bad == [], f"{scene.name} step {k}: {bad[:5]}"
        checked += 1
    assert checked > 0


@needs_ref
def test_oracle_matches_reference_small_pile():
    """A 6x6x6 drop-and-settle pile: many fast pairs, refreshes, impacts and colours."""
    scene = scenes.pile(side=6, half_room=4.0, steps=150)
    w = refworld.RefWorld(scene)
    refreshed = impacts = 0
    for k in range(scene.steps):
        rec = w.step()
        if rec is None:
            continue
        refreshed += len(rec.refreshes)
        impacts += len(rec.lists["impact"]) > 0
        b, c, lists = refworld.oracle_step(rec.pre_bodies, rec.pre_contacts, w.settings, scene.dt,
                                           refworld.recorded_refresh(rec.refreshes))
        bad = parity.compare(parity.record_dict(rec), b, c, lists)
        bad += parity.compare_touches(rec, c, lists["touch_events"], lists["hits"], lists["touching"])
        assert bad == [], f"step {k}: {bad[:5]}"
    assert refreshed > 0 and impacts > 0


def test_maxf_minf_special_values():
    """maxf/minf follow std::fmax/std::fmin (src/internal.h:26-45): NaN yields the other operand, and
    of two zeros max gives -0 only when both are -0, min gives +0 only when both are +0."""
    lib = refworld._load_oracle()
    nan = float("nan")
    assert lib.tfo_maxf(nan, 2.0) == 2.0 and lib.tfo_maxf(2.0, nan) == 2.0
    assert lib.tfo_minf(nan, -1.0) == -1.0
    assert math.copysign(1, lib.tfo_maxf(-0.0, 0.0)) == 1 and math.copysign(1, lib.tfo_maxf(0.0, -0.0)) == 1
    assert math.copysign(1, lib.tfo_maxf(-0.0, -0.0)) == -1
    assert math.copysign(1, lib.tfo_minf(-0.0, 0.0)) == -1 and math.copysign(1, lib.tfo_minf(0.0, 0.0)) == 1
    assert lib.tfo_maxf(3.0, 1.0) == 3.0 and lib.tfo_minf(3.0, 1.0) == 1.0


def test_make_soft_closed_form():
    """makeSoft (src/solver.cpp:255-262): the soft constraint coefficients of a damped spring of
    frequency f and damping ratio z over step h: bias = w / (2z + hw), mass = hw(2z + hw)/(1 + hw(2z + hw)),
    impulse = 1/(1 + hw(2z + hw)), w = 2 pi f; zero stiffness gives (0, 1, 0)."""
    lib = refworld._load_oracle()
    out = np.zeros(3, np.float32)
    for f, z, h in ((30.0, 10.0, 1 / 240), (60.0, 10.0, 1 / 480), (5.0, 0.5, 1 / 60)):
        lib.tfo_make_soft(f, z, h, out.ctypes.data_as(refworld._fp))
        w = 2 * math.pi * f
        a1 = 2 * z + h * w
        a2 = h * w * a1
        np.testing.assert_allclose(out, [w / a1, a2 / (1 + a2), 1 / (1 + a2)], rtol=2e-6)
    lib.tfo_make_soft(0.0, 10.0, 1 / 240, out.ctypes.data_as(refworld._fp))
    assert out.tolist() == [0.0, 1.0, 0.0]


def _body(q, inertia, locks=0, lock_frame=(0, 0, 0, 1), inv_mass=0.5):
    b = np.zeros((), _abi.BODY_DTYPE)
    b["q"] = q
    b["inertia"] = np.asarray(inertia, np.float32).reshape(9)
    b["lock_frame"] = lock_frame
    b["locks"] = locks
    b["inv_mass"] = inv_mass
    b["type"] = _abi.DYNAMIC
    b["awake"] = 1
    b["alive"] = 1
    return b


def test_update_inverse_mass_against_numpy():
    """updateInverseMass (src/world.cpp:62-92): the world inverse inertia is (R I R^T)^-1 and, for a
    planar body, the inverse inertia about the plane normal is n . I_w^-1 n restricted to the free axis."""
    lib = refworld._load_oracle()
    rng = np.random.default_rng(3)
    for _ in range(20):
        q = rng.normal(size=4).astype(np.float32)
        q /= np.linalg.norm(q)
        a = rng.normal(size=(3, 3))
        inertia = (a @ a.T + 3 * np.eye(3)).astype(np.float32)
        b = _body(q, inertia)
        lib.tfo_update_inverse_mass(b.ctypes.data)
        x, y, z, w = q.astype(np.float64)
        r = np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
                      [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
                      [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)]])
        want = np.linalg.inv(r @ inertia.astype(np.float64) @ r.T)
        np.testing.assert_allclose(b["inv_inertia_world"].reshape(3, 3), want, rtol=2e-4, atol=2e-5)
        np.testing.assert_array_equal(b["inv_mass_world"].reshape(3, 3), np.eye(3, dtype=np.float32) * 0.5)
    # planar: free to turn about z only, so the inverse inertia about z is 1 / I_zz
    b = _body((0, 0, 0, 1), np.diag([2.0, 3.0, 4.0]), locks=_abi.LOCK_PLANAR)
    lib.tfo_update_inverse_mass(b.ctypes.data)
    assert b["plane_inv_inertia"] == np.float32(0.25) and b["plane_inv_mass"] == np.float32(0.5)
    np.testing.assert_array_equal(b["inv_mass_world"].reshape(3, 3), np.diag([0.5, 0.5, 0.0]).astype(np.float32))


def _brute_colors(edges, moving, nb):
    """Greedy first-fit in list order, written out independently of the oracle."""
    used = [set() for _ in range(nb)]
    out = []
    for a, b in edges:
        taken = (used[a] if moving[a] else set()) | (used[b] if moving[b] else set())
        col = next((k for k in range(32) if k not in taken), 32)
        if col < 32:
            if moving[a]:
                used[a].add(col)
            if moving[b]:
                used[b].add(col)
        out.append(col)
    return out


def test_coloring_brute_force_and_invariants():
    """colorContacts (src/lanes.cpp:356-374) on random contact graphs: equal to a brute-force greedy
    colouring, no two contacts of one colour share a moving body, and each contact's colour is the
    lowest one free on its moving bodies among the earlier contacts."""
    lib = refworld._load_oracle()
    rng = np.random.default_rng(11)
    for trial in range(30):
        nb = int(rng.integers(3, 40))
        nc = int(rng.integers(1, 200))
        bodies = np.zeros(nb, _abi.BODY_DTYPE)
        bodies["type"] = np.where(rng.random(nb) < 0.15, _abi.STATIC, _abi.DYNAMIC)
        bodies["awake"] = rng.random(nb) < 0.9
        bodies["alive"] = 1
        contacts = np.zeros(nc, _abi.CONTACT_DTYPE)
        edges = []
        for i in range(nc):
            a, b = rng.choice(nb, 2, replace=False)
            contacts[i]["body_a"], contacts[i]["body_b"] = a, b
            edges.append((int(a), int(b)))
        moving = (bodies["type"] == _abi.DYNAMIC) & (bodies["awake"] == 1)
        lst = np.arange(nc, dtype=np.uint32)
        colors = np.zeros(nc, np.uint8)
        lib.tfo_color(bodies.ctypes.data, nb, contacts.ctypes.data, lst.ctypes.data, nc, colors.ctypes.data)
        assert colors.tolist() == _brute_colors(edges, moving, nb)
        for col in range(32):
            seen = set()
            for (a, b), c in zip(edges, colors):
                if c != col:
                    continue
                for x in (a, b):
                    if moving[x]:
                        assert x not in seen
                        seen.add(x)


@needs_ref
def test_invariants_hold_and_catch_planted_mistakes():
    """The size-independent checks used at full size: clean on a real step of a pile, and each one
    fails when a mistake is planted in the oracle's results."""
    scene = scenes.pile(side=6, half_room=4.0, steps=60)
    w = refworld.RefWorld(scene)
    rec = None
    for _ in range(60):
        rec = w.step() or rec
    b, c, lists = refworld.oracle_step(rec.pre_bodies, rec.pre_contacts, w.settings, scene.dt,
                                       refworld.recorded_refresh(rec.refreshes))
    assert parity.invariants(rec.pre_bodies, rec.pre_contacts, c, b, lists) == []
    coloured = [ci for ci in lists["rows"] if lists["colors"][ci] < 32]
    assert len(coloured) > 10
    bad_colors = dict(lists, colors=lists["colors"].copy())
    bad_colors["colors"][coloured[-1]] = (bad_colors["colors"][coloured[-1]] + 1) % 32
    assert any("colour" in m for m in parity.invariants(rec.pre_bodies, rec.pre_contacts, c, b, bad_colors))
    rows = lists["rows"].copy()
    rows[[0, -1]] = rows[[-1, 0]]
    assert any("row order" in m for m in parity.invariants(rec.pre_bodies, rec.pre_contacts, c, b, dict(lists, rows=rows)))
    awake = np.nonzero((b["type"] == 2) & (rec.pre_bodies["awake"] != 0))[0]
    bad_b = b.copy()
    bad_b["island"][awake[-1]] += 1
    assert any("island" in m for m in parity.invariants(rec.pre_bodies, rec.pre_contacts, c, bad_b, lists))


@needs_ref
def test_edges_scene_reaches_its_branches():
    """The edge scene exercises what the other scenes leave alone (its recipe says which lines): per
    step, the reference's own solve-block records show damped, scaled and forced bodies, speeds
    clamped (`kept` below 1), kinematic bodies touching dynamic ones, bodies carried below the sleep
    threshold by a kinematic platform that keep a zero sleep time, and restitution applied to planar
    and to 3D contacts outside the colours."""
    scene = scenes.edges()
    w = refworld.RefWorld(scene)
    seen = dict(kept=0, kinematic=0, planar_bounce=0, rest_bounce=0, forced=0, damped=0, scaled=0)
    carried = [i for i, b in enumerate(scene.bodies) if b["type"] == scenes.DYNAMIC and b["pos"][2] == 3 and b["pos"][0] < -5]
    for _ in range(scene.steps):
        rec = w.step()
        if rec is None:
            continue
        pb, pc = rec.pre_bodies, rec.pre_contacts
        awake = (pb["type"] == scenes.DYNAMIC) & (pb["awake"] == 1)
        seen["kept"] += int((rec.post_bodies["kept"] < 1).sum())
        seen["forced"] += int((awake & (np.abs(pb["force"]).sum(1) > 0) & (np.abs(pb["torque"]).sum(1) > 0)).sum())
        seen["damped"] += int((awake & (pb["linear_damping"] > 0)).sum())
        seen["scaled"] += int((awake & (pb["gravity_scale"] != 1)).sum())
        for ci in rec.lists["active"]:
            if pc["count"][ci] > 0 and scenes.KINEMATIC in (pb["type"][pc["body_a"][ci]], pb["type"][pc["body_b"][ci]]):
                seen["kinematic"] += 1
        post = rec.post_contacts
        for ci in rec.lists["rest"]:
            n = post["count"][ci]
            p = post["points"][ci][:n]
            if post["restitution"][ci] > 0 and (p["relative_velocity"] < -1).any() and (p["max_normal_impulse"] != 0).any():
                seen["planar_bounce" if post["planar"][ci] else "rest_bounce"] += 1
        for i in carried:
            sb = rec.step_bodies[i]
            if sb["awake"] and np.linalg.norm(sb["v"]) < 0.05:
                assert sb["sleep_time"] == 0, "a body carried by a moving kinematic body kept its sleep timer"
    assert all(v > 0 for v in seen.values()), seen


@needs_ref
@pytest.mark.parametrize("make", [lambda: scenes.pile(side=8, half_room=6.0, steps=150), scenes.tumble, scenes.edges,
                                  lambda: scenes.pyramid(rows=12, steps=150)],
                         ids=["pile8", "tumble", "edges", "pyramid12"])
def test_oracle_contact_stream_matches_reference(make):
    """The contact stream (reuse of the contacts that still fit, points measured again, new manifolds
    with impulses carried from the closest old contact, resting pairs kept, all merged in key order)
    against what the reference's updateContacts left, every step, chained: each step streams onto the
    contacts and reuse data the oracle's own last step left (a mid-step refresh leaves the contacts it
    collided again not reusable, src/world.cpp:1110, 1117)."""
    scene = make()
    w = refworld.RefWorld(scene)
    kept = fresh = 0
    prev = None
    for k in range(scene.steps):
        rec = w.step()
        if rec is None:
            prev = None
            continue
        old, old_ext = (rec.old_contacts, rec.old_ext) if prev is None else prev
        c, e = refworld.oracle_stream(rec.old_bodies, old, old_ext, rec.pairs, rec.manifolds, w.settings)
        bad = parity.compare_stream(rec, c, e)
        assert bad == [], f"{scene.name} step {k}: {bad[:5]}"
        fresh += len(rec.manifolds)
        kept += int(rec.pre_ext["reusable"].sum())
        e = e.copy()
        _, c2, _ = refworld.oracle_step(rec.pre_bodies, c, w.settings, rec.dt, refworld.recorded_refresh(rec.refreshes), ext=e)
        prev = (c2, e)
    assert fresh > 0 and kept > 0
