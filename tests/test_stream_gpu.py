"""The contact stream on the GPU (tfs_stream_contacts) against the reference's updateContacts, with
the contacts resident from step to step.

Each step the reference records the state its updateContacts started from (bodies, contacts, their
reuse data), its candidate pairs with their materials and the narrowphase's manifolds, and what it
left. The GPU keeps its contacts: only the first step uploads them; every later step uploads the
bodies (poses may have moved on the host after the solve, e.g. by continuous collision), streams the
pairs and manifolds, and the result must equal the reference's contacts and reuse data bit for bit;
then the step is solved on those resident contacts and checked against the reference as every parity
test is (lists, colours, row order, islands, sleep, velocities, poses, impulses, touch events)."""
import numpy as np
import pytest

from twofold_amd import scenes
from twofold_amd.solver import Solver
import parity
import refworld

pytestmark = pytest.mark.gpu

needs_ref = pytest.mark.skipif(not refworld.ref_available(), reason="reference checker not built (make -C oracle ref)")


def recorded(refreshes):
    by = {r["substep"]: r for r in refreshes}

    def fresh(substep, remaining, again, contacts, bodies):
        r = by[substep]
        assert np.array_equal(again, r["again"]), "the GPU asked to refresh other contacts"
        return r["fresh"]

    return fresh


@needs_ref
@pytest.mark.parametrize("make", [lambda: scenes.pile(side=8, half_room=6.0, steps=150), scenes.tumble, scenes.edges,
                                  lambda: scenes.pyramid(rows=12, steps=150)],
                         ids=["pile8", "tumble", "edges", "pyramid12"])
@pytest.mark.parametrize("resident", [False, True], ids=["uploaded", "resident"])
def test_stream_keeps_contacts_resident_and_matches_reference(make, resident):
    """`uploaded`: every step starts from the reference's own contacts and reuse data; `resident`: only
    the first does, later steps stream onto the contacts the GPU's last step left."""
    scene = make()
    print(f"stream: {scene.name}, resident {resident}", flush=True)
    w = refworld.RefWorld(scene)
    first = True
    kept = 0
    with Solver(1024, 4096, settings=w.settings) as s:
        for k in range(scene.steps):
            rec = w.step()
            if rec is None:
                first = True   # the reference skipped its solve block; start over from its state
                continue
            if first or not resident:
                s.upload(rec.old_bodies, rec.old_contacts)
                s.upload_ext(rec.old_ext)
                first = False
            else:
                s.upload_bodies(rec.old_bodies)
            s.stream(rec.pairs, rec.manifolds)
            bad = parity.compare_stream(rec, s.contacts(), s.ext())
            assert bad == [], f"{scene.name} step {k}: stream {bad[:5]}"
            kept += int(rec.pre_ext["reusable"].sum())
            s.upload_bodies(rec.pre_bodies)
            s.step(rec.dt, recorded(rec.refreshes) if rec.refreshes else None)
            case = parity.record_dict(rec)
            got_c = s.contacts()
            bad = parity.compare(case, s.bodies(), got_c, s.lists())
            bad += parity.compare_touches(rec, got_c, *s.touches())
            assert bad == [], f"{scene.name} step {k}: solve {bad[:5]}"
    assert kept > 0


def test_simulation_with_streamed_contacts():
    """The host world driving the solver with the contacts resident (Simulation(stream=True)): an 8x8x8
    pile dropped for 150 steps; the contacts the device keeps are as many as the host world's own
    narrowphase finds, a growing share of them reused (still fitting, measured again) once the pile
    rests, and the bodies stay finite and inside the walls."""
    from twofold_amd.world import Simulation

    with Simulation(scenes.pile(side=8, half_room=6.0), device_refresh=True, stream=True) as sim:
        reused = 0
        for k in range(150):
            sim.step()
            if k >= 100:
                reused += int(sim.solver.ext()["reusable"].sum())
        bodies = sim.solver.bodies()
        contacts = sim.solver.contacts()
        host, _ = sim.world.collide(bodies.copy(), sim.dt)
    assert np.isfinite(bodies["com"]).all() and (np.abs(bodies["com"][:, [0, 2]]) < 7.0).all()
    assert len(contacts) > 500 and abs(len(contacts) - len(host)) <= 0.05 * len(host), (len(contacts), len(host))
    assert reused > 0


def test_streamed_simulation_matches_the_oracle_step_for_step():
    """The host world driving the GPU with resident contacts against the oracle driven the same way (the
    same candidate pairs and manifolds from the host world at the GPU's poses, the same refresh callback):
    a 10x10x10 pile dropped for 200 steps, bodies and contacts equal bit for bit after every step."""
    import ctypes

    from twofold_amd import _abi
    from twofold_amd.world import Simulation

    lib = refworld._load_oracle()
    with Simulation(scenes.pile(side=10, half_room=7.0), stream=True) as sim:
        hw = sim.world
        fn = ctypes.cast(hw.refresh_fn, refworld.REFRESH_FN)
        ext = np.zeros(0, _abi.EXT_DTYPE)
        contacts = np.zeros(0, _abi.CONTACT_DTYPE)
        for k in range(200):
            bodies = sim.solver.bodies() if k else sim.bodies.copy()
            before = bodies.copy()
            pairs, manifolds, woke = hw.pairs(bodies, hw.dt)
            contacts, ext = refworld.oracle_stream(before, contacts, ext, pairs, manifolds, hw.settings)
            ext = ext.copy()
            sim.solver.stream(pairs, manifolds)
            if woke:
                sim.solver.upload_bodies(bodies)
            sim.solver.step_native(hw.dt, hw.refresh_fn, hw.h)
            b = np.ascontiguousarray(bodies.copy())
            c = np.ascontiguousarray(contacts.copy())
            n = max(len(c), 1)
            arrays = [np.zeros(n, np.uint32) for _ in range(6)]
            colors = np.zeros(n, np.uint8)
            L = refworld._Lists()
            for name, a in zip(("active", "fast", "fast_solid", "impact", "by_color", "impact_by_color"), arrays):
                setattr(L, name, a.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)))
            L.colors = colors.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8))
            ev = np.zeros(n, np.uint32)
            hits = np.zeros(n, _abi.HIT_DTYPE)
            L.touch_events = ev.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32))
            L.hits = hits.ctypes.data
            L.ext = ext.ctypes.data
            assert lib.tfo_step(b.ctypes.data, len(b), c.ctypes.data, len(c), hw.settings.ctypes.data, np.float32(hw.dt), fn,
                                ctypes.c_void_p(hw.h), ctypes.byref(L)) == 0
            contacts = c
            gb, gc, ge = sim.solver.bodies(), sim.solver.contacts(), sim.solver.ext()
            for name, want, got in (("bodies", b, gb), ("contacts", c, gc), ("ext", ext, ge)):
                fields = [f for f in want.dtype.names if f != "pad_" and
                          np.ascontiguousarray(want[f]).tobytes() != np.ascontiguousarray(got[f]).tobytes()]
                assert not fields, f"step {k}: {name} differ in {fields}"
            if k % 50 == 49:
                print(f"streamed pile: step {k + 1}, {len(c)} contacts, {L.n_colors} colours", flush=True)
