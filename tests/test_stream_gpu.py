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
