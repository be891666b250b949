"""Batched worlds (tfs_set_worlds) against each world run alone by the reference.

Worlds of uneven body counts (the batched workload's recipe at test sizes) are stepped by the
reference one world at a time, recording the exact inputs of each world's solve block; the GPU then
solves all of them in one resident set, and every world's lists, colours, row order, island ids,
sleep flags, velocities, poses and impulses must match its own run bit for bit."""
import numpy as np
import pytest

from twofold_amd import scenes
from twofold_amd.solver import Solver
import parity
import refworld

pytestmark = pytest.mark.gpu

needs_ref = pytest.mark.skipif(not refworld.ref_available(), reason="reference checker not built (make -C oracle ref)")


def concatenate(recs):
    """One resident set from per-world records: bodies, world starts, contacts with global body
    indices, contact starts, and the refresh callback that replays every world's recorded manifolds."""
    bodies = np.concatenate([r.pre_bodies for r in recs])
    bs = np.cumsum([0] + [len(r.pre_bodies) for r in recs]).astype(np.uint32)
    parts = []
    for r, off in zip(recs, bs[:-1]):
        c = r.pre_contacts.copy()
        c["body_a"] += off
        c["body_b"] += off
        parts.append(c)
    contacts = np.concatenate(parts)
    cs = np.cumsum([0] + [len(r.pre_contacts) for r in recs]).astype(np.int64)

    def refresh(substep, remaining, again, c, b):
        want, fresh = [], []
        for r, off in zip(recs, cs[:-1]):
            for x in r.refreshes:
                if x["substep"] == substep:
                    want.append(x["again"].astype(np.int64) + off)
                    fresh.append(x["fresh"])
        assert np.array_equal(np.concatenate(want), again), "the GPU asked to refresh other contacts"
        return np.concatenate(fresh)

    return bodies, bs, contacts, cs, refresh


def world_view(k, bs, cs, bodies, contacts, lists):
    """World k's part of the batched results, in the world's own indices."""
    b = bodies[bs[k]:bs[k + 1]].copy()
    c = contacts[cs[k]:cs[k + 1]].copy()
    c["body_a"] -= bs[k]
    c["body_b"] -= bs[k]
    out = {}
    for name in parity.LISTS:
        x = np.asarray(lists[name]).astype(np.int64)
        out[name] = (x[(x >= cs[k]) & (x < cs[k + 1])] - cs[k]).astype(np.uint32)
    out["colors"] = np.asarray(lists["colors"])[cs[k]:cs[k + 1]]
    return b, c, out


@needs_ref
def test_batched_worlds_match_each_world_alone():
    sizes = scenes.batch_sizes(12, seed=0x5EED, lo=8, hi=160)
    worlds = [scenes.batch_world(n, 0x1000 + k, steps=50) for k, n in enumerate(sizes)]
    refs = [refworld.RefWorld(w) for w in worlds]
    settings = refs[0].settings
    s = Solver(64, 64, settings=settings)
    checked = refreshed = 0
    try:
        for step in range(50):
            recs = [r.step() for r in refs]
            live = [rec for rec in recs if rec is not None]
            if len(live) < 2:
                continue
            bodies, bs, contacts, cs, refresh = concatenate(live)
            print(f"batched step {step}: {len(live)} worlds, {len(bodies)} bodies, {len(contacts)} contacts", flush=True)
            s.upload_worlds(bodies, bs, contacts)
            s.step(np.float32(worlds[0].dt), refresh)
            got_b, got_c, lists = s.bodies(), s.contacts(), s.lists()
            for k, rec in enumerate(live):
                b, c, l = world_view(k, bs, cs, got_b, got_c, lists)
                bad = parity.compare(parity.record_dict(rec), b, c, l)
                assert bad == [], f"step {step} world {k} ({len(rec.pre_bodies)} bodies): {bad[:5]}"
            refreshed += sum(len(r.refreshes) > 0 for r in live)
            checked += 1
    finally:
        s.close()
    assert checked > 20 and refreshed > 0
