"""tfs_step_noalloc, the step that allocates nothing, against tfs_step on the same inputs.

The worlds live in device memory that the test allocates (hipMalloc through the library's own HIP
runtime); the entry point
takes every scratch buffer from its fixed array. Two calls in a row: a 6x6x6 drop pile whose step
collides fast pairs again on the device, then a smaller world (the ten-box stack) that finds the first
call's leftovers in the scratch. Bodies and contacts must match tfs_step's bit for bit."""
import numpy as np
import pytest

from twofold_amd import _abi, scenes
from twofold_amd._abi import BODY_DTYPE, CONTACT_DTYPE, SHAPE_DTYPE
from twofold_amd.solver import DeviceArray, Solver, device_synchronize, load_library
from twofold_amd.world import Simulation
import golden_io

pytestmark = pytest.mark.gpu

TFS_E_CAPACITY = -6


def _lib():
    return load_library()


def noalloc_step(bodies, contacts, shapes, settings, dt):
    lib = _lib()
    db, dc = DeviceArray(bodies), DeviceArray(contacts)
    ds = DeviceArray(shapes) if shapes is not None else None
    s = np.ascontiguousarray(np.asarray(settings, _abi.SETTINGS_DTYPE).reshape(()))
    try:
        rc = lib.tfs_step_noalloc(0, db.ptr, len(bodies), dc.ptr, len(contacts), ds.ptr if ds is not None else None,
                                  s.ctypes.data, np.float32(dt), None)
        assert rc == 0, f"tfs_step_noalloc returned {rc}"
        device_synchronize()
        return db.get(), dc.get()
    finally:
        for x in (db, dc, ds):
            if x is not None:
                x.close()


def plain_step(bodies, contacts, shapes, settings, dt):
    with Solver(len(bodies), max(len(contacts), 1), settings=settings) as s:
        s.upload(bodies, contacts)
        s.set_body_shapes(shapes)
        s.step(np.float32(dt))
        return s.bodies(), s.contacts()


def differing(a, b):
    return [f for f in a.dtype.names if f != "pad_" and
            np.ascontiguousarray(a[f]).view(np.uint8).tobytes() != np.ascontiguousarray(b[f]).view(np.uint8).tobytes()]


def test_noalloc_step_matches_tfs_step_twice():
    # first call: a dropping pile with device refreshes
    sim = Simulation(scenes.pile(side=6, half_room=4.0), device_refresh=True)
    try:
        for _ in range(40):
            sim.step()
        bodies = sim.solver.bodies()
        contacts, _ = sim.world.collide(bodies.copy(), sim.dt, sim.contacts)
        shapes, settings, dt = sim.device_refresh, sim.settings, sim.dt
    finally:
        sim.close()
    assert len(bodies) <= 2048 and len(contacts) <= 8192
    print(f"noalloc call 1: {len(bodies)} bodies, {len(contacts)} contacts", flush=True)
    want_b, want_c = plain_step(bodies, contacts, shapes, settings, dt)
    got_b, got_c = noalloc_step(bodies, contacts, shapes, settings, dt)
    assert differing(want_b, got_b) == [] and differing(want_c, got_c) == []
    assert differing(bodies, got_b) != []   # the step moved something

    # second call: a smaller world, no shapes (its step needs no refresh)
    case = golden_io.load("stack-10")[0]
    settings = _abi.default_settings(substeps=4)
    print(f"noalloc call 2: {len(case['pre_bodies'])} bodies, {len(case['pre_contacts'])} contacts", flush=True)
    want_b, want_c = plain_step(case["pre_bodies"], case["pre_contacts"], None, settings, case["dt"])
    got_b, got_c = noalloc_step(case["pre_bodies"], case["pre_contacts"], None, settings, case["dt"])
    assert differing(want_b, got_b) == [] and differing(want_c, got_c) == []


def test_noalloc_step_refuses_a_world_beyond_its_scratch():
    """Sizes beyond the fixed scratch are refused before any device work."""
    lib = _lib()
    buf = DeviceArray(np.zeros(1024, np.uint8))
    s = _abi.default_settings()
    try:
        rc = lib.tfs_step_noalloc(0, buf.ptr, 2049, buf.ptr, 0, None, s.ctypes.data, np.float32(1 / 60), None)
        assert rc == TFS_E_CAPACITY
        rc = lib.tfs_step_noalloc(0, buf.ptr, 1, buf.ptr, 8193, None, s.ctypes.data, np.float32(1 / 60), None)
        assert rc == TFS_E_CAPACITY
    finally:
        buf.close()
