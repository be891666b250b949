"""The device refresh (tfs_set_body_shapes): the same scene stepped with its mid-step refreshes
collided on the device and through the host world's refresh callback gives the same bodies and
contacts bit for bit after every step, since both run one narrowphase source (narrow.h). Scenes: an
8x8x8 box/sphere pile in walls over its drop and landing, the tumble scene (fast pairs, impacts,
restitution), 16 batched worlds solved together, and the bouncing sphere (sphere on a half-space with
restitution 1). Shapes the device does not collide are refused."""
import numpy as np
import pytest

from twofold_amd import scenes
from twofold_amd._abi import SHAPE_DTYPE
from twofold_amd.solver import SolverError
from twofold_amd.world import Simulation

pytestmark = pytest.mark.gpu

CASES = {
    "pile8": (lambda: scenes.pile(side=8), 150),
    "tumble": (lambda: scenes.tumble(), 120),
    "batched16": (lambda: [scenes.batch_world(n, 0x5EED0000 + k) for k, n in enumerate(scenes.batch_sizes(16))], 60),
    "bounce": (lambda: scenes.bounce(), 120),
}


@pytest.mark.parametrize("name", list(CASES))
def test_device_refresh_matches_the_host_callback(name):
    make, steps = CASES[name]
    host = Simulation(make())
    dev = Simulation(make(), device_refresh=True)
    assert dev.device_refresh is not None and len(dev.device_refresh) == len(dev.bodies)
    host.solver.set_profiling(True)
    dev.solver.set_profiling(True)
    for k in range(steps):
        host.step()
        dev.step()
        hb, db = host.solver.bodies(), dev.solver.bodies()
        if hb.tobytes() != db.tobytes() or host.contacts.tobytes() != dev.contacts.tobytes():
            bad = [f for f in hb.dtype.names if hb[f].tobytes() != db[f].tobytes()]
            pytest.fail(f"{name}: step {k} differs (body fields {bad}, contacts "
                        f"{'equal' if host.contacts.tobytes() == dev.contacts.tobytes() else 'differ'})")
    # the host callback ran (family 3 waits for its pairs); the device refresh ran no callback
    host_refreshes, _, _ = host.solver.profile(3)
    dev_waits, _, _ = dev.solver.profile(3)
    dev_refreshes, _, dev_pairs = dev.solver.profile(4)
    print(f"{name}: {steps} steps, {host_refreshes} refreshes on the host, {dev_refreshes} on the device "
          f"({dev_pairs:.0f} pairs)", flush=True)
    assert dev_waits == 0 and dev_refreshes == host_refreshes
    if name != "bounce":
        assert dev_refreshes > 0
    host.solver.close()
    dev.solver.close()


def test_shapes_the_device_does_not_collide_are_refused():
    with pytest.raises(SolverError):
        Simulation(scenes.planar_pile_small(), device_refresh=True)   # prisms and cylinders
    sim = Simulation(scenes.bounce())
    bad = np.zeros(len(sim.bodies), SHAPE_DTYPE)
    bad["kind"] = 2   # a cylinder
    with pytest.raises(SolverError):
        sim.solver.set_body_shapes(bad)
    bad["kind"] = 1   # a box with no size
    with pytest.raises(SolverError):
        sim.solver.set_body_shapes(bad)
    # a shape count that is not the body count fails the step, before it changes anything
    good = sim.world.body_shapes()
    sim.solver.set_body_shapes(np.concatenate([good, good[:1]]))
    sim.collide()
    with pytest.raises(SolverError):
        sim.solve()
    sim.solver.close()
