"""The joint recorder of the reference checker (oracle/ref_capture.cpp): the joints and the world body's
velocities it records at the solve block's end are the world's after the step (nothing after the solve
block changes them, src/world.cpp:1692-1730), each step starts from the last one's, and the rows do
work (impulses, reactions) on the joint scenes the GPU tests use."""
import numpy as np
import pytest

from twofold_amd import scenes
import refworld

pytestmark = pytest.mark.skipif(not refworld.ref_available(), reason="reference checker not built (make -C oracle ref)")


@pytest.mark.parametrize("make", [scenes.chain, scenes.joint_mix], ids=lambda f: f.__name__)
def test_recorded_joints_are_the_worlds(make):
    scene = make()
    w = refworld.RefWorld(scene)
    last = None
    worked = 0
    for k in range(60):
        rec = w.step()
        assert rec is not None
        assert rec.post_joints.tobytes() == w.joints().tobytes(), k
        if last is not None:
            carried = ("linear_impulse", "angular_impulse", "axial_impulse", "lower_impulse", "upper_impulse",
                       "motor_impulse", "friction_impulse", "broken")
            for f in carried:
                assert rec.pre_joints[f].tobytes() == last[f].tobytes(), (k, f)
        worked += int(np.any(rec.post_joints["force"] != 0))
        last = rec.post_joints.copy()
    assert worked == 60
    assert len(w.joints()) == len(scene.joints)
