"""Joint rows on the GPU against the live reference (src/joint.cpp, src/world.cpp:1623-1691, 1210-1223):
each step of a reference run with joints is replayed on the GPU from the inputs its solve block saw
(bodies, contacts, joints with their carried impulses, the world body's velocities) and the step's
lists, colours, poses, velocities, inverse masses, islands, sleep flags, contact impulses, and every
joint's impulses, row directions, force, torque and broken flag must equal the reference's bit for bit.
Scenes: two chains of point joints (one held to the world, one hanging from a moving kinematic body)
folding onto a pile, and a scene with a hinged bridge (friction, motor), a fixed joint, a rigid and a
sprung, limited distance joint, a limited slider with motor and friction, and a fixed joint that breaks."""
import numpy as np
import pytest

from twofold_amd import scenes
from twofold_amd.solver import Solver
import parity
import refworld
from test_gpu_parity import recorded

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not refworld.ref_available(), reason="reference checker not built (make -C oracle ref)")]

JOINT_FIELDS = ("linear_impulse", "angular_impulse", "axial_impulse", "motor_impulse", "lower_impulse",
                "upper_impulse", "friction_impulse", "swing_impulse", "friction_turn", "row_force", "row_torque",
                "force", "torque", "broken", "alive")


def compare_joints(rec, joints, world):
    bad = []
    for f in JOINT_FIELDS:
        want, got = rec.post_joints[f], joints[f]
        d = np.nonzero([a.tobytes() != b.tobytes() for a, b in zip(want, got)])[0]
        if len(d):
            bad.append(f"joint {f} at {d[:5].tolist()}")
    if rec.post_world.tobytes() != world.tobytes():
        bad.append("world body velocities")
    return bad


@pytest.mark.parametrize("make", [scenes.chain, scenes.joint_mix], ids=lambda f: f.__name__)
def test_joints_match_reference_every_step(make):
    scene = make()
    w = refworld.RefWorld(scene)
    s = Solver(len(scene.bodies), 1024)
    checked = broken = 0
    try:
        for k in range(scene.steps):
            rec = w.step()
            if rec is None:
                continue
            case = parity.record_dict(rec)
            s.set_settings(w.settings)
            s.upload(case["pre_bodies"], case["pre_contacts"])
            s.upload_joints(rec.pre_joints, rec.pre_world)
            s.step(case["dt"], recorded(case["refreshes"]) if case["refreshes"] else None)
            b, c, lists = s.bodies(), s.contacts(), s.lists()
            joints, world = s.joints()
            bad = parity.compare(case, b, c, lists) + compare_joints(rec, joints, world)
            assert bad == [], f"{scene.name} step {k}: {bad[:6]}"
            checked += 1
            broken = int(joints["broken"].sum())
    finally:
        s.close()
    print(f"{scene.name}: {checked} steps with {len(rec.post_joints)} joints bit-exact; {broken} broken at the end",
          flush=True)
    assert checked == scene.steps
    if make is scenes.joint_mix:
        assert broken == 1


def test_unsupported_joints_are_refused():
    """A limited hinge and a swing-twist joint read angles through atan2: refused before any step."""
    scene = scenes.joint_mix()
    w = refworld.RefWorld(scene)
    j = w.joints()
    s = Solver(len(scene.bodies), 64)
    try:
        s.upload(w.bodies(), w.contacts())
        for kind, limited in ((scenes.JOINT_HINGE, 1), (5, 0)):
            bad = j.copy()
            bad["type"][1] = kind
            bad["limited"][1] = limited
            with pytest.raises(RuntimeError, match="atan2"):
                s.upload_joints(bad)
        s.upload_joints(j)
    finally:
        s.close()


def test_joints_naming_bodies_no_longer_resident_are_refused():
    """Joints uploaded for more bodies than are resident after a later body upload: the step refuses
    them (TFS_E_STATE) before any kernel reads a body out of range."""
    scene = scenes.chain()
    w = refworld.RefWorld(scene)
    bodies, joints = w.bodies(), w.joints()
    s = Solver(len(bodies), 64)
    try:
        s.upload(bodies, w.contacts())
        s.upload_joints(joints)
        s.upload(bodies[:4], w.contacts()[:0])
        with pytest.raises(RuntimeError, match="upload them again"):
            s.step(scene.dt, None)
    finally:
        s.close()


def test_joints_beside_several_worlds_are_refused():
    """Each of the reference's worlds has a world body of its own; joints beside several worlds side by
    side (tfs_set_worlds) are refused at the step (TFS_E_STATE)."""
    scene = scenes.chain()
    w = refworld.RefWorld(scene)
    bodies, joints = w.bodies(), w.joints()
    both = np.concatenate([bodies, bodies])
    s = Solver(len(both), 64)
    try:
        s.upload_worlds(both, np.array([0, len(bodies), len(both)], np.uint32), w.contacts()[:0])
        s.upload_joints(joints)
        with pytest.raises(RuntimeError, match="one world at a time"):
            s.step(scene.dt, None)
    finally:
        s.close()
