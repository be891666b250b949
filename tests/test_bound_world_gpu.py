"""The drop-in, compiled: the reference engine built with the binding (oracle/ref_bound.cpp), whose
World::step runs its solve block and its sleep pass on the GPU through the C ABI (tfs_step with the
sleep pass off, then tfs_sleep where the reference calls updateSleep, after the continuous pass and the
sensors) and its mid-step refresh through the refresh callback with the reference's own narrowphase,
stepped side by side with the unmodified reference from the same scene. After every step the two
worlds' bodies, contacts and joints (every field the ABI carries), and the step's hit events, are equal
bit for bit: the binding changes nothing a user of World can observe."""
import os

import numpy as np
import pytest

from twofold_amd import scenes
import refworld

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not os.path.exists(refworld.BOUND_SO),
                                 reason="the bound reference is not built (make -C oracle bound)")]

SCENES = {
    "pyramid": lambda: scenes.pyramid(),
    "pile8": lambda: scenes.pile(side=8, half_room=6.0, steps=200),
    "tumble": lambda: scenes.tumble(),
    "edges": lambda: scenes.edges(),
    "planar-mix": lambda: scenes.planar_mix(),
    "rolling": lambda: scenes.rolling(),
    "chain": lambda: scenes.chain(),
    "joint-mix": lambda: scenes.joint_mix(),
}


@pytest.mark.parametrize("name", list(SCENES))
def test_bound_world_steps_as_the_reference(name):
    scene = SCENES[name]()
    plain = refworld.RefWorld(scene)
    bound = refworld.RefWorld(scene, lib_path=refworld.BOUND_SO)
    try:
        asleep = 0
        for k in range(scene.steps):
            plain.step(record=False)
            bound.step(record=False)
            err = bound.lib.tfb_error()
            assert not err, f"{name} step {k}: {err.decode()}"
            pb, bb = plain.bodies(), bound.bodies()
            if pb.tobytes() != bb.tobytes():
                bad = [f for f in pb.dtype.names if pb[f].tobytes() != bb[f].tobytes()]
                pytest.fail(f"{name} step {k}: body fields {bad} differ")
            pc, bc = plain.contacts(), bound.contacts()
            if pc.tobytes() != bc.tobytes():
                bad = [f for f in pc.dtype.names if pc[f].tobytes() != bc[f].tobytes()]
                pytest.fail(f"{name} step {k}: contact fields {bad} differ ({len(pc)} / {len(bc)} contacts)")
            pj, bj = plain.joints(), bound.joints()
            if pj.tobytes() != bj.tobytes():
                bad = [f for f in pj.dtype.names if pj[f].tobytes() != bj[f].tobytes()]
                pytest.fail(f"{name} step {k}: joint fields {bad} differ")
            ph, bh = plain.hits(), bound.hits()
            assert ph.tobytes() == bh.tobytes(), f"{name} step {k}: hit events differ"
            asleep = int(np.sum((pb["type"] == 2) & (pb["awake"] == 0)))
        print(f"{name}: {scene.steps} steps through World::step, bodies, contacts and joints equal every step; "
              f"{asleep} bodies asleep at the end", flush=True)
    finally:
        del plain, bound
        refworld._load_ref(refworld.BOUND_SO).tfb_release()
