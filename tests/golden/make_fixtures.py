"""Records golden fixtures from the reference engine built from its own sources
(oracle/_ref/libtwofold_ref.so, made by `make -C oracle ref`): for each small scene, the solve
block's input and output on a few chosen steps - the first, the first with a mid-step refresh, the
first with an impact pass, the first where a body falls asleep, and one late step.

Run in the build container: python tests/golden/make_fixtures.py
"""
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from twofold_amd import scenes  # noqa: E402
import golden_io  # noqa: E402
import refworld  # noqa: E402

SCENES = [scenes.pyramid(), scenes.stack(), scenes.bounce(), scenes.rolling(), scenes.planar_mix(), scenes.tumble()]


def pick(scene):
    w = refworld.RefWorld(scene)
    chosen = {}
    late = scene.steps * 2 // 3
    for k in range(scene.steps):
        rec = w.step()
        if rec is None:
            continue
        tags = []
        if not chosen:
            tags.append("first")
        if rec.refreshes and "refresh" not in chosen.values():
            tags.append("refresh")
        if len(rec.lists["impact"]) and "impact" not in chosen.values():
            tags.append("impact")
        slept = (rec.pre_bodies["awake"] == 1) & (rec.step_bodies["awake"] == 0)
        if slept.any() and "sleep" not in chosen.values():
            tags.append("sleep")
        if k >= late and "late" not in chosen.values() and len(rec.lists["active"]):
            tags.append("late")
        if tags:
            chosen[k] = tags[0]
            golden_io.save(os.path.join(HERE, f"{scene.name}-step{k:04d}.npz"), scene.name, k, rec, w.settings)
            print(scene.name, k, tags, len(rec.pre_contacts), "contacts", flush=True)


if __name__ == "__main__":
    for old in os.listdir(HERE):
        if old.endswith(".npz"):
            os.remove(os.path.join(HERE, old))
    for s in SCENES:
        pick(s)
