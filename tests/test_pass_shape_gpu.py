"""The pass shape (grid blocks, tail width, axes loaded or made), its autotuning and the launch of the
grid-barrier passes (plain or cooperative): every shape and launch gives the same bodies, contacts and
lists bit for bit, and tuning runs each candidate in turn, then keeps one of them. Scenes: an 8x8x8 box/sphere pile in walls over its first 180 steps (the drop and the
landing: many colours, refreshes, impacts) and the tumble scene over 120 (fast pairs, restitution)."""
import numpy as np
import pytest

from twofold_amd import scenes
from twofold_amd.world import Simulation

pytestmark = pytest.mark.gpu

def run(scene, shape=None, tune_rounds=0, cooperative=False, device_refresh=False, graph=False):
    steps = 180 if scene.name.startswith("pile") else 120
    sim = Simulation(scene, device_refresh=device_refresh)
    sim.solver.set_cooperative(cooperative)
    sim.solver.set_graph(graph)
    if shape is not None:
        sim.solver.set_pass_shape(*shape)
    seen = []
    if tune_rounds:
        sim.solver.autotune(tune_rounds)
    colours = 0
    for _ in range(steps):
        sim.step()
        seen.append(sim.solver.pass_shape())
        colours = max(colours, sim.solver.counts()["colors"])
    assert colours >= 4 and len(sim.contacts) >= 20, (colours, len(sim.contacts))
    out = (sim.solver.bodies().tobytes(), sim.solver.contacts().tobytes(), sim.solver.lists()["colors"].tobytes())
    sim.solver.close()
    return out, seen


@pytest.fixture(scope="module", params=["pile", "tumble"])
def scene(request):
    return scenes.pile(side=8) if request.param == "pile" else scenes.tumble()


@pytest.fixture(scope="module")
def default(scene):
    return run(scene)


def test_every_shape_gives_the_same_bits(scene, default):
    for shape in [(64, 0, False), (128, 128, True), (96, 64, True), (32, 256, False)]:
        print(f"{scene.name}: shape {shape}", flush=True)
        got, seen = run(scene, shape)
        assert seen[-1]["blocks"] <= shape[0] and seen[-1]["tail_rows"] == shape[1] and seen[-1]["made_axes"] == shape[2]
        assert got == default[0], shape


def test_cooperative_launch_gives_the_same_bits(scene, default):
    got, _ = run(scene, cooperative=True)
    assert got == default[0]


def test_autotune_tries_every_candidate_then_keeps_one(scene, default):
    got, seen = run(scene, tune_rounds=2)
    c = seen[0]["candidates"]
    assert c >= 6
    left = [2 * c] + [s["tuning_steps_left"] for s in seen]
    assert all(b in (a, a - 1) for a, b in zip(left, left[1:]))   # one a step, none for a step with no pass
    assert left[-1] == 0 and sum(a != b for a, b in zip(left, left[1:])) == 2 * c
    final = seen[-1]
    assert final["blocks"] in (64, 128, 256) and final["tail_rows"] in (0, 64, 128)
    assert got == default[0]


def test_fused_substeps_and_graphs_give_the_same_bits(scene):
    """The fused substep launch (one persistent kernel per substep) and the substep loop captured in a
    HIP graph, launched once per step with the eager steps' synchronisations between them, against the
    separate launches, step for step to the end of the run: with the refreshes on the device, so that
    no host callback keeps the loop out of the graph."""
    base, _ = run(scene, device_refresh=True)
    for shape, graph in (((128, 64, False, True), False), ((64, 64, True, True), False), ((128, 64, False, False), True),
                         ((128, 64, False, True), True)):
        print(f"{scene.name}: shape {shape} graph {graph}", flush=True)
        got, seen = run(scene, shape, device_refresh=True, graph=graph)
        assert seen[-1]["fused"] == shape[3] and seen[-1]["graph"] == graph
        assert got == base, (shape, graph)
