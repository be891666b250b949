"""The host world that feeds the solver on the GPU box (twofold_amd/csrc/scene.cpp): mass properties
and manifolds of simple configurations with closed-form answers, and its contacts against the
reference narrowphase's on the same poses."""
import numpy as np
import pytest

from twofold_amd import _abi, scenes
from twofold_amd.world import HostWorld
import refworld


def single(kind_a, dims_a, pos_a, kind_b, dims_b, pos_b, rot_b=scenes.IDENTITY, static_a=False):
    s = scenes.Scene("pair")
    if kind_a == scenes.HALF_SPACE:
        s.ground(0.5)
    else:
        a = s.body(_abi.STATIC if static_a else _abi.DYNAMIC, pos=pos_a)
        s.shape(a, kind_a, dims_a, density=1000.0)
    b = s.body(_abi.DYNAMIC, pos=pos_b, rot=rot_b)
    s.shape(b, kind_b, dims_b, density=1000.0)
    return s


def contacts_of(scene):
    w = HostWorld(scene)
    b = w.bodies()
    c, _ = w.collide(b, scene.dt)
    return b, c


def test_mass_properties():
    s = scenes.Scene("m")
    a = s.body(_abi.DYNAMIC)
    s.shape(a, scenes.BOX, (0.5, 1.0, 1.5), density=2.0)
    b = s.body(_abi.DYNAMIC, pos=(5, 0, 0))
    s.shape(b, scenes.SPHERE, (0.5,), density=1000.0)
    bodies = HostWorld(s).bodies()
    m = 2.0 * 1 * 2 * 3
    np.testing.assert_allclose(1 / bodies[0]["inv_mass"], m, rtol=1e-6)
    np.testing.assert_allclose(bodies[0]["inertia"][[0, 4, 8]], [m * (4 + 9) / 12, m * (1 + 9) / 12, m * (1 + 4) / 12],
                               rtol=1e-6)
    ms = 1000.0 * 4 / 3 * np.pi * 0.125
    np.testing.assert_allclose(1 / bodies[1]["inv_mass"], ms, rtol=1e-5)
    np.testing.assert_allclose(bodies[1]["inertia"][0], 0.4 * ms * 0.25, rtol=1e-5)


def test_box_resting_on_ground_has_four_points_at_zero():
    _, c = contacts_of(single(scenes.HALF_SPACE, (), None, scenes.BOX, (0.5, 0.5, 0.5), (0, 0.5, 0)))
    assert len(c) == 1 and c[0]["count"] == 4
    np.testing.assert_allclose(c[0]["normal"], [0, 1, 0], atol=1e-6)
    np.testing.assert_allclose(c[0]["points"]["separation"][:4], 0, atol=1e-6)
    np.testing.assert_allclose(np.abs(c[0]["points"]["anchor_b"][:4, [0, 2]]), 0.5, atol=1e-6)


def test_sphere_pairs_closed_form():
    _, c = contacts_of(single(scenes.SPHERE, (0.5,), (0, 1, 0), scenes.SPHERE, (0.25,), (0, 1.74, 0)))
    assert len(c) == 1 and c[0]["count"] == 1
    np.testing.assert_allclose(c[0]["points"]["separation"][0], -0.01, atol=1e-6)
    np.testing.assert_allclose(c[0]["normal"], [0, 1, 0], atol=1e-6)
    _, c = contacts_of(single(scenes.BOX, (0.5, 0.5, 0.5), (0, 1, 0), scenes.SPHERE, (0.25,), (0.76, 1.2, 0),
                              static_a=True))
    assert len(c) == 1 and c[0]["count"] == 1
    np.testing.assert_allclose(c[0]["normal"], [1, 0, 0], atol=1e-6)
    np.testing.assert_allclose(c[0]["points"]["separation"][0], 0.01, atol=1e-6)


def test_stacked_boxes_face_contact():
    rot = scenes.axis_angle((0, 1, 0), 0.3)
    _, c = contacts_of(single(scenes.BOX, (0.5, 0.5, 0.5), (0, 0.5, 0), scenes.BOX, (0.4, 0.3, 0.4), (0.1, 1.295, 0),
                              rot_b=rot, static_a=True))
    assert len(c) == 1 and c[0]["count"] == 4
    np.testing.assert_allclose(c[0]["normal"], [0, 1, 0], atol=1e-5)
    np.testing.assert_allclose(c[0]["points"]["separation"][:4], -0.005, atol=1e-5)


@pytest.mark.skipif(not refworld.ref_available(), reason="reference checker not built")
def test_manifolds_match_reference_shape_on_a_pile():
    """Same bodies at the same poses: the host world finds the same pairs in contact as the
    reference narrowphase, with normals within a few degrees and point counts mostly equal."""
    scene = scenes.pile(side=6, half_room=4.0, steps=60)
    w = refworld.RefWorld(scene)
    for _ in range(59):
        w.step(record=False)
    rec = w.step()
    ref = rec.pre_contacts
    host = HostWorld(scene)
    bodies = rec.pre_bodies.copy()
    mine, _ = host.collide(bodies, scene.dt)
    rk = {int(k): c for k, c in zip(ref["key"], ref)}
    mk = {int(k): c for k, c in zip(mine["key"], mine)}
    touching = [k for k, c in rk.items() if c["count"] and (c["points"]["separation"][:c["count"]] < 0).any()]
    assert len(touching) > 50
    found = [k for k in touching if k in mk]
    assert len(found) >= 0.97 * len(touching)
    cos = [float(np.dot(rk[k]["normal"], mk[k]["normal"])) for k in found]
    assert np.median(cos) > 0.999
    same = np.mean([rk[k]["count"] == mk[k]["count"] for k in found])
    assert same > 0.7


def test_groups_keep_worlds_apart():
    """Two worlds in one host world (same coordinates, separate collision groups) give each world's
    own contacts, with body indices and keys shifted by the first world's bodies."""
    worlds = [scenes.stack(6), scenes.hub(3)]
    alone = []
    for sc in worlds:
        w = HostWorld(sc)
        alone.append(w.collide(w.bodies(), sc.dt)[0])
    both = HostWorld(worlds)
    got, _ = both.collide(both.bodies(), worlds[0].dt)
    off = both.body_start[1]
    assert len(got) == len(alone[0]) + len(alone[1]) and len(alone[1]) > 0
    second = alone[1].copy()
    second["body_a"] += off
    second["body_b"] += off
    second["key"] += (np.uint64(off) << np.uint64(32)) + np.uint64(off)
    want = np.concatenate([alone[0], second])
    for f in want.dtype.names:  # field by field: the struct's trailing padding is not data
        assert np.ascontiguousarray(want[f]).tobytes() == np.ascontiguousarray(got[f]).tobytes(), f
