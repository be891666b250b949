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


def planar_single(kind_a, dims_a, pos_a, kind_b, dims_b, pos_b, static_a=False):
    """Two bodies locked to the xy plane (or a static first one), identity rotation."""
    s = scenes.Scene("planar-pair")
    a = s.body(_abi.STATIC if static_a else _abi.DYNAMIC, pos=pos_a, locks=0 if static_a else _abi.LOCK_PLANAR)
    s.shape(a, kind_a, dims_a, density=1000.0)
    b = s.body(_abi.DYNAMIC, pos=pos_b, locks=_abi.LOCK_PLANAR)
    s.shape(b, kind_b, dims_b, density=1000.0)
    return s


SQUARE = [-0.5, 0.5, 0.5, -0.5, 0.5, 0.5, -0.5, 0.5, -0.5, -0.5]   # zMin, zMax, then x0, y0, ... (clockwise)


def test_cylinder_and_prism_mass_properties():
    s = scenes.Scene("m")
    a = s.body(_abi.DYNAMIC)
    s.shape(a, scenes.CYLINDER, (0.5, 0.25), density=1000.0)
    b = s.body(_abi.DYNAMIC, pos=(5, 0, 0))
    s.shape(b, scenes.PRISM, SQUARE, density=1000.0)
    c = s.body(_abi.DYNAMIC, pos=(-5, 0, 0))
    s.shape(c, scenes.BOX, (0.5, 0.5, 0.5), density=1000.0)
    bodies = HostWorld(s).bodies()
    m = 1000.0 * np.pi * 0.25 * 0.5
    np.testing.assert_allclose(1 / bodies[0]["inv_mass"], m, rtol=1e-5)
    np.testing.assert_allclose(bodies[0]["inertia"][[0, 4, 8]], [m * (3 * 0.25 + 0.25) / 12, m * (3 * 0.25 + 0.25) / 12,
                                                                 m * 0.25 / 2], rtol=1e-5)
    # a unit square prism (given clockwise) weighs and turns as the unit box
    np.testing.assert_allclose(1 / bodies[1]["inv_mass"], 1 / bodies[2]["inv_mass"], rtol=1e-6)
    np.testing.assert_allclose(bodies[1]["inertia"], bodies[2]["inertia"], rtol=1e-5, atol=1e-6)


def test_cylinder_and_prism_on_the_ground():
    _, c = contacts_of(single(scenes.HALF_SPACE, (), None, scenes.CYLINDER, (0.3, 0.5), (0, 0.3, 0)))
    assert len(c) == 1 and c[0]["count"] == 2
    np.testing.assert_allclose(c[0]["normal"], [0, 1, 0], atol=1e-6)
    np.testing.assert_allclose(c[0]["points"]["separation"][:2], 0, atol=1e-6)
    np.testing.assert_allclose(sorted(c[0]["points"]["anchor_b"][:2, 2]), [-0.5, 0.5], atol=1e-6)
    np.testing.assert_allclose(c[0]["points"]["anchor_b"][:2, 1], -0.3, atol=1e-6)
    _, c = contacts_of(single(scenes.HALF_SPACE, (), None, scenes.PRISM, SQUARE, (0, 0.49, 0)))
    assert len(c) == 1 and c[0]["count"] == 4
    np.testing.assert_allclose(c[0]["points"]["separation"][:4], -0.01, atol=1e-6)


def test_extruded_pairs_closed_form():
    cases = [(scenes.PRISM, SQUARE, scenes.PRISM, SQUARE, 0.99, 4, -0.01),
             (scenes.PRISM, SQUARE, scenes.CYLINDER, (0.5, 0.5), 0.99, 2, -0.01),
             (scenes.CYLINDER, (0.5, 0.5), scenes.PRISM, SQUARE, 0.99, 2, -0.01),
             (scenes.CYLINDER, (0.5, 0.5), scenes.CYLINDER, (0.25, 0.5), 0.76, 2, 0.01),
             (scenes.BOX, (0.5, 0.5, 0.5), scenes.CYLINDER, (0.5, 0.5), 1.0, 2, 0.0)]
    for ka, da, kb, db, x, count, sep in cases:
        _, c = contacts_of(planar_single(ka, da, (0, 1, 0), kb, db, (x, 1, 0)))
        assert len(c) == 1 and c[0]["count"] == count, (ka, kb)
        np.testing.assert_allclose(c[0]["normal"], [1, 0, 0], atol=1e-6)
        np.testing.assert_allclose(c[0]["points"]["separation"][:count], sep, atol=1e-6)
        np.testing.assert_allclose(sorted(set(np.round(c[0]["points"]["anchor_a"][:count, 2], 6))), [-0.5, 0.5],
                                   atol=1e-6)


def test_free_upright_boxes_meet_along_z():
    """Axis-aligned boxes that are not locked to the plane collide in 3D: one behind the other in z
    gives a face contact along z, not a 2D one."""
    s = scenes.Scene("z-pair")
    for z in (0.0, 0.99):
        b = s.body(_abi.DYNAMIC, pos=(0, 1, z))
        s.shape(b, scenes.BOX, (0.5, 0.5, 0.5), density=1000.0)
    _, c = contacts_of(s)
    assert len(c) == 1 and c[0]["count"] == 4
    np.testing.assert_allclose(c[0]["normal"], [0, 0, 1], atol=1e-6)
    np.testing.assert_allclose(c[0]["points"]["separation"][:4], -0.01, atol=1e-6)


def test_sphere_against_prism_and_cylinder():
    for kind, dims in ((scenes.PRISM, SQUARE), (scenes.CYLINDER, (0.5, 0.5))):
        _, c = contacts_of(single(kind, dims, (0, 1, 0), scenes.SPHERE, (0.25,), (0, 1.74, 0.2), static_a=True))
        assert len(c) == 1 and c[0]["count"] == 1
        np.testing.assert_allclose(c[0]["normal"], [0, 1, 0], atol=1e-6)
        np.testing.assert_allclose(c[0]["points"]["separation"][0], -0.01, atol=1e-6)
        # against the end cap
        _, c = contacts_of(single(kind, dims, (0, 1, 0), scenes.SPHERE, (0.25,), (0.1, 1.0, 0.76), static_a=True))
        assert len(c) == 1
        np.testing.assert_allclose(c[0]["normal"], [0, 0, 1], atol=1e-6)
        np.testing.assert_allclose(c[0]["points"]["separation"][0], 0.01, atol=1e-6)


@pytest.mark.skipif(not refworld.ref_available(), reason="reference checker not built")
def test_planar_mix_matches_reference_shape():
    """The planar workload's mix (boxes, prisms, cylinders, free spheres) at the reference's poses after
    it has dropped: the same touching pairs, with normals within a few degrees."""
    scene = scenes.planar_pile(side=8, spheres=10, steps=120)
    w = refworld.RefWorld(scene)
    for _ in range(119):
        w.step(record=False)
    rec = w.step()
    ref = rec.pre_contacts
    mine, _ = HostWorld(scene).collide(rec.pre_bodies.copy(), scene.dt)
    rk = {int(k): c for k, c in zip(ref["key"], ref)}
    mk = {int(k): c for k, c in zip(mine["key"], mine)}
    touching = [k for k, c in rk.items() if c["count"] and (c["points"]["separation"][:c["count"]] < -1e-3).any()]
    assert len(touching) > 40
    found = [k for k in touching if k in mk]
    assert len(found) >= 0.95 * len(touching), (len(found), len(touching))
    cos = [float(np.dot(rk[k]["normal"], mk[k]["normal"])) for k in found]
    assert np.median(cos) > 0.999


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
