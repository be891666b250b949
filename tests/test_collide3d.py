"""The restated box / sphere / half-space narrowphase (twofold_amd/csrc/collide3d.h, host build in
libtwofold_scene.so) against the reference's own `collide` (src/collide3d.cpp:1778-1811, built from its
sources into oracle/_ref), bit for bit, on seeded pairs chosen to reach every branch: resting and
tilted boxes on half-spaces (more than four points within the margin, so reduceManifold picks four),
box faces (A or B as the reference face), edge pairs (clipped, and the one-point closest-segment case),
separated boxes whose queries are unsure (the GJK gap test), spheres inside and outside boxes (face,
edge and corner regions), sphere pairs, and fast pairs (Motion::fast)."""
import ctypes
import math
import os

import numpy as np
import pytest

import refworld

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SCENE_SO = os.path.join(ROOT, "twofold_amd", "libtwofold_scene.so")

_fp = ctypes.POINTER(ctypes.c_float)
SPHERE, BOX, HALF = 0, 1, 4


def _libs():
    if not refworld.ref_available():
        pytest.skip("the reference build (oracle/_ref) is not present")
    ref = ctypes.CDLL(refworld.REF_SO)
    ref.tfr_collide.restype = ctypes.c_int
    ref.tfr_collide.argtypes = [ctypes.c_int, _fp, _fp, ctypes.c_int, _fp, _fp, ctypes.c_float, ctypes.c_float, _fp,
                                ctypes.c_float, ctypes.c_int, _fp, _fp]
    ours = ctypes.CDLL(SCENE_SO)
    ours.tfw_collide3d.restype = ctypes.c_int
    ours.tfw_collide3d.argtypes = [ctypes.c_int, _fp, _fp, ctypes.c_int, _fp, _fp, ctypes.c_float, ctypes.c_float,
                                   _fp, ctypes.c_float, ctypes.c_int, _fp]
    return ref, ours


def _p(a):
    return a.ctypes.data_as(_fp)


def _quat(axis, angle):
    axis = np.asarray(axis, np.float64)
    n = np.linalg.norm(axis)
    if n == 0:
        return np.array([0, 0, 0, 1], np.float32)
    s = math.sin(0.5 * angle) / n
    return np.array([axis[0] * s, axis[1] * s, axis[2] * s, math.cos(0.5 * angle)], np.float32)


def _rot(q, v):
    q = q.astype(np.float64)
    u = q[:3]
    t = 2.0 * np.cross(u, v)
    return v + q[3] * t + np.cross(u, t)


def _cases(rng, n):
    """(kind A, dims A, pose A, kind B, dims B, pose B, margin, motion, dt, fast) for n pairs."""
    ground_q = np.array([-math.sqrt(0.5), 0, 0, math.sqrt(0.5)], np.float32)  # local z to world y
    out = []
    for i in range(n):
        family = i % 9
        half_a = rng.uniform(0.2, 0.6, 3).astype(np.float32)
        half_b = rng.uniform(0.2, 0.6, 3).astype(np.float32)
        qa = _quat(rng.normal(size=3), rng.uniform(0, math.pi))
        qb = _quat(rng.normal(size=3), rng.uniform(0, math.pi))
        if family == 0:  # box on the ground, tilted a little or a lot
            tilt = rng.choice([0.0, 1e-3, 0.02, 0.3, rng.uniform(0, math.pi)])
            qb = _quat(rng.normal(size=3), tilt)
            pb = np.array([rng.uniform(-3, 3), 0, rng.uniform(-3, 3)], np.float32)
            low = min(_rot(qb, (np.array(c) * 2 - 1) * half_b)[1] for c in np.ndindex(2, 2, 2))
            pb[1] = -low + rng.uniform(-0.05, 0.03)
            case = (HALF, np.zeros(3, np.float32), np.zeros(3, np.float32), ground_q, BOX, half_b, pb, qb)
        elif family == 1:  # sphere on the ground, or the ground as shape B
            r = np.float32(rng.uniform(0.2, 0.5))
            pb = np.array([rng.uniform(-3, 3), r + rng.uniform(-0.05, 0.03), rng.uniform(-3, 3)], np.float32)
            if rng.uniform() < 0.5:
                case = (HALF, np.zeros(3, np.float32), np.zeros(3, np.float32), ground_q, SPHERE,
                        np.array([r, 0, 0], np.float32), pb, qb)
            else:
                case = (SPHERE, np.array([r, 0, 0], np.float32), pb, qb, HALF, np.zeros(3, np.float32),
                        np.zeros(3, np.float32), ground_q)
        elif family in (2, 3):  # two boxes near touching, any orientations (faces, edges, unsure gaps)
            d = rng.normal(size=3)
            d /= np.linalg.norm(d)
            reach = np.linalg.norm(half_a) + np.linalg.norm(half_b)
            pb = (d * rng.uniform(0.45, 1.0) * reach).astype(np.float32)
            case = (BOX, half_a, np.zeros(3, np.float32), qa, BOX, half_b, pb, qb)
        elif family == 4:  # a box stacked on a box, nearly aligned (faces of A or of B)
            qa = _quat(rng.normal(size=3), rng.choice([0.0, 1e-4, 0.01, 0.05]))
            qb = _quat(rng.normal(size=3), rng.choice([0.0, 1e-4, 0.01, 0.05]))
            pb = np.array([rng.uniform(-0.3, 0.3), half_a[1] + half_b[1] + rng.uniform(-0.03, 0.03),
                           rng.uniform(-0.3, 0.3)], np.float32)
            case = (BOX, half_a, np.zeros(3, np.float32), qa, BOX, half_b, pb, qb)
        elif family == 5:  # crossed edges: boxes turned 45 degrees about two axes
            qa = _quat([1, 0, 0], math.pi / 4 + rng.uniform(-0.05, 0.05))
            qb = _quat([0, 0, 1], math.pi / 4 + rng.uniform(-0.05, 0.05))
            ea = _rot(qa, np.array([0, half_a[1], half_a[2]])).max()
            eb = _rot(qb, np.array([half_b[0], half_b[1], 0])).max()
            pb = np.array([rng.uniform(-0.05, 0.05), ea + eb + rng.uniform(-0.04, 0.03), rng.uniform(-0.05, 0.05)],
                          np.float32)
            case = (BOX, half_a, np.zeros(3, np.float32), qa, BOX, half_b, pb, qb)
        elif family == 6:  # sphere and box: centre inside, near a face, an edge or a corner
            r = np.float32(rng.uniform(0.2, 0.5))
            corner = rng.integers(-1, 2, 3).astype(np.float32)
            local = corner * half_a + rng.normal(size=3) * 0.1 * (1 + (corner != 0) * r * 3)
            pb = _rot(qa, local).astype(np.float32)
            sphere = (SPHERE, np.array([r, 0, 0], np.float32), pb, qb)
            box = (BOX, half_a, np.zeros(3, np.float32), qa)
            case = box + sphere if rng.uniform() < 0.5 else sphere + box
        elif family == 8:  # boxes corner to corner or edge to edge across small gaps on two or three axes:
            # the closed-form queries are unsure and the GJK gap test decides
            qa = _quat(rng.normal(size=3), rng.choice([0.0, 1e-3, 0.03, 0.1]))
            qb = _quat(rng.normal(size=3), rng.choice([0.0, 1e-3, 0.03, 0.1]))
            sign = rng.choice([-1.0, 1.0], 3)
            gaps = rng.uniform(-0.01, 0.08, 3)
            gaps[rng.integers(0, 3)] *= rng.choice([0.0, 1.0, -5.0])
            pb = (sign * (half_a + half_b + gaps)).astype(np.float32)
            case = (BOX, half_a, np.zeros(3, np.float32), qa, BOX, half_b, pb, qb)
        else:  # two spheres
            ra, rb = np.float32(rng.uniform(0.2, 0.5)), np.float32(rng.uniform(0.2, 0.5))
            d = rng.normal(size=3)
            pb = (d / np.linalg.norm(d) * (ra + rb + rng.uniform(-0.1, 0.05))).astype(np.float32)
            if rng.uniform() < 0.05:
                pb = np.zeros(3, np.float32)
            case = (SPHERE, np.array([ra, 0, 0], np.float32), np.zeros(3, np.float32), qa, SPHERE,
                    np.array([rb, 0, 0], np.float32), pb, qb)
        ka, da, pa, qa_, kb, db, pb_, qb_ = case
        fast = rng.uniform() < 0.3
        speed = 8.0 if fast else 0.5
        motion = np.concatenate([rng.normal(size=3) * speed, rng.normal(size=3), pa, rng.normal(size=3) * speed,
                                 rng.normal(size=3), pb_]).astype(np.float32)
        dt = np.float32(rng.choice([1 / 60, 0.75 / 60, 0.5 / 60, 0.25 / 60]))
        margin = np.float32(0.02 + (rng.uniform(0, 0.2) if fast else 0.0))
        out.append((ka, da, np.concatenate([pa, qa_]).astype(np.float32), kb, db,
                    np.concatenate([pb_, qb_]).astype(np.float32), margin, motion, dt, fast))
    return out


def test_restated_narrowphase_is_the_references_bit_for_bit():
    ref, ours = _libs()
    rng = np.random.default_rng(20261204)
    solid = np.zeros(8, np.float32)
    want = np.zeros(24, np.float32)
    got = np.zeros(24, np.float32)
    counts = np.zeros(5, int)
    for case in _cases(rng, 27000):
        ka, da, pa, kb, db, pb, margin, motion, dt, fast = case
        want[:] = 0
        assert ref.tfr_collide(ka, _p(da), _p(pa), kb, _p(db), _p(pb), margin, np.float32(0.005), _p(motion), dt,
                               int(fast), _p(solid), _p(want)) == 0
        # the restatement takes the sizes the reference keeps after buildShape
        sa = solid[:3].copy() if ka == BOX else np.array([solid[3], 0, 0], np.float32)
        sb = solid[4:7].copy() if kb == BOX else np.array([solid[7], 0, 0], np.float32)
        got[:] = 0
        assert ours.tfw_collide3d(ka, _p(sa), _p(pa), kb, _p(sb), _p(pb), margin, np.float32(0.005), _p(motion), dt,
                                  int(fast), _p(got)) == 0
        n = int(want[3])
        counts[n] += 1
        used = 4 + 5 * n
        if want[:4].view(np.uint32).tolist() != got[:4].view(np.uint32).tolist() or \
                want[4:used].view(np.uint32).tolist() != got[4:used].view(np.uint32).tolist():
            raise AssertionError(f"case {case} differs:\nreference {want[:used]}\nrestated  {got[:used]}")
    # every manifold size appears, so every branch that makes one was reached
    assert counts[0] > 100 and counts[1] > 100 and counts[2] > 100 and counts[4] > 100, counts


def test_box_polyhedron_tables_match_the_reference_box():
    """The restated box's face loops and edges give the reference's manifolds for a box resting flat
    on the ground on each of its six faces (every face as the incident face)."""
    ref, ours = _libs()
    solid = np.zeros(8, np.float32)
    want = np.zeros(24, np.float32)
    got = np.zeros(24, np.float32)
    half = np.array([0.3, 0.4, 0.5], np.float32)
    ground = np.array([0, 0, 0, -math.sqrt(0.5), 0, 0, math.sqrt(0.5)], np.float32)
    motion = np.zeros(18, np.float32)
    for axis, angle in (([1, 0, 0], 0), ([1, 0, 0], math.pi / 2), ([1, 0, 0], math.pi), ([1, 0, 0], -math.pi / 2),
                        ([0, 0, 1], math.pi / 2), ([0, 0, 1], -math.pi / 2)):
        q = _quat(axis, angle)
        low = min(_rot(q, (np.array(c) * 2 - 1) * half)[1] for c in np.ndindex(2, 2, 2))
        pose = np.array([0.1, -low - 0.001, -0.2, *q], np.float32)
        ref.tfr_collide(HALF, _p(np.zeros(3, np.float32)), _p(ground), BOX, _p(half), _p(pose), np.float32(0.02),
                        np.float32(0.005), _p(motion), np.float32(1 / 240), 0, _p(solid), _p(want))
        ours.tfw_collide3d(HALF, _p(np.zeros(3, np.float32)), _p(ground), BOX, _p(solid[4:7].copy()), _p(pose),
                           np.float32(0.02), np.float32(0.005), _p(motion), np.float32(1 / 240), 0, _p(got))
        assert int(want[3]) == 4
        assert want.view(np.uint32).tolist() == got.view(np.uint32).tolist()
