"""Scene recipes for the workloads the solver is measured and checked on.

A recipe is plain data: bodies (motion type, pose, velocities, locks) and shapes (kind, dimensions,
material). The same recipe can be loaded into any world that accepts that data. All arithmetic is
float32, in the reference's operation order, so a recipe gives the same bits everywhere.

Random draws use the reference viewer's xorshift32 generator (viewer/scenes.cpp:22-32) with its
seed, drawn in the order each recipe documents.
"""
import ctypes
import ctypes.util
import math

import numpy as np

from ._abi import DYNAMIC, KINEMATIC, LOCK_PLANAR, STATIC  # noqa: F401

f32 = np.float32
_libm = ctypes.CDLL(ctypes.util.find_library("m") or "libm.so.6")
_libm.sinf.restype = ctypes.c_float
_libm.sinf.argtypes = [ctypes.c_float]
_libm.cosf.restype = ctypes.c_float
_libm.cosf.argtypes = [ctypes.c_float]
_libm.sqrtf.restype = ctypes.c_float
_libm.sqrtf.argtypes = [ctypes.c_float]

PI = f32(3.14159265358979)

SPHERE, BOX, CYLINDER, PRISM, HALF_SPACE = 0, 1, 2, 3, 4


class Rng:
    """xorshift32 as viewer/scenes.cpp:22-32."""

    def __init__(self, seed=0x9E3779B9):
        self.s = seed & 0xFFFFFFFF

    def next(self):
        s = self.s
        s ^= (s << 13) & 0xFFFFFFFF
        s ^= s >> 17
        s ^= (s << 5) & 0xFFFFFFFF
        self.s = s
        return f32(s & 0xFFFFFF) / f32(16777216.0)

    def range(self, lo, hi):
        lo, hi = f32(lo), f32(hi)
        return lo + (hi - lo) * self.next()

    def normal(self, sigma=1.0):
        """One N(0, sigma) draw by Box-Muller on two uniform draws."""
        u1 = max(float(self.next()), 1e-7)
        u2 = float(self.next())
        return f32(sigma * math.sqrt(-2.0 * math.log(u1)) * math.cos(2.0 * math.pi * u2))


def normalize3(v):
    v = np.asarray(v, f32)
    length = _libm.sqrtf(f32(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]))
    length = f32(length)
    if length > 0:
        inv = f32(1.0) / length
        return np.array([v[0] * inv, v[1] * inv, v[2] * inv], f32)
    return np.zeros(3, f32)


def axis_angle(axis, angle):
    """Quat::axisAngle (include/twofold/math.h:146-149) in float32 with libm's sinf/cosf."""
    half = f32(0.5) * f32(angle)
    a = normalize3(axis) * f32(_libm.sinf(half))
    return np.array([a[0], a[1], a[2], f32(_libm.cosf(half))], f32)


def from_to(a, b):
    """Quat::fromTo (include/twofold/math.h:152-158) for the unit vectors used here."""
    a = np.asarray(a, f32)
    b = np.asarray(b, f32)
    d = f32(a[0] * b[0] + a[1] * b[1] + a[2] * b[2])
    if d < f32(-0.999999):
        raise ValueError("opposite vectors are not used by these recipes")
    c = np.array([a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]], f32)
    s = f32(_libm.sqrtf(f32((f32(1.0) + d) * f32(2.0))))
    return np.array([c[0] / s, c[1] / s, c[2] / s, f32(0.5) * s], f32)


IDENTITY = np.array([0, 0, 0, 1], f32)
UP = from_to((0, 0, 1), (0, 1, 0))  # the ground's rotation in the reference tests (tests/test_world.cpp:14)


class Scene:
    """Bodies and shapes of one world, in creation order."""

    def __init__(self, name, substeps=4, dt=1.0 / 60.0, steps=1, settings=None):
        self.name = name
        self.substeps = substeps
        self.dt = f32(dt)
        self.steps = steps
        self.bodies = []
        self.shapes = []
        self.joints = []
        # WorldSettings that differ from the defaults (tfs_settings field names)
        self.settings = dict(settings or {})

    def body(self, type=DYNAMIC, pos=(0, 0, 0), rot=IDENTITY, v=(0, 0, 0), w=(0, 0, 0), locks=0,
             lock_frame=IDENTITY, awake=True, linear_damping=0.0, angular_damping=0.0, gravity_scale=1.0,
             force=(0, 0, 0), torque=(0, 0, 0)):
        """A body; `force` and `torque` are applied before every step (World::addForce / addTorque
        without waking, src/world.cpp:2064-2088: they last one step and a sleeping body ignores them)."""
        self.bodies.append(dict(
            type=type, pos=np.asarray(pos, f32), rot=np.asarray(rot, f32), v=np.asarray(v, f32),
            w=np.asarray(w, f32), locks=locks, lock_frame=np.asarray(lock_frame, f32), awake=awake,
            linear_damping=f32(linear_damping), angular_damping=f32(angular_damping),
            gravity_scale=f32(gravity_scale), force=np.asarray(force, f32), torque=np.asarray(torque, f32)))
        return len(self.bodies) - 1

    def forces(self):
        """(body, force, torque) of every body with a force or torque applied each step."""
        return [(i, b["force"], b["torque"]) for i, b in enumerate(self.bodies)
                if b["force"].any() or b["torque"].any()]

    def shape(self, body, kind, dims=(), pos=(0, 0, 0), rot=IDENTITY, density=1.0, friction=0.5,
              restitution=0.0, rolling=0.0):
        self.shapes.append(dict(
            body=body, kind=kind, dims=np.asarray(dims, f32), pos=np.asarray(pos, f32),
            rot=np.asarray(rot, f32), density=f32(density), friction=f32(friction),
            restitution=f32(restitution), rolling=f32(rolling)))
        return len(self.shapes) - 1

    def joint(self, kind, a, b, anchor=(0, 0, 0), anchor_b=(0, 0, 0), axis=(1, 0, 0), limited=False, lower=0.0,
              upper=0.0, motor_speed=0.0, max_motor=0.0, friction=0.0, length=-1.0, hertz=0.0, damping_ratio=0.0,
              break_force=math.inf, break_torque=math.inf, collide_connected=False):
        """A joint as World::createJoint takes it (include/twofold/world.h:135-210): kind JOINT_*, body a
        (None: the world), body b, world-space anchors and axis at the bodies' creation poses; the
        hinge's and slider's limits, motor and friction, the distance joint's length (negative: the
        distance the anchors start at), spring and limits (its min and max distance as lower, upper)."""
        p = np.array([*anchor, *anchor_b, *axis, 1.0 if limited else 0.0, lower, upper, motor_speed, max_motor,
                      friction, length, hertz, damping_ratio, break_force, break_torque,
                      1.0 if collide_connected else 0.0], f32)
        self.joints.append(dict(kind=kind, a=-1 if a is None else a, b=b, p=p))
        return len(self.joints) - 1

    def ground(self, friction=0.6, restitution=0.0, rolling=0.0):
        g = self.body(STATIC, rot=UP)
        self.shape(g, HALF_SPACE, friction=friction, restitution=restitution, rolling=rolling)
        return g


JOINT_POINT, JOINT_FIXED, JOINT_HINGE, JOINT_SLIDER, JOINT_DISTANCE = range(5)


def chain(links=10, steps=240):
    """A chain of boxes (half-extents 0.3 x 0.1 x 0.1) held end to end by point joints, its first link
    held to the world at (0, 8, 0), laid out level along x so it swings down and folds onto a pile of
    spheres and boxes on a half-space; a second chain hangs from a kinematic body moving along z."""
    s = Scene("chain", substeps=4, steps=steps)
    s.ground(0.6)
    prev = None
    for k in range(links):
        x = f32(0.3) + f32(0.6) * f32(k)
        b = s.body(pos=(x, 8.0, 0.0))
        s.shape(b, BOX, (0.3, 0.1, 0.1), density=500.0)
        anchor = (x - f32(0.3), 8.0, 0.0)
        s.joint(JOINT_POINT, prev, b, anchor=anchor)
        prev = b
    for i in range(3):
        for j in range(3):
            b = s.body(pos=(1.0 + 1.1 * i, 0.6 + 0.0 * j, -1.1 + 1.1 * j))
            if (i + j) % 2:
                s.shape(b, SPHERE, (0.5,), density=300.0)
            else:
                s.shape(b, BOX, (0.5, 0.5, 0.5), density=300.0)
    mover = s.body(KINEMATIC, pos=(-3.0, 6.0, -2.0), v=(0.0, 0.0, 1.0))
    s.shape(mover, BOX, (0.2, 0.2, 0.2))
    prev = mover
    for k in range(4):
        y = f32(5.6) - f32(0.8) * f32(k)
        b = s.body(pos=(-3.0, y, -2.0))
        s.shape(b, SPHERE, (0.2,), density=800.0)
        s.joint(JOINT_POINT, prev, b, anchor=(-3.0, y + f32(0.4), -2.0))
        prev = b
    return s


def joint_mix(steps=240):
    """Every joint the device solves, each beside boxes and spheres that strike it: a bridge of planks
    joined by hinges (unlimited, one with friction, one with a motor) whose ends are held to the world by
    point joints; two boxes welded by a fixed joint tumbling onto it; a box on a rigid rope (distance
    joint) from the world and a sphere on a sprung distance joint with limits; a slider along x with
    limits, motor and friction carrying a box; and a fixed joint that breaks under a falling block."""
    s = Scene("joint-mix", substeps=4, steps=steps)
    s.ground(0.6)
    planks = []
    for k in range(6):
        b = s.body(pos=(f32(-2.5) + f32(k), 3.0, 0.0))
        s.shape(b, BOX, (0.45, 0.05, 0.5), density=400.0)
        planks.append(b)
    s.joint(JOINT_POINT, None, planks[0], anchor=(-3.0, 3.0, 0.0))
    for k in range(5):
        x = f32(-2.0) + f32(k)
        s.joint(JOINT_HINGE, planks[k], planks[k + 1], anchor=(x, 3.0, 0.0), axis=(0.0, 0.0, 1.0),
                friction=0.5 if k == 1 else 0.0, motor_speed=0.5 if k == 3 else 0.0, max_motor=2.0 if k == 3 else 0.0)
    s.joint(JOINT_POINT, None, planks[-1], anchor=(3.0, 3.0, 0.0))
    a = s.body(pos=(-0.5, 5.0, 0.0), w=(0.0, 0.0, 2.0))
    s.shape(a, BOX, (0.25, 0.25, 0.25), density=600.0)
    b = s.body(pos=(0.1, 5.0, 0.0))
    s.shape(b, SPHERE, (0.25,), density=600.0)
    s.joint(JOINT_FIXED, a, b, anchor=(-0.2, 5.0, 0.0))
    rope = s.body(pos=(5.0, 4.0, 0.0), v=(0.0, 0.0, 2.0))
    s.shape(rope, BOX, (0.3, 0.3, 0.3), density=500.0)
    s.joint(JOINT_DISTANCE, None, rope, anchor=(5.0, 7.0, 0.0), anchor_b=(5.0, 4.3, 0.0))
    bob = s.body(pos=(7.0, 4.0, 0.0), v=(1.0, 0.0, 0.0))
    s.shape(bob, SPHERE, (0.3,), density=500.0)
    s.joint(JOINT_DISTANCE, None, bob, anchor=(7.0, 6.0, 0.0), anchor_b=(7.0, 4.0, 0.0), length=1.5, hertz=3.0,
            damping_ratio=0.3, limited=True, lower=1.0, upper=2.5)
    base = s.body(STATIC, pos=(0.0, 0.25, 4.0))
    s.shape(base, BOX, (3.0, 0.25, 0.5))
    cart = s.body(pos=(0.0, 0.8, 4.0))
    s.shape(cart, BOX, (0.3, 0.3, 0.3), density=400.0)
    s.joint(JOINT_SLIDER, base, cart, anchor=(0.0, 0.8, 4.0), axis=(1.0, 0.0, 0.0), limited=True, lower=-1.5,
            upper=1.0, motor_speed=1.0, max_motor=20.0, friction=1.0)
    post = s.body(pos=(-5.0, 0.5, 3.0))
    s.shape(post, BOX, (0.5, 0.5, 0.5), density=200.0)
    top = s.body(pos=(-5.0, 1.5, 3.0))
    s.shape(top, BOX, (0.5, 0.5, 0.5), density=200.0)
    s.joint(JOINT_FIXED, post, top, anchor=(-5.0, 1.0, 3.0), break_force=400.0, break_torque=300.0)
    block = s.body(pos=(-5.0, 5.0, 3.0), v=(0.0, -6.0, 0.0))
    s.shape(block, BOX, (0.6, 0.6, 0.6), density=2000.0)
    return s


def pyramid(rows=20, density=1000.0, friction=0.6, spacing=1.02, steps=600):
    """Boxes of half-extent 0.5 stacked in rows of rows, rows-1, ... on a half-space; row r at
    y = 0.5 + r, box i at x = (i - (n-1)/2) * spacing (the viewer's pyramid, viewer/scenes.cpp:76-87)."""
    s = Scene(f"pyramid-{rows}", substeps=4, steps=steps)
    s.ground(friction)
    for row in range(rows):
        n = rows - row
        for i in range(n):
            x = (f32(i) - f32(0.5) * f32(n - 1)) * f32(spacing)
            b = s.body(DYNAMIC, pos=(x, f32(0.5) + f32(row), 0))
            s.shape(b, BOX, (0.5, 0.5, 0.5), density=density, friction=friction)
    return s


def pile(side=32, spacing=1.2, jitter=0.1, walls=True, seed=0x9E3779B9, density=1000.0,
         friction_range=(0.3, 0.8), restitution_range=None, steps=300, half_room=20.0):
    """The drop-and-settle pile: side^3 bodies on a jittered grid above a half-space inside four
    static walls. Bodies are created y-major (for y, z, x). Per body the draws are: jitter x, y, z;
    then a box's three half-extents U[0.2, 0.5] (even creation index) or a sphere's radius
    U[0.2, 0.5] (odd); friction; rotation axis x, y, z in U[-1, 1]; angle in U[0, pi]; then the
    restitution when a range is given."""
    s = Scene(f"pile-{side ** 3}", substeps=4, steps=steps)
    rng = Rng(seed)
    s.ground(0.6)
    if walls:
        _walls(s, half_room)
    _pile_bodies(s, rng, side, side ** 3, spacing, jitter, density, friction_range, restitution_range)
    return s


def batch_world(n, seed, steps=240):
    """One world of the batched workload: the pile recipe on a ceil(cbrt(n))^3 grid truncated to its
    first n bodies, inside walls sized to the grid, with one friction U[0.2, 1.0] and one
    restitution U[0, 0.3] (the world's first two draws) on all its shapes."""
    side = 1
    while side ** 3 < n:
        side += 1
    s = Scene(f"world-{n}-{seed:08x}", substeps=4, steps=steps)
    rng = Rng(seed)
    fr, e = rng.range(0.2, 1.0), rng.range(0.0, 0.3)
    s.ground(fr, restitution=e)
    _walls(s, f32(0.6) * f32(side) + f32(1.0), friction=fr, restitution=e)
    _pile_bodies(s, rng, side, n, 1.2, 0.1, 1000.0, (fr, fr), (e, e))
    return s


def _walls(s, half_room, friction=0.6, restitution=0.0):
    h = f32(half_room)
    t = f32(0.5)
    for x in (-1, 1):
        wb = s.body(STATIC, pos=(f32(x) * (h + t), h, 0))
        s.shape(wb, BOX, (t, h, h + t), friction=friction, restitution=restitution)
    for z in (-1, 1):
        wb = s.body(STATIC, pos=(0, h, f32(z) * (h + t)))
        s.shape(wb, BOX, (h + t, h, t), friction=friction, restitution=restitution)


def _pile_bodies(s, rng, side, count, spacing, jitter, density, friction_range, restitution_range):
    k = 0
    half = f32(0.5) * f32(side - 1)
    for y in range(side):
        for z in range(side):
            for x in range(side):
                if k == count:
                    return
                jx, jy, jz = rng.range(-jitter, jitter), rng.range(-jitter, jitter), rng.range(-jitter, jitter)
                pos = ((f32(x) - half) * f32(spacing) + jx, f32(1.0) + f32(spacing) * f32(y) + jy,
                       (f32(z) - half) * f32(spacing) + jz)
                if k % 2 == 0:
                    kind, dims = BOX, (rng.range(0.2, 0.5), rng.range(0.2, 0.5), rng.range(0.2, 0.5))
                else:
                    kind, dims = SPHERE, (rng.range(0.2, 0.5),)
                fr = rng.range(*friction_range)
                axis = (rng.range(-1, 1), rng.range(-1, 1), rng.range(-1, 1))
                angle = rng.range(0, PI)
                e = rng.range(*restitution_range) if restitution_range else f32(0)
                b = s.body(DYNAMIC, pos=pos, rot=axis_angle(axis, angle))
                s.shape(b, kind, dims, density=density, friction=fr, restitution=e)
                k += 1


def batch_sizes(worlds, seed=0x9E3779B9, lo=8, hi=1024):
    """Body counts of the batched workload: round(exp(U[ln lo, ln hi])) per world."""
    rng = Rng(seed)
    out = []
    for _ in range(worlds):
        u = rng.range(math.log(lo), math.log(hi))
        out.append(int(round(math.exp(float(u)))))
    return out


def stack(n=10, steps=240):
    """n unit boxes in a column on a half-space (tests/test_world.cpp:235-244)."""
    s = Scene(f"stack-{n}", steps=steps)
    s.ground(0.6)
    for i in range(n):
        b = s.body(DYNAMIC, pos=(0, f32(0.5) + f32(i), 0))
        s.shape(b, BOX, (0.5, 0.5, 0.5), friction=0.6)
    return s


def bounce(restitution=1.0, height=3.0, radius=0.5, steps=120):
    """A sphere dropped on a half-space with restitution (tests/test_world.cpp:246-267)."""
    s = Scene("bounce", steps=steps)
    s.ground(0.6, restitution=restitution)
    b = s.body(DYNAMIC, pos=(0, height, 0))
    s.shape(b, SPHERE, (radius,), restitution=restitution)
    return s


def rolling(speed=4.0, resistance=0.2, steps=240):
    """A rolling ball and box pushed along a half-space with rolling resistance
    (tests/test_world.cpp:1023-1059)."""
    s = Scene("rolling", steps=steps)
    s.ground(0.6, rolling=resistance)
    b = s.body(DYNAMIC, pos=(0, 0.5, 0), v=(speed, 0, 0), w=(0, 0, -speed / 0.5))
    s.shape(b, SPHERE, (0.5,), friction=0.6, rolling=resistance)
    c = s.body(DYNAMIC, pos=(0, 0.5, 3), v=(speed, 0, 0))
    s.shape(c, BOX, (0.5, 0.5, 0.5), friction=0.6, rolling=resistance)
    return s


def planar_mix(steps=240):
    """Planar boxes and a disc locked to the xy plane, a planar rolling disc, and a free sphere
    landing on them (tests/test_world.cpp:410-480, tests/test_planar.cpp:75-117)."""
    s = Scene("planar-mix", steps=steps)
    s.ground(0.6)
    for i in range(4):
        b = s.body(DYNAMIC, pos=(f32(-2.0) + f32(1.1) * f32(i), f32(0.5), 0), locks=LOCK_PLANAR)
        s.shape(b, BOX, (0.5, 0.5, 0.5), friction=0.6)
    top = s.body(DYNAMIC, pos=(-1.45, 1.6, 0), locks=LOCK_PLANAR, w=(0, 0, 1.0))
    s.shape(top, BOX, (0.6, 0.25, 0.5), friction=0.6)
    disc = s.body(DYNAMIC, pos=(4, 0.5, 0), locks=LOCK_PLANAR, w=(0, 0, -6.0))
    s.shape(disc, CYLINDER, (0.5, 0.25), friction=0.6, rolling=0.05)
    ball = s.body(DYNAMIC, pos=(-1.0, 3.5, 0.2))
    s.shape(ball, SPHERE, (0.4,), friction=0.6, restitution=0.3)
    return s


def tumble(n=24, seed=7, steps=180):
    """Mixed boxes and spheres thrown fast at each other and a half-space: fast pairs, impacts,
    restitution and friction all in play."""
    s = Scene("tumble", steps=steps)
    rng = Rng(seed)
    s.ground(0.6, restitution=0.2)
    for i in range(n):
        pos = (rng.range(-3, 3), rng.range(1, 6), rng.range(-3, 3))
        v = (rng.range(-8, 8), rng.range(-12, 2), rng.range(-8, 8))
        w = (rng.range(-6, 6), rng.range(-6, 6), rng.range(-6, 6))
        rot = axis_angle((rng.range(-1, 1), rng.range(-1, 1), rng.range(-1, 1)), rng.range(0, PI))
        b = s.body(DYNAMIC, pos=pos, rot=rot, v=v, w=w)
        if i % 2 == 0:
            s.shape(b, BOX, (rng.range(0.2, 0.5), rng.range(0.2, 0.5), rng.range(0.2, 0.5)),
                    friction=rng.range(0.2, 1.0), restitution=rng.range(0, 0.4))
        else:
            s.shape(b, SPHERE, (rng.range(0.2, 0.5),), friction=rng.range(0.2, 1.0),
                    restitution=rng.range(0, 0.4))
    return s


def hub(n=7, steps=90):
    """A dynamic slab carrying n x n spheres on a half-space. The slab has n * n + 1 contacts, so once
    its 32 colours are taken the rest of its contacts go to the serial rest list
    (src/lanes.cpp:356-374) and are solved one at a time after the coloured rows."""
    s = Scene(f"hub-{n}", steps=steps)
    s.ground(0.6)
    half = f32(0.25) * f32(n) + f32(0.05)
    slab = s.body(DYNAMIC, pos=(0, 0.25, 0))
    s.shape(slab, BOX, (half, 0.25, half), density=500.0, friction=0.5)
    for i in range(n):
        for k in range(n):
            x = (f32(i) - f32(n - 1) / f32(2)) * f32(0.5)
            z = (f32(k) - f32(n - 1) / f32(2)) * f32(0.5)
            b = s.body(DYNAMIC, pos=(x, 0.7, z))
            s.shape(b, SPHERE, (0.2,), density=1000.0, friction=0.5)
    return s


def edges(steps=240):
    """The branches of the path that the other scenes leave alone, in one world:
    - damping, gravity scale, and a force and a torque applied every step (integrateVelocities,
      src/world.cpp:1141-1151);
    - a lowered speed cap (7 m/s, 9 rad/s) that thrown and falling bodies reach, so limitSpeed clamps
      and `kept` drops below 1 (src/solver.cpp:30-41; src/world.cpp:1150, 1162, 1684);
    - a kinematic platform sliding at 3 cm/s (below the sleep threshold) carrying two boxes, and a
      kinematic turntable carrying a sphere: kinematic-dynamic contacts with static softness, and the
      kinematic wake of updateSleep (src/world.cpp:1204-1207);
    - restitution on contacts outside the colours: planar boxes and a planar disc dropped with
      restitution 0.5 (solvePlanar's list), a ball and a box with rolling resistance dropped with
      restitution (the 3D rest list), and a slab under 7 x 7 dropped spheres (more contacts than
      colours), all landing faster than the restitution threshold (applyRestitution,
      src/solver.cpp:423-460).
    The ground has restitution 0.4 and friction 0.6."""
    s = Scene("edges", steps=steps, settings=dict(max_linear_speed=7.0, max_angular_speed=9.0))
    s.ground(0.6, restitution=0.4)
    # damping, gravity scale, forces and torques
    b = s.body(DYNAMIC, pos=(-6, 2.0, -4), v=(1.0, 0, 0), w=(0, 3.0, 0), linear_damping=0.8, angular_damping=0.5)
    s.shape(b, BOX, (0.4, 0.3, 0.5), density=800.0, friction=0.5)
    b = s.body(DYNAMIC, pos=(-4, 3.0, -4), gravity_scale=0.25, angular_damping=0.3)
    s.shape(b, SPHERE, (0.35,), density=900.0, friction=0.7)
    b = s.body(DYNAMIC, pos=(-2, 1.5, -4), gravity_scale=1.7, linear_damping=0.1)
    s.shape(b, BOX, (0.3, 0.3, 0.3), density=1000.0, friction=0.6)
    b = s.body(DYNAMIC, pos=(0, 0.6, -4), force=(400.0, 0.0, -150.0), torque=(0.0, 60.0, 0.0))
    s.shape(b, BOX, (0.4, 0.4, 0.4), density=1000.0, friction=0.4)
    b = s.body(DYNAMIC, pos=(2, 0.8, -4), force=(-90.0, 250.0, 40.0), torque=(35.0, 0.0, -20.0), linear_damping=0.2)
    s.shape(b, SPHERE, (0.3,), density=1000.0, friction=0.6)
    # the speed cap: thrown and falling bodies
    b = s.body(DYNAMIC, pos=(4, 1.0, -4), v=(12.0, 3.0, 0.0), w=(0.0, 15.0, 4.0))
    s.shape(b, BOX, (0.3, 0.3, 0.3), density=1000.0, friction=0.6)
    b = s.body(DYNAMIC, pos=(6, 6.0, -4), w=(20.0, 0.0, 0.0))
    s.shape(b, SPHERE, (0.3,), density=1000.0, friction=0.6, restitution=0.3)
    # a sliding kinematic platform and a turntable
    plat = s.body(KINEMATIC, pos=(-6, 0.25, 3), v=(0.03, 0.0, 0.0))
    s.shape(plat, BOX, (1.5, 0.25, 1.5), friction=0.8)
    for k in range(2):
        b = s.body(DYNAMIC, pos=(f32(-6.6) + f32(1.2) * f32(k), 0.8, 3))
        s.shape(b, BOX, (0.3, 0.3, 0.3), density=1000.0, friction=0.8)
    table = s.body(KINEMATIC, pos=(-2, 0.2, 3), w=(0.0, 0.8, 0.0))
    s.shape(table, BOX, (1.0, 0.2, 1.0), friction=0.9)
    b = s.body(DYNAMIC, pos=(-1.4, 0.75, 3))
    s.shape(b, SPHERE, (0.35,), density=1000.0, friction=0.9, rolling=0.05)
    # restitution on the planar rest list
    for k in range(3):
        b = s.body(DYNAMIC, pos=(f32(1.5) + f32(1.3) * f32(k), f32(1.2) + f32(0.4) * f32(k), 3), locks=LOCK_PLANAR,
                   w=(0, 0, f32(0.5) * f32(k)))
        s.shape(b, BOX, (0.5, 0.4, 0.5), density=1000.0, friction=0.6, restitution=0.5)
    disc = s.body(DYNAMIC, pos=(6, 1.6, 3), locks=LOCK_PLANAR, w=(0, 0, -3.0))
    s.shape(disc, CYLINDER, (0.5, 0.25), density=1000.0, friction=0.6, restitution=0.5, rolling=0.05)
    # restitution on the 3D rest list: rolling resistance
    b = s.body(DYNAMIC, pos=(-6, 2.0, 8), v=(2.0, 0, 0), w=(0, 0, -4.0))
    s.shape(b, SPHERE, (0.5,), density=1000.0, friction=0.6, restitution=0.6, rolling=0.1)
    b = s.body(DYNAMIC, pos=(-3, 1.5, 8), v=(1.5, 0, 0))
    s.shape(b, BOX, (0.4, 0.4, 0.4), density=1000.0, friction=0.6, restitution=0.5, rolling=0.1)
    # restitution past the colours: a slab under 49 dropped spheres
    n = 7
    half = f32(0.25) * f32(n) + f32(0.05)
    slab = s.body(DYNAMIC, pos=(3, 0.25, 8))
    s.shape(slab, BOX, (half, 0.25, half), density=500.0, friction=0.5, restitution=0.6)
    for i in range(n):
        for k in range(n):
            x = f32(3) + (f32(i) - f32(n - 1) / f32(2)) * f32(0.5)
            z = f32(8) + (f32(k) - f32(n - 1) / f32(2)) * f32(0.5)
            b = s.body(DYNAMIC, pos=(x, 1.3, z))
            s.shape(b, SPHERE, (0.2,), density=1000.0, friction=0.5, restitution=0.6)
    return s


def prism_profile(rng, sides, radius):
    """A convex polygon of `sides` vertices on a circle of `radius`, vertex k at angle
    (k + U[0.15, 0.85]) * 2 pi / sides (sorted random angles, no sliver), moved so that its centroid
    is the origin. Returns x0, y0, x1, y1, ... (counter-clockwise, float32)."""
    xs, ys = [], []
    for k in range(sides):
        a = (f32(k) + rng.range(0.15, 0.85)) * f32(2.0 * PI) / f32(sides)
        xs.append(f32(radius * np.cos(a)))
        ys.append(f32(radius * np.sin(a)))
    # centroid of the polygon (shoelace), in float64, then rounded
    area = cx = cy = 0.0
    for i in range(sides):
        j = (i + 1) % sides
        c = float(xs[i]) * float(ys[j]) - float(xs[j]) * float(ys[i])
        area += c
        cx += (float(xs[i]) + float(xs[j])) * c
        cy += (float(ys[i]) + float(ys[j])) * c
    cx /= 3.0 * area
    cy /= 3.0 * area
    out = []
    for x, y in zip(xs, ys):
        out += [f32(float(x) - cx), f32(float(y) - cy)]
    return out


def planar_pile(side=256, spheres=4096, spacing=1.3, seed=0x9E3779B9, steps=240):
    """The planar path at scale (the planar + 3D workload): side x side bodies locked to the xy plane
    (LockPlanar, z = 0) on a grid in x-y from y = 1, inside a static U container (half-space floor,
    two walls of z half-depth 5 m), with free spheres (radius 0.3) dropped above them. Planar bodies:
    40% boxes of half-extents (U[0.2, 0.6], U[0.2, 0.6], 0.5) turned about z by U[0, pi]; 40% prisms
    of 3-8 sides (prism_profile, radius U[0.2, 0.6], z from -0.5 to 0.5); 20% cylinders of radius
    U[0.2, 0.6] and half height 0.5 whose axis is the plane normal. Friction 0.6. Draws per planar
    body, rows bottom-up: the kind (U[0, 1)), then box: half-extent x, y, angle; prism: sides
    (3 + floor(6 U)), radius, one draw per vertex; cylinder: radius. Per sphere: x, y offset, z."""
    s = Scene(f"planar-{side * side}", substeps=8, steps=steps)
    rng = Rng(seed)
    s.ground(0.6)
    half = f32(0.5) * f32(side - 1)
    wall_x = (half + f32(1.0)) * f32(spacing)
    top = f32(1.0) + f32(spacing) * f32(side)
    for x in (-1, 1):
        wb = s.body(STATIC, pos=(f32(x) * (wall_x + f32(0.5)), top * f32(0.5), 0))
        s.shape(wb, BOX, (0.5, top * f32(0.5) + f32(4.0), 5.0), friction=0.6)
    for j in range(side):
        for i in range(side):
            pos = ((f32(i) - half) * f32(spacing), f32(1.0) + f32(spacing) * f32(j), 0)
            u = rng.next()
            if u < 0.4:
                hx, hy, angle = rng.range(0.2, 0.6), rng.range(0.2, 0.6), rng.range(0, PI)
                b = s.body(DYNAMIC, pos=pos, rot=axis_angle((0, 0, 1), angle), locks=LOCK_PLANAR)
                s.shape(b, BOX, (hx, hy, 0.5), density=1000.0, friction=0.6)
            elif u < 0.8:
                sides = 3 + min(int(rng.next() * 6), 5)
                radius = rng.range(0.2, 0.6)
                b = s.body(DYNAMIC, pos=pos, locks=LOCK_PLANAR)
                s.shape(b, PRISM, [-0.5, 0.5] + prism_profile(rng, sides, radius), density=1000.0, friction=0.6)
            else:
                b = s.body(DYNAMIC, pos=pos, locks=LOCK_PLANAR)
                s.shape(b, CYLINDER, (rng.range(0.2, 0.6), 0.5), density=1000.0, friction=0.6)
    for k in range(spheres):
        x = rng.range(-wall_x + f32(1.0), wall_x - f32(1.0))
        y = top + f32(2.0) + rng.range(0.0, 64.0)
        z = rng.range(-0.3, 0.3)
        b = s.body(DYNAMIC, pos=(x, y, z))
        s.shape(b, SPHERE, (0.3,), density=1000.0, friction=0.6)
    return s


def planar_pile_small(steps=150):
    """A 10 x 10 planar pile with 12 spheres: a rest list of up to ~160 planar contacts."""
    return planar_pile(side=10, spheres=12, steps=steps)


def throw_world(seed, steps=32, jitter=0.005):
    """One world of the adjoint workload: 64 boxes of half-extent 0.25 in a 4 x 4 x 4 block of touching
    boxes (spacing 0.5, created y-major: for y, z, x) and a wall of 10 such boxes (2 high, 5 wide)
    against the block's +x face, offset half a box in z. No static body: everything is free and falls
    together. Draws, in order: the world's friction U[0.2, 1.0] (on every shape); per block box a
    jitter x, y, z in U[-jitter, jitter], then v0 ~ N(0, 2 m/s) per axis and w0 ~ N(0, 1 rad/s) per
    axis (Box-Muller, two uniform draws each); per wall box a jitter x, y, z, then v0 and w0 ~ N(0,
    0.05) per axis. Restitution is 0. The jitter keeps the points off zero separation, where the bias
    switches form, and the wall's small velocities keep its contacts off exactly zero relative
    velocity, where a clamp switches: both are kinks of the loss."""
    s = Scene(f"throw-{seed:08x}", substeps=4, steps=steps)
    rng = Rng(seed)
    fr = rng.range(0.2, 1.0)
    half = (0.25, 0.25, 0.25)
    for y in range(4):
        for z in range(4):
            for x in range(4):
                j = [rng.range(-jitter, jitter) for _ in range(3)]
                p = ((f32(x) - f32(1.5)) * f32(0.5) + j[0], (f32(y) - f32(1.5)) * f32(0.5) + j[1],
                     (f32(z) - f32(1.5)) * f32(0.5) + j[2])
                v = [rng.normal(2.0) for _ in range(3)]
                w = [rng.normal(1.0) for _ in range(3)]
                b = s.body(DYNAMIC, pos=p, v=v, w=w)
                s.shape(b, BOX, half, density=1000.0, friction=fr)
    for y in range(2):
        for z in range(5):
            j = [rng.range(-jitter, jitter) for _ in range(3)]
            p = (f32(1.25) + j[0], (f32(y) - f32(0.5)) * f32(0.5) + j[1], (f32(z) - f32(2.0)) * f32(0.5) + j[2])
            v = [rng.normal(0.05) for _ in range(3)]
            w = [rng.normal(0.05) for _ in range(3)]
            b = s.body(DYNAMIC, pos=p, v=v, w=w)
            s.shape(b, BOX, half, density=1000.0, friction=fr)
    return s
