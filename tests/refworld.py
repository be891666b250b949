"""ctypes wrappers for the two checkers under oracle/ (test infrastructure).

RefWorld drives the reference engine built from its own sources (oracle/_ref/libtwofold_ref.so) and
reads back what its solve block saw and did each step. Oracle runs the CPU restatement
(oracle/liboracle.so) on the same POD arrays the HIP path takes.
"""
import ctypes
import os

import numpy as np

from twofold_amd import _abi
from twofold_amd._abi import BODY_DTYPE, CONTACT_DTYPE

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF_SO = os.path.join(ROOT, "oracle", "_ref", "libtwofold_ref.so")
# the same engine with the drop-in binding compiled in (oracle/ref_bound.cpp): solve block and sleep pass
# on the GPU through the C ABI
BOUND_SO = os.path.join(ROOT, "oracle", "_ref", "libtwofold_bound.so")
ORACLE_SO = os.path.join(ROOT, "oracle", "liboracle.so")

_fp = ctypes.POINTER(ctypes.c_float)
_vp = ctypes.c_void_p


def _ptr(a):
    return a.ctypes.data_as(_vp) if a is not None else None


def _f(v):
    return np.ascontiguousarray(v, np.float32).ctypes.data_as(_fp)


_refs = {}


def ref_available():
    return os.path.exists(REF_SO)


def _load_ref(path=REF_SO):
    if path not in _refs:
        lib = ctypes.CDLL(path)
        lib.tfr_create.restype = _vp
        lib.tfr_create.argtypes = [_vp]
        lib.tfr_destroy.argtypes = [_vp]
        lib.tfr_copy.argtypes = [_vp, _vp]
        lib.tfr_set_continuous.argtypes = [_vp, ctypes.c_int]
        lib.tfr_add_body.restype = ctypes.c_int
        lib.tfr_add_body.argtypes = [_vp, ctypes.c_int, _fp, _fp, _fp, _fp, ctypes.c_int, _fp, ctypes.c_int,
                                     ctypes.c_float, ctypes.c_float, ctypes.c_float]
        lib.tfr_add_shape.restype = ctypes.c_int
        lib.tfr_add_shape.argtypes = [_vp, ctypes.c_int, ctypes.c_int, _fp, ctypes.c_int, _fp, _fp,
                                      ctypes.c_float, ctypes.c_float, ctypes.c_float, ctypes.c_float]
        lib.tfr_step.restype = ctypes.c_int
        lib.tfr_step.argtypes = [_vp, ctypes.c_float, ctypes.c_int]
        lib.tfr_profile.argtypes = [_vp, _fp]
        lib.tfr_counts.argtypes = [_vp, _vp]
        lib.tfr_set_bodies.argtypes = [_vp, _vp, ctypes.c_uint32]
        lib.tfr_add_force.argtypes = [_vp, ctypes.c_int, _fp, _fp]
        for name in ("tfr_num_bodies", "tfr_num_contacts", "tfr_cap_num_refreshes"):
            getattr(lib, name).restype = ctypes.c_uint32
            getattr(lib, name).argtypes = [_vp]
        for name in ("tfr_bodies", "tfr_contacts"):
            getattr(lib, name).restype = ctypes.c_uint32
            getattr(lib, name).argtypes = [_vp, _vp, ctypes.c_uint32]
        for name in ("tfr_cap_bodies", "tfr_cap_contacts", "tfr_cap_list"):
            getattr(lib, name).restype = ctypes.c_uint32
            getattr(lib, name).argtypes = [_vp, ctypes.c_int, _vp, ctypes.c_uint32]
        lib.tfr_cap_colors.restype = ctypes.c_uint32
        lib.tfr_cap_colors.argtypes = [_vp, _vp, ctypes.c_uint32]
        lib.tfr_cap_ctx.argtypes = [_vp, _fp]
        for name in ("tfr_cap_old_bodies", "tfr_cap_pairs", "tfr_cap_manifolds"):
            getattr(lib, name).restype = ctypes.c_uint32
            getattr(lib, name).argtypes = [_vp, _vp, ctypes.c_uint32]
        lib.tfr_cap_ext.restype = ctypes.c_uint32
        lib.tfr_cap_ext.argtypes = [_vp, ctypes.c_int, _vp, ctypes.c_uint32]
        lib.tfr_hits.restype = ctypes.c_uint32
        lib.tfr_hits.argtypes = [_vp, _vp, ctypes.c_uint32]
        lib.tfr_body_shapes.restype = ctypes.c_uint32
        lib.tfr_body_shapes.argtypes = [_vp, _vp, ctypes.c_uint32]
        lib.tfr_cap_refresh.restype = ctypes.c_uint32
        lib.tfr_cap_refresh.argtypes = [_vp, ctypes.c_uint32, ctypes.POINTER(ctypes.c_int), _fp, _vp, _vp,
                                        ctypes.c_uint32]
        lib.tfr_add_joint.restype = ctypes.c_int
        lib.tfr_add_joint.argtypes = [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, _fp]
        lib.tfr_cap_joints.restype = ctypes.c_uint32
        lib.tfr_cap_joints.argtypes = [_vp, ctypes.c_int, _vp, ctypes.c_uint32, _fp]
        lib.tfr_joints.restype = ctypes.c_uint32
        lib.tfr_joints.argtypes = [_vp, _vp, ctypes.c_uint32]
        _abi.check_sizes(lib, "tfr")
        if path == BOUND_SO:
            lib.tfb_error.restype = ctypes.c_char_p
        _refs[path] = lib
    return _refs[path]


class StepRecord:
    """What the reference's solve block saw and did in one step."""

    __slots__ = ("pre_bodies", "pre_contacts", "post_bodies", "post_contacts", "step_bodies", "lists",
                 "colors", "ctx", "refreshes", "dt", "after_contacts", "hits", "touching", "old_bodies",
                 "old_contacts", "old_ext", "pre_ext", "pairs", "manifolds", "pre_joints", "post_joints",
                 "pre_world", "post_world")


class RefWorld:
    """The reference engine, one world, built from a scene recipe."""

    def __init__(self, scene, settings=None, continuous=True, lib_path=REF_SO):
        self.lib = _load_ref(lib_path)
        s = settings if settings is not None else _abi.default_settings(substeps=scene.substeps,
                                                                        **getattr(scene, "settings", {}))
        self.settings = s
        self.forces = scene.forces() if hasattr(scene, "forces") else []
        self.h = self.lib.tfr_create(s.ctypes.data_as(_vp))
        self.dt = scene.dt
        if not continuous:
            self.lib.tfr_set_continuous(self.h, 0)
        for b in scene.bodies:
            self.lib.tfr_add_body(self.h, b["type"], _f(b["pos"]), _f(b["rot"]), _f(b["v"]), _f(b["w"]),
                                  b["locks"], _f(b["lock_frame"]), 1 if b["awake"] else 0,
                                  b["linear_damping"], b["angular_damping"], b["gravity_scale"])
        for sh in scene.shapes:
            dims = sh["dims"] if len(sh["dims"]) else np.zeros(1, np.float32)
            r = self.lib.tfr_add_shape(self.h, sh["body"], sh["kind"], _f(dims), len(sh["dims"]), _f(sh["pos"]),
                                       _f(sh["rot"]), sh["density"], sh["friction"], sh["restitution"],
                                       sh["rolling"])
            if r < 0:
                raise ValueError("shape rejected")
        for j in getattr(scene, "joints", []):
            if self.lib.tfr_add_joint(self.h, j["kind"], j["a"], j["b"], _f(j["p"])) < 0:
                raise ValueError("joint rejected")

    def __del__(self):
        if getattr(self, "h", None):
            self.lib.tfr_destroy(self.h)
            self.h = None

    def set_bodies(self, bodies):
        b = np.ascontiguousarray(bodies, BODY_DTYPE)
        self.lib.tfr_set_bodies(self.h, _ptr(b), len(b))

    def counts(self):
        """Active contacts, their points, impact contacts, their points, of the last step."""
        out = np.zeros(4, np.uint32)
        self.lib.tfr_counts(self.h, _ptr(out))
        return out

    def step(self, dt=None, record=True):
        dt = self.dt if dt is None else np.float32(dt)
        for body, force, torque in self.forces:
            self.lib.tfr_add_force(self.h, body, _f(force), _f(torque))
        solved = self.lib.tfr_step(self.h, dt, 1 if record else 2)
        if not record:
            return None
        if not solved:
            return None
        return self._record(dt)

    def profile(self):
        out = np.zeros(7, np.float32)
        self.lib.tfr_profile(self.h, out.ctypes.data_as(_fp))
        return dict(zip(("total", "broadphase", "narrowphase", "solver", "continuous", "sleep", "touching"),
                        out.tolist()))

    def hits(self):
        n = self.lib.tfr_hits(self.h, None, 0)
        out = np.zeros(n, _abi.HIT_DTYPE)
        self.lib.tfr_hits(self.h, _ptr(out), n)
        return out

    def body_shapes(self):
        """Each body's one shape as the device refresh takes it (SHAPE_DTYPE), with the reference's sizes;
        kind 255 where a body has another geometry or more than one shape."""
        n = self.lib.tfr_num_bodies(self.h)
        out = np.zeros(n, _abi.SHAPE_DTYPE)
        self.lib.tfr_body_shapes(self.h, _ptr(out), n)
        return out

    def bodies(self):
        n = self.lib.tfr_num_bodies(self.h)
        out = np.zeros(n, BODY_DTYPE)
        self.lib.tfr_bodies(self.h, _ptr(out), n)
        return out

    def joints(self):
        n = self.lib.tfr_joints(self.h, None, 0)
        out = np.zeros(n, _abi.JOINT_DTYPE)
        self.lib.tfr_joints(self.h, _ptr(out), n)
        return out

    def contacts(self):
        n = self.lib.tfr_num_contacts(self.h)
        out = np.zeros(n, CONTACT_DTYPE)
        self.lib.tfr_contacts(self.h, _ptr(out), n)
        return out

    def _record(self, dt):
        lib, h = self.lib, self.h
        r = StepRecord()
        r.dt = dt
        nb = lib.tfr_cap_bodies(h, 0, None, 0)
        nc = lib.tfr_cap_contacts(h, 0, None, 0)
        def bodies(which):
            a = np.zeros(nb, BODY_DTYPE)
            lib.tfr_cap_bodies(h, which, _ptr(a), nb)
            return a
        def contacts(which):
            a = np.zeros(nc, CONTACT_DTYPE)
            lib.tfr_cap_contacts(h, which, _ptr(a), nc)
            return a
        r.pre_bodies, r.post_bodies, r.step_bodies = bodies(0), bodies(1), bodies(2)
        r.pre_contacts, r.post_contacts = contacts(0), contacts(1)
        # updateContacts: the state it started from, its candidate pairs and their manifolds, the reuse data
        def grab(fn, dtype, *which):
            n = fn(h, *which, None, 0)
            a = np.zeros(n, dtype)
            fn(h, *which, _ptr(a), n)
            return a
        r.old_bodies = grab(lib.tfr_cap_old_bodies, BODY_DTYPE)
        r.old_contacts = grab(lib.tfr_cap_contacts, CONTACT_DTYPE, 2)
        r.old_ext = grab(lib.tfr_cap_ext, _abi.EXT_DTYPE, 2)
        r.pre_ext = grab(lib.tfr_cap_ext, _abi.EXT_DTYPE, 0)
        r.pairs = grab(lib.tfr_cap_pairs, _abi.PAIR_DTYPE)
        nj = lib.tfr_cap_joints(h, 0, None, 0, None)
        r.pre_joints, r.post_joints = np.zeros(nj, _abi.JOINT_DTYPE), np.zeros(nj, _abi.JOINT_DTYPE)
        r.pre_world, r.post_world = np.zeros(6, np.float32), np.zeros(6, np.float32)
        lib.tfr_cap_joints(h, 0, _ptr(r.pre_joints), nj, r.pre_world.ctypes.data_as(_fp))
        lib.tfr_cap_joints(h, 1, _ptr(r.post_joints), nj, r.post_world.ctypes.data_as(_fp))
        r.manifolds = grab(lib.tfr_cap_manifolds, _abi.MANIFOLD_DTYPE)
        # after the whole step: the contacts' touched flags as reportTouches left them, its hits and count
        r.after_contacts = self.contacts()
        r.hits = self.hits()
        r.touching = int(self.profile()["touching"])
        names = ("active", "fast", "fast_solid", "impact", "rows", "rest", "impact_rows", "impact_rest")
        r.lists = {}
        for k, name in enumerate(names):
            n = lib.tfr_cap_list(h, k, None, 0)
            a = np.zeros(n, np.uint32)
            lib.tfr_cap_list(h, k, _ptr(a), n)
            r.lists[name] = a
        r.colors = np.zeros(nc, np.uint8)
        lib.tfr_cap_colors(h, _ptr(r.colors), nc)
        r.ctx = np.zeros(10, np.float32)
        lib.tfr_cap_ctx(h, r.ctx.ctypes.data_as(_fp))
        r.refreshes = []
        for k in range(lib.tfr_cap_num_refreshes(h)):
            sub = ctypes.c_int()
            rem = np.zeros(1, np.float32)
            n = lib.tfr_cap_refresh(h, k, ctypes.byref(sub), rem.ctypes.data_as(_fp), None, None, 0)
            again = np.zeros(n, np.uint32)
            fresh = np.zeros(n, CONTACT_DTYPE)
            lib.tfr_cap_refresh(h, k, None, None, _ptr(again), _ptr(fresh), n)
            r.refreshes.append(dict(substep=sub.value, remaining=rem[0], again=again, fresh=fresh))
        return r


# ---- the oracle -------------------------------------------------------------------------------

REFRESH_FN = ctypes.CFUNCTYPE(ctypes.c_int, _vp, ctypes.c_int, ctypes.c_float, ctypes.c_uint32,
                              ctypes.POINTER(ctypes.c_uint32), _vp, _vp, _vp)


class _Lists(ctypes.Structure):
    _fields_ = [(n, ctypes.POINTER(ctypes.c_uint32)) for n in
                ("active", "fast", "fast_solid", "impact", "by_color", "impact_by_color")] + [
        ("colors", ctypes.POINTER(ctypes.c_uint8))] + [
        (n, ctypes.c_uint32) for n in ("n_active", "n_fast", "n_fast_solid", "n_impact", "n_rest", "n_impact_rest",
                                       "n_colors", "points_active", "points_impact")] + [
        ("touch_events", ctypes.POINTER(ctypes.c_uint32)), ("hits", ctypes.c_void_p)] + [
        (n, ctypes.c_uint32) for n in ("n_touch_events", "n_hits", "touching")] + [("ext", ctypes.c_void_p)]


_oracle = None


def _load_oracle():
    global _oracle
    if _oracle is None:
        lib = ctypes.CDLL(ORACLE_SO)
        lib.tfo_step.restype = ctypes.c_int
        lib.tfo_step.argtypes = [_vp, ctypes.c_uint32, _vp, ctypes.c_uint32, _vp, ctypes.c_float, REFRESH_FN, _vp,
                                 ctypes.POINTER(_Lists)]
        lib.tfo_update_inverse_mass.argtypes = [_vp]
        lib.tfo_make_soft.argtypes = [ctypes.c_float, ctypes.c_float, ctypes.c_float, _fp]
        lib.tfo_maxf.restype = ctypes.c_float
        lib.tfo_maxf.argtypes = [ctypes.c_float, ctypes.c_float]
        lib.tfo_minf.restype = ctypes.c_float
        lib.tfo_minf.argtypes = [ctypes.c_float, ctypes.c_float]
        lib.tfo_color.argtypes = [_vp, ctypes.c_uint32, _vp, _vp, ctypes.c_uint32, _vp]
        _oracle = lib
    return _oracle


def recorded_refresh(refreshes):
    """A refresh callback that hands back the manifolds the reference's narrowphase found."""
    by_sub = {r["substep"]: r for r in refreshes}

    def cb(user, substep, remaining, n_again, again, contacts, bodies, fresh):
        r = by_sub.get(substep)
        if r is None or len(r["again"]) != n_again:
            return -1
        got = np.ctypeslib.as_array(again, shape=(n_again,))
        if not np.array_equal(got, r["again"]):
            return -1
        out = np.ctypeslib.as_array(ctypes.cast(fresh, ctypes.POINTER(ctypes.c_uint8)),
                                    shape=(n_again * CONTACT_DTYPE.itemsize,))
        out[:] = r["fresh"].view(np.uint8)
        return 0

    return cb


def oracle_stream(bodies, old, old_ext, pairs, manifolds, settings):
    """tfo_stream_contacts: the contacts and reuse data the contact stream leaves."""
    lib = _load_oracle()
    lib.tfo_stream_contacts.restype = ctypes.c_uint32
    lib.tfo_stream_contacts.argtypes = [_vp, _vp, _vp, ctypes.c_uint32, _vp, ctypes.c_uint32, _vp, _vp, _vp, _vp]
    cap = max(len(old) + len(manifolds), 1)
    out = np.zeros(cap, CONTACT_DTYPE)
    ext = np.zeros(cap, _abi.EXT_DTYPE)
    args = [np.ascontiguousarray(x) for x in (bodies, old, old_ext, pairs, manifolds)]
    n = lib.tfo_stream_contacts(_ptr(args[0]), _ptr(args[1]), _ptr(args[2]), len(old), _ptr(args[3]), len(pairs),
                                _ptr(args[4]), settings.ctypes.data_as(_vp), _ptr(out), _ptr(ext))
    return out[:n], ext[:n]


def oracle_step(bodies, contacts, settings, dt, refresh=None, ext=None):
    """Runs tfo_step in place on copies; returns (bodies, contacts, lists dict). `ext` (the contacts'
    reuse data) is updated in place when given."""
    lib = _load_oracle()
    b = np.ascontiguousarray(bodies.copy())
    c = np.ascontiguousarray(contacts.copy())
    n = max(len(c), 1)
    arrays = {k: np.zeros(n, np.uint32) for k in ("active", "fast", "fast_solid", "impact", "by_color",
                                                  "impact_by_color")}
    colors = np.zeros(n, np.uint8)
    L = _Lists()
    for k, a in arrays.items():
        setattr(L, k, a.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)))
    L.colors = colors.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8))
    events = np.zeros(n, np.uint32)
    hits = np.zeros(n, _abi.HIT_DTYPE)
    L.touch_events = events.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32))
    L.hits = hits.ctypes.data
    L.ext = ext.ctypes.data if ext is not None else None
    cb = REFRESH_FN(refresh) if refresh is not None else REFRESH_FN()
    rc = lib.tfo_step(_ptr(b), len(b), _ptr(c), len(c), settings.ctypes.data_as(_vp), np.float32(dt), cb, None,
                      ctypes.byref(L))
    if rc != 0:
        raise RuntimeError(f"tfo_step failed: {rc}")
    lists = dict(
        active=arrays["active"][:L.n_active], fast=arrays["fast"][:L.n_fast],
        fast_solid=arrays["fast_solid"][:L.n_fast_solid], impact=arrays["impact"][:L.n_impact],
        rows=arrays["by_color"][:L.n_active], rest=arrays["by_color"][L.n_active - L.n_rest:L.n_active],
        impact_rows=arrays["impact_by_color"][:L.n_impact],
        impact_rest=arrays["impact_by_color"][L.n_impact - L.n_impact_rest:L.n_impact],
        colors=colors[:len(contacts)], n_colors=L.n_colors, points_active=L.points_active,
        points_impact=L.points_impact, touch_events=events[:L.n_touch_events], hits=hits[:L.n_hits],
        touching=L.touching)
    return b, c, lists
