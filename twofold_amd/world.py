"""A host world that drives the GPU solver end to end (libtwofold_scene.so + libtwofold_hip.so).

HostWorld holds boxes, spheres and half-spaces built from a scene recipe (twofold_amd.scenes) and runs
the broadphase and narrowphase on the host; Simulation steps it the way World::step does
(src/world.cpp:1564-1727): contacts from the narrowphase, then the solve block and the sleep pass on
the GPU, with fast pairs collided again mid-step through the native refresh callback. Only the solver
is the product's hot path; this module is how a user (and the benchmark) feeds it.
"""
import ctypes
import os

import numpy as np

from . import _abi
from ._abi import BODY_DTYPE, CONTACT_DTYPE, MANIFOLD_DTYPE, PAIR_DTYPE, SETTINGS_DTYPE, SHAPE_DTYPE
from .solver import REFRESH_FN, Solver, SolverError

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libtwofold_scene.so")

_vp = ctypes.c_void_p
_fp = ctypes.POINTER(ctypes.c_float)
_lib = None


def _load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise SolverError(f"{LIB_PATH} is missing: build it with __graft_entry__.build()")
        lib = ctypes.CDLL(LIB_PATH)
        lib.tfw_create.restype = _vp
        lib.tfw_create.argtypes = [_vp]
        lib.tfw_destroy.argtypes = [_vp]
        lib.tfw_add_body.restype = ctypes.c_int
        lib.tfw_add_body.argtypes = [_vp, ctypes.c_int, _fp, _fp, _fp, _fp, ctypes.c_int, _fp, ctypes.c_int,
                                     ctypes.c_float, ctypes.c_float, ctypes.c_float]
        lib.tfw_add_shape.restype = ctypes.c_int
        lib.tfw_add_shape.argtypes = [_vp, ctypes.c_int, ctypes.c_int, _fp, ctypes.c_int, _fp, _fp, ctypes.c_float,
                                      ctypes.c_float, ctypes.c_float, ctypes.c_float]
        lib.tfw_set_group.restype = ctypes.c_int
        lib.tfw_set_group.argtypes = [_vp, ctypes.c_uint32, ctypes.c_uint32]
        lib.tfw_bodies.restype = ctypes.c_uint32
        lib.tfw_bodies.argtypes = [_vp, _vp, ctypes.c_uint32]
        lib.tfw_collide.restype = ctypes.c_uint32
        lib.tfw_collide.argtypes = [_vp, _vp, ctypes.c_uint32, ctypes.c_float, _vp, ctypes.c_uint32, _vp,
                                    ctypes.c_uint32, ctypes.POINTER(ctypes.c_uint32)]
        lib.tfw_pairs.restype = ctypes.c_uint32
        lib.tfw_pairs.argtypes = [_vp, _vp, ctypes.c_uint32, ctypes.c_float, _vp, _vp, ctypes.c_uint32,
                                  ctypes.POINTER(ctypes.c_uint32), ctypes.POINTER(ctypes.c_uint32)]
        lib.tfw_body_shapes.restype = ctypes.c_int
        lib.tfw_body_shapes.argtypes = [_vp, _vp, ctypes.c_uint32]
        _lib = lib
    return _lib


def _f(v):
    return np.ascontiguousarray(v, np.float32).ctypes.data_as(_fp)


class HostWorld:
    """Bodies and shapes of a scene recipe, with a narrowphase for them. Given a list of recipes, it
    holds them as independent worlds, one after the other (`body_start`), each its own collision
    group."""

    def __init__(self, scene, settings=None):
        self.lib = _load()
        worlds = list(scene) if isinstance(scene, (list, tuple)) else [scene]
        first = worlds[0]
        self.settings = settings if settings is not None else _abi.default_settings(substeps=first.substeps)
        self.h = self.lib.tfw_create(self.settings.ctypes.data_as(_vp))
        self.dt = first.dt
        self.body_start = [0]
        for g, sc in enumerate(worlds):
            base = self.body_start[-1]
            for b in sc.bodies:
                i = self.lib.tfw_add_body(self.h, b["type"], _f(b["pos"]), _f(b["rot"]), _f(b["v"]), _f(b["w"]),
                                          b["locks"], _f(b["lock_frame"]), 1 if b["awake"] else 0,
                                          b["linear_damping"], b["angular_damping"], b["gravity_scale"])
                if len(worlds) > 1:
                    self.lib.tfw_set_group(self.h, i, g)
            for sh in sc.shapes:
                dims = sh["dims"] if len(sh["dims"]) else np.zeros(1, np.float32)
                if self.lib.tfw_add_shape(self.h, base + sh["body"], sh["kind"], _f(dims), len(sh["dims"]),
                                          _f(sh["pos"]), _f(sh["rot"]), sh["density"], sh["friction"],
                                          sh["restitution"], sh["rolling"]) < 0:
                    raise ValueError("the host world takes one box, sphere or half-space per body, at its origin")
            self.body_start.append(base + len(sc.bodies))
        self.worlds = len(worlds)
        # the native refresh callback, handed straight to the solver
        self.refresh_fn = REFRESH_FN(("tfw_refresh", self.lib))

    def close(self):
        if getattr(self, "h", None):
            self.lib.tfw_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def bodies(self):
        n = self.lib.tfw_bodies(self.h, None, 0)
        out = np.zeros(n, BODY_DTYPE)
        self.lib.tfw_bodies(self.h, out.ctypes.data_as(_vp), n)
        return out

    def body_shapes(self):
        """The bodies' shapes (SHAPE_DTYPE) for the device refresh, or None if a body's shape is a
        cylinder or a prism, which only this host world collides."""
        n = self.lib.tfw_body_shapes(self.h, None, 0)
        if n < 0:
            return None
        out = np.zeros(n, SHAPE_DTYPE)
        self.lib.tfw_body_shapes(self.h, out.ctypes.data_as(_vp), n)
        return out

    def collide(self, bodies, dt, prev=None):
        """Contacts at the bodies' poses, in key order, impulses carried from prev. May wake bodies in
        place; returns (contacts, woken)."""
        prev = np.zeros(0, CONTACT_DTYPE) if prev is None else np.ascontiguousarray(prev, CONTACT_DTYPE)
        woke = ctypes.c_uint32()
        cap = max(4 * len(bodies), 1024)
        while True:
            out = np.zeros(cap, CONTACT_DTYPE)
            n = self.lib.tfw_collide(self.h, bodies.ctypes.data_as(_vp), len(bodies), np.float32(dt),
                                     prev.ctypes.data_as(_vp), len(prev), out.ctypes.data_as(_vp), cap,
                                     ctypes.byref(woke))
            if n <= cap:
                return out[:n], woke.value
            cap = n


    def pairs(self, bodies, dt, pinned=None):
        """The contact stream's input at the bodies' poses: candidate pairs (PAIR_DTYPE, key order) and
        their manifolds (MANIFOLD_DTYPE). May wake bodies in place; returns (pairs, manifolds, woken).
        `pinned` (a dict this call fills and reuses) keeps the two arrays in pinned host memory, from
        which the stream's upload goes by DMA."""
        nm, woke = ctypes.c_uint32(), ctypes.c_uint32()
        cap = max(4 * len(bodies), 1024)
        while True:
            if pinned is not None:
                if pinned.get("cap", 0) < cap:
                    from .solver import PinnedArray
                    for k in ("pairs", "manifolds"):
                        if k in pinned:
                            pinned[k].close()
                    pinned["pairs"] = PinnedArray(PAIR_DTYPE, cap)
                    pinned["manifolds"] = PinnedArray(MANIFOLD_DTYPE, cap)
                    pinned["cap"] = cap
                pairs, mf = pinned["pairs"].view(cap), pinned["manifolds"].view(cap)
            else:
                pairs = np.zeros(cap, PAIR_DTYPE)
                mf = np.zeros(cap, MANIFOLD_DTYPE)
            n = self.lib.tfw_pairs(self.h, bodies.ctypes.data_as(_vp), len(bodies), np.float32(dt),
                                   pairs.ctypes.data_as(_vp), mf.ctypes.data_as(_vp), cap, ctypes.byref(nm),
                                   ctypes.byref(woke))
            if n <= cap:
                return pairs[:n], mf[:nm.value], woke.value
            cap = n


class Simulation:
    """A scene (or a list of scenes: independent worlds solved together) stepped with the host
    narrowphase and the GPU solver."""

    def __init__(self, scene, settings=None, device=0, device_refresh=False, stream=False):
        """`stream`: the contacts stay on the device from step to step; each step sends the candidate
        pairs and their manifolds (tfs_stream_contacts) instead of the contacts."""
        self.stream = stream
        self.pinned = {}
        self.world = HostWorld(scene, settings)
        self.settings = self.world.settings
        self.dt = self.world.dt
        self.bodies = self.world.bodies()
        self.contacts = np.zeros(0, CONTACT_DTYPE)
        self.solver = Solver(len(self.bodies), max(4 * len(self.bodies), 64), device=device, settings=self.settings)
        self.upload(self.bodies, self.contacts)
        self.steps = 0
        # device_refresh: mid-step refreshes collide on the device with the same narrowphase
        # (spheres, boxes and half-spaces only; None keeps the host callback)
        self.device_refresh = self.world.body_shapes() if device_refresh else None
        if device_refresh and self.device_refresh is None:
            raise SolverError("the device refresh collides spheres, boxes and half-spaces only")
        self.solver.set_body_shapes(self.device_refresh)

    def upload(self, bodies, contacts):
        if self.world.worlds > 1:
            self.solver.upload_worlds(bodies, self.world.body_start, contacts)
        else:
            self.solver.upload(bodies, contacts)

    def collide(self):
        """The narrowphase for the next step; uploads its contacts (and bodies it woke), or streams its
        pairs and manifolds onto the resident contacts."""
        self.bodies = self.solver.bodies() if self.steps else self.bodies
        if self.stream and self.world.worlds == 1:
            pairs, manifolds, woke = self.world.pairs(self.bodies, self.dt, self.pinned)
            self.solver.stream(pairs, manifolds)
            if woke:
                self.solver.upload_bodies(self.bodies)
            return
        contacts, woke = self.world.collide(self.bodies, self.dt, self.contacts)
        if woke or not self.steps:
            self.upload(self.bodies, contacts)
        else:
            self.solver.upload_contacts(contacts)
        self.contacts = contacts

    def solve(self):
        """The solve block and sleep pass on the GPU, over the uploaded contacts."""
        self.solver.step_native(self.dt, self.world.refresh_fn, self.world.h)
        if not self.stream:
            self.contacts = self.solver.contacts()
        self.steps += 1

    def step(self):
        self.collide()
        self.solve()

    def close(self):
        """Releases the solver's device state, the pinned buffers and the host world."""
        self.solver.close()
        for k in ("pairs", "manifolds"):
            if k in self.pinned:
                self.pinned[k].close()
        self.pinned = {}
        self.world.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
