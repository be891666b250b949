"""Host-side binding of the MI355X contact solver (libtwofold_hip.so, include/twofold_solver.h).

`Solver` mirrors the part of twofold::World that the solve path owns (src/world.cpp:1564-1727): it
takes the bodies and the contact manifolds the narrowphase produced, runs one step of the solve
block and the sleep pass on the GPU, and hands back bodies, contacts and the integer results
(lists, colours, row order). Errors raise SolverError with the library's message; there is no CPU
fallback.
"""
import atexit
import ctypes
import os
import sys
import weakref

import numpy as np

from . import _abi
from ._abi import (BODY_DTYPE, CONTACT_DTYPE, EXT_DTYPE, HIT_DTYPE, JOINT_DTYPE, MANIFOLD_DTYPE, PAIR_DTYPE, SETTINGS_DTYPE,
                   SHAPE_DTYPE)

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libtwofold_hip.so")

_vp = ctypes.c_void_p
ABI_VERSION = 3  # TFS_ABI_VERSION of include/twofold_solver.h this binding mirrors
REFRESH_FN = ctypes.CFUNCTYPE(ctypes.c_int, _vp, ctypes.c_int, ctypes.c_float, ctypes.c_uint32,
                              ctypes.POINTER(ctypes.c_uint32), _vp, _vp, _vp)

EXPORTS = ("tfs_abi_version", "tfs_sizeof_body", "tfs_sizeof_contact", "tfs_sizeof_settings", "tfs_default_settings",
           "tfs_create", "tfs_destroy", "tfs_last_error", "tfs_set_settings", "tfs_upload_bodies",
           "tfs_upload_contacts", "tfs_set_worlds", "tfs_step", "tfs_download_bodies", "tfs_download_contacts", "tfs_download_list",
           "tfs_last_step_ms", "tfs_last_step_counts", "tfs_set_profiling", "tfs_read_profile", "tfs_fixed_begin",
           "tfs_fixed_step", "tfs_adjoint_seed", "tfs_adjoint_step", "tfs_adjoint_read", "tfs_check_minmax",
           "tfs_autotune", "tfs_pass_shape", "tfs_set_pass_shape", "tfs_set_cooperative", "tfs_set_body_shapes",
           "tfs_step_noalloc", "tfs_download_touches", "tfs_set_graph", "tfs_stream_contacts", "tfs_upload_contact_ext",
           "tfs_download_contact_ext", "tfs_download_poses", "tfs_set_step_sleep", "tfs_sleep", "tfs_upload_joints",
           "tfs_download_joints", "tfs_sizeof_joint")


class SolverError(RuntimeError):
    pass


_lib = None


def load_library(path=LIB_PATH):
    """Loads the HIP library; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise SolverError(f"{path} is missing: build it with __graft_entry__.build() (hipcc, gfx950)")
    lib = ctypes.CDLL(path)
    lib.tfs_abi_version.restype = ctypes.c_int
    for name in ("tfs_sizeof_body", "tfs_sizeof_contact", "tfs_sizeof_settings"):
        getattr(lib, name).restype = ctypes.c_size_t
    lib.tfs_default_settings.argtypes = [_vp]
    lib.tfs_create.argtypes = [ctypes.c_int, ctypes.c_uint32, ctypes.c_uint32, ctypes.POINTER(_vp)]
    lib.tfs_destroy.argtypes = [_vp]
    lib.tfs_last_error.restype = ctypes.c_char_p
    lib.tfs_last_error.argtypes = [_vp]
    lib.tfs_set_settings.argtypes = [_vp, _vp]
    lib.tfs_upload_bodies.argtypes = [_vp, _vp, ctypes.c_uint32]
    lib.tfs_upload_contacts.argtypes = [_vp, _vp, ctypes.c_uint32]
    lib.tfs_set_worlds.argtypes = [_vp, ctypes.c_uint32, _vp]
    lib.tfs_step.argtypes = [_vp, ctypes.c_float, REFRESH_FN, _vp]
    lib.tfs_download_bodies.argtypes = [_vp, _vp, ctypes.c_uint32]
    lib.tfs_download_contacts.argtypes = [_vp, _vp, ctypes.c_uint32]
    lib.tfs_download_list.argtypes = [_vp, ctypes.c_int, _vp, ctypes.c_uint32, ctypes.POINTER(ctypes.c_uint32)]
    lib.tfs_last_step_ms.argtypes = [_vp, ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_float)]
    lib.tfs_last_step_counts.argtypes = [_vp, _vp]
    lib.tfs_set_profiling.argtypes = [_vp, ctypes.c_int]
    lib.tfs_fixed_begin.argtypes = [_vp, ctypes.c_float, ctypes.c_uint32]
    lib.tfs_fixed_step.argtypes = [_vp]
    lib.tfs_adjoint_seed.argtypes = [_vp, _vp]
    lib.tfs_adjoint_step.argtypes = [_vp]
    lib.tfs_adjoint_read.argtypes = [_vp, _vp, _vp, _vp]
    lib.tfs_check_minmax.argtypes = [ctypes.c_int, _vp, _vp, ctypes.c_uint32, _vp]
    lib.tfs_autotune.argtypes = [_vp, ctypes.c_uint32]
    lib.tfs_pass_shape.argtypes = [_vp, _vp]
    lib.tfs_set_pass_shape.argtypes = [_vp, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_int, ctypes.c_int]
    lib.tfs_set_graph.argtypes = [_vp, ctypes.c_int]
    lib.tfs_stream_contacts.argtypes = [_vp, _vp, ctypes.c_uint32, _vp, ctypes.c_uint32, ctypes.POINTER(ctypes.c_uint32)]
    lib.tfs_upload_contact_ext.argtypes = [_vp, _vp, ctypes.c_uint32]
    lib.tfs_download_poses.argtypes = [_vp, _vp, ctypes.c_uint32]
    lib.tfs_download_contact_ext.argtypes = [_vp, _vp, ctypes.c_uint32]
    lib.tfs_set_cooperative.argtypes = [_vp, ctypes.c_int]
    lib.tfs_set_step_sleep.argtypes = [_vp, ctypes.c_int]
    lib.tfs_sleep.argtypes = [_vp, ctypes.c_float]
    lib.tfs_set_body_shapes.argtypes = [_vp, ctypes.c_uint32, _vp]
    lib.tfs_upload_joints.argtypes = [_vp, _vp, ctypes.c_uint32, _vp]
    lib.tfs_download_joints.argtypes = [_vp, _vp, ctypes.c_uint32, _vp]
    lib.tfs_download_touches.argtypes = [_vp, _vp, ctypes.c_uint32, ctypes.POINTER(ctypes.c_uint32), _vp,
                                         ctypes.c_uint32, ctypes.POINTER(ctypes.c_uint32), ctypes.POINTER(ctypes.c_uint32)]
    lib.tfs_step_noalloc.argtypes = [ctypes.c_int, _vp, ctypes.c_uint32, _vp, ctypes.c_uint32, _vp, _vp, ctypes.c_float,
                                     _vp]
    lib.tfs_read_profile.argtypes = [_vp, ctypes.c_int, ctypes.POINTER(ctypes.c_uint64), ctypes.POINTER(ctypes.c_double),
                                     ctypes.POINTER(ctypes.c_double)]
    if lib.tfs_abi_version() != ABI_VERSION:
        raise SolverError("ABI version mismatch")
    _abi.check_sizes(lib, "tfs")
    if lib.tfs_sizeof_settings() != SETTINGS_DTYPE.itemsize:
        raise SolverError("tfs_settings layout mismatch")
    _lib = lib
    return lib


def device_minmax(a, b, device=0):
    """The solves' device max and min of float32 pairs: (max, min) arrays (tfs_check_minmax)."""
    lib = load_library()
    a = np.ascontiguousarray(a, np.float32)
    b = np.ascontiguousarray(b, np.float32)
    if a.shape != b.shape or a.ndim != 1:
        raise SolverError("a and b must be 1-D arrays of one length")
    out = np.zeros(2 * len(a), np.float32)
    r = lib.tfs_check_minmax(device, _ptr(a), _ptr(b), len(a), _ptr(out))
    if r != 0:
        raise SolverError(f"tfs_check_minmax failed ({r})")
    return out[0::2].copy(), out[1::2].copy()


def _ptr(a):
    return a.ctypes.data_as(_vp)


_hip = None


def hip_runtime():
    """The HIP runtime the library itself links (libamdhip64 from ROCm), loaded once. Callers that need
    device memory or a device-wide sync beside the library use this one: a second HIP runtime in the
    process (torch bundles its own) may not see the GPU once the first is up."""
    global _hip
    if _hip is None:
        load_library()
        hip = ctypes.CDLL("libamdhip64.so.7")
        hip.hipMalloc.argtypes = [ctypes.POINTER(_vp), ctypes.c_size_t]
        hip.hipFree.argtypes = [_vp]
        hip.hipMemcpy.argtypes = [_vp, _vp, ctypes.c_size_t, ctypes.c_int]
        hip.hipDeviceSynchronize.argtypes = []
        _hip = hip
    return _hip


def device_synchronize():
    """Waits for all work on the current device (hipDeviceSynchronize)."""
    rc = hip_runtime().hipDeviceSynchronize()
    if rc != 0:
        raise SolverError(f"hipDeviceSynchronize failed ({rc})")


class PinnedArray:
    """A numpy structured array in pinned host memory (hipHostMalloc), which HIP copies to and from the
    device by DMA at full speed; `view(n)` is its first n records."""

    def __init__(self, dtype, capacity):
        self.hip = hip_runtime()
        self.hip.hipHostMalloc.argtypes = [ctypes.POINTER(_vp), ctypes.c_size_t, ctypes.c_uint]
        self.hip.hipHostFree.argtypes = [_vp]
        self.dtype = np.dtype(dtype)
        self.capacity = max(int(capacity), 1)
        self.ptr = _vp()
        rc = self.hip.hipHostMalloc(ctypes.byref(self.ptr), self.capacity * self.dtype.itemsize, 0)
        if rc != 0:
            raise SolverError(f"hipHostMalloc failed ({rc})")
        raw = (ctypes.c_char * (self.capacity * self.dtype.itemsize)).from_address(self.ptr.value)
        self.array = np.frombuffer(raw, self.dtype)

    def view(self, n):
        return self.array[:n]

    def close(self):
        if self.ptr:
            self.array = None
            self.hip.hipHostFree(self.ptr)
            self.ptr = _vp()

    def __del__(self):
        if not sys.is_finalizing():
            self.close()


class DeviceArray:
    """A numpy structured array copied into device memory of its own (hipMalloc), for the entry points
    that take device pointers; `get()` copies it back."""

    def __init__(self, a):
        self.hip = hip_runtime()
        a = np.ascontiguousarray(a)
        self.dtype, self.n = a.dtype, len(a)
        self.ptr = _vp()
        rc = self.hip.hipMalloc(ctypes.byref(self.ptr), max(a.nbytes, 1))
        if rc != 0:
            raise SolverError(f"hipMalloc failed ({rc})")
        if a.nbytes:
            self._copy(self.ptr, a.ctypes.data, a.nbytes, 1)

    def _copy(self, dst, src, n, kind):
        rc = self.hip.hipMemcpy(dst, src, n, kind)
        if rc != 0:
            raise SolverError(f"hipMemcpy failed ({rc})")

    def get(self):
        out = np.zeros(self.n, self.dtype)
        if out.nbytes:
            self._copy(out.ctypes.data, self.ptr, out.nbytes, 2)
        return out

    def close(self):
        if self.ptr:
            self.hip.hipFree(self.ptr)
            self.ptr = _vp()

    def __del__(self):
        if not sys.is_finalizing():
            self.close()


# Solvers still open. Their device memory is released before the interpreter tears anything down:
# at interpreter exit an object's __del__ can run after other libraries (torch, a profiler) have
# already shut the HIP runtime down, and a hipFree then crashes inside exit().
_live = weakref.WeakSet()


@atexit.register
def close_all():
    """Closes every solver still open (run at interpreter exit, before module teardown)."""
    for s in list(_live):
        s.close()


class Solver:
    """One world's solver state, resident on a GPU."""

    def __init__(self, max_bodies=1, max_contacts=1, device=0, settings=None):
        self.lib = load_library()
        h = _vp()
        rc = self.lib.tfs_create(device, max_bodies, max_contacts, ctypes.byref(h))
        if rc != 0:
            raise SolverError(f"tfs_create failed ({rc}); is a GPU visible?")
        self.h = h
        self.nb = 0
        self.nc = 0
        self._cb = None
        _live.add(self)
        if settings is not None:
            self.set_settings(settings)

    def close(self):
        """Releases the device state (tfs_destroy). Safe to call more than once."""
        if getattr(self, "h", None):
            self.lib.tfs_destroy(self.h)
            self.h = None
        _live.discard(self)

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        # during interpreter shutdown the runtime may already be gone: leave the memory to the
        # process's end rather than call into it (close_all has closed every solver before that)
        if sys.is_finalizing():
            return
        self.close()

    def _check(self, rc, what):
        if rc != 0:
            msg = self.lib.tfs_last_error(self.h).decode(errors="replace")
            raise SolverError(f"{what} failed ({rc}): {msg}")

    def set_settings(self, settings):
        s = np.ascontiguousarray(np.asarray(settings, SETTINGS_DTYPE).reshape(()))
        self._check(self.lib.tfs_set_settings(self.h, _ptr(s)), "tfs_set_settings")

    def upload(self, bodies, contacts):
        b = np.ascontiguousarray(bodies, BODY_DTYPE)
        c = np.ascontiguousarray(contacts, CONTACT_DTYPE)
        self._check(self.lib.tfs_upload_bodies(self.h, _ptr(b), len(b)), "tfs_upload_bodies")
        self._check(self.lib.tfs_upload_contacts(self.h, _ptr(c), len(c)), "tfs_upload_contacts")
        self.nb, self.nc = len(b), len(c)

    def upload_bodies(self, bodies):
        """Bodies only: the resident contacts stay (the contact stream's path)."""
        b = np.ascontiguousarray(bodies, BODY_DTYPE)
        self._check(self.lib.tfs_upload_bodies(self.h, _ptr(b), len(b)), "tfs_upload_bodies")
        self.nb = len(b)

    def set_worlds(self, body_start):
        """Independent worlds over the resident bodies (tfs_set_worlds); upload contacts after."""
        bs = np.ascontiguousarray(body_start, np.uint32)
        self._check(self.lib.tfs_set_worlds(self.h, len(bs) - 1, _ptr(bs)), "tfs_set_worlds")
        self.nc = 0

    def upload_worlds(self, bodies, body_start, contacts):
        """Independent worlds: their bodies and contacts concatenated world after world, the contacts'
        body indices global (tfs_upload_bodies, tfs_set_worlds, tfs_upload_contacts)."""
        b = np.ascontiguousarray(bodies, BODY_DTYPE)
        self._check(self.lib.tfs_upload_bodies(self.h, _ptr(b), len(b)), "tfs_upload_bodies")
        self.nb = len(b)
        self.set_worlds(body_start)
        self.upload_contacts(contacts)

    def upload_contacts(self, contacts):
        c = np.ascontiguousarray(contacts, CONTACT_DTYPE)
        self._check(self.lib.tfs_upload_contacts(self.h, _ptr(c), len(c)), "tfs_upload_contacts")
        self.nc = len(c)

    def step_native(self, dt, fn, user):
        """One step with a native refresh callback (a REFRESH_FN pointer and its user pointer)."""
        self._check(self.lib.tfs_step(self.h, ctypes.c_float(dt), fn, user), "tfs_step")

    def step(self, dt, refresh=None):
        """One step. `refresh(substep, remaining, again, contacts, bodies) -> fresh contacts` is called
        when fast contacts need new manifolds mid-step: `again` the contacts' indices, `contacts` those
        contacts and `bodies` their two bodies each (2k, 2k + 1), as tfs_refresh_fn hands them."""
        if refresh is None:
            cb = REFRESH_FN()
        else:
            def cb_impl(user, substep, remaining, n_again, again, contacts, bodies, fresh):
                try:
                    ag = np.ctypeslib.as_array(again, shape=(n_again,)).copy()
                    ct = np.frombuffer((ctypes.c_char * (n_again * CONTACT_DTYPE.itemsize)).from_address(contacts),
                                       CONTACT_DTYPE).copy()
                    bd = np.frombuffer((ctypes.c_char * (2 * n_again * BODY_DTYPE.itemsize)).from_address(bodies),
                                       BODY_DTYPE).copy()
                    out = np.ascontiguousarray(refresh(substep, remaining, ag, ct, bd), CONTACT_DTYPE)
                    if len(out) != n_again:
                        return -1
                    ctypes.memmove(fresh, out.ctypes.data, out.nbytes)
                    return 0
                except Exception:  # an exception must not cross the C boundary
                    return -1
            cb = REFRESH_FN(cb_impl)
        self._cb = cb
        self._check(self.lib.tfs_step(self.h, ctypes.c_float(dt), cb, None), "tfs_step")

    def poses(self):
        """The bodies' poses and velocities, compact (POSE_DTYPE, tfs_download_poses)."""
        out = np.zeros(self.nb, _abi.POSE_DTYPE)
        self._check(self.lib.tfs_download_poses(self.h, _ptr(out), self.nb), "tfs_download_poses")
        return out

    def bodies(self):
        out = np.zeros(self.nb, BODY_DTYPE)
        self._check(self.lib.tfs_download_bodies(self.h, _ptr(out), self.nb), "tfs_download_bodies")
        return out

    def contacts(self):
        out = np.zeros(self.nc, CONTACT_DTYPE)
        self._check(self.lib.tfs_download_contacts(self.h, _ptr(out), self.nc), "tfs_download_contacts")
        return out

    def list(self, kind):
        n = ctypes.c_uint32()
        self._check(self.lib.tfs_download_list(self.h, kind, None, 0, ctypes.byref(n)), "tfs_download_list")
        out = np.zeros(n.value, np.uint8 if kind == _abi.LIST_COLORS else np.uint32)
        self._check(self.lib.tfs_download_list(self.h, kind, _ptr(out), n.value, ctypes.byref(n)), "tfs_download_list")
        return out

    def lists(self):
        """The integer results of the last step, named as the oracle names them."""
        k = _abi
        rows, imp_rows = self.list(k.LIST_ROWS), self.list(k.LIST_IMPACT_ROWS)
        colors = self.list(k.LIST_COLORS)
        rest = rows[colors[rows] >= k.COLORS] if len(rows) else rows
        imp_rest = imp_rows[colors[imp_rows] >= k.COLORS] if len(imp_rows) else imp_rows
        return dict(active=self.list(k.LIST_ACTIVE), fast=self.list(k.LIST_FAST), fast_solid=self.list(k.LIST_FAST_SOLID),
                    impact=self.list(k.LIST_IMPACT), rows=rows, rest=rest, impact_rows=imp_rows, impact_rest=imp_rest,
                    colors=colors)

    def last_ms(self):
        a, b = ctypes.c_float(), ctypes.c_float()
        self._check(self.lib.tfs_last_step_ms(self.h, ctypes.byref(a), ctypes.byref(b)), "tfs_last_step_ms")
        return a.value, b.value

    def set_profiling(self, on):
        self._check(self.lib.tfs_set_profiling(self.h, 1 if on else 0), "tfs_set_profiling")

    def profile(self, family):
        """(launches, milliseconds, algorithmic bytes) summed since profiling was turned on; family 0 the
        substep start (velocities integrated, every row's warm start), 1 solve, 2 restitution (event-timed); 3 waiting for the refresh's contacts on the host,
        4 the refresh callback (host clock; bytes = pairs collided again); 5 the solve passes over the
        rest lists (event-timed); 6 the reverse sweep's passes of tfs_adjoint_step (event-timed)."""
        n, ms, by = ctypes.c_uint64(), ctypes.c_double(), ctypes.c_double()
        self._check(self.lib.tfs_read_profile(self.h, family, ctypes.byref(n), ctypes.byref(ms), ctypes.byref(by)),
                    "tfs_read_profile")
        return n.value, ms.value, by.value

    # ---- pass shape and its autotuning ----

    def autotune(self, rounds):
        """Tunes the pass shape over the next rounds x candidates steps (0 stops tuning)."""
        self._check(self.lib.tfs_autotune(self.h, rounds), "tfs_autotune")

    def pass_shape(self):
        """The pass shape in use: dict of blocks, tail_rows (0 = no tail), made_axes, tuning_steps_left,
        candidates (0 when not tuning)."""
        out = np.zeros(7, np.uint32)
        self._check(self.lib.tfs_pass_shape(self.h, _ptr(out)), "tfs_pass_shape")
        return dict(blocks=int(out[0]), tail_rows=int(out[1]), made_axes=bool(out[2]), tuning_steps_left=int(out[3]),
                    candidates=int(out[4]), fused=bool(out[5]), graph=bool(out[6]))

    def set_cooperative(self, on):
        """Launch the grid-barrier passes cooperatively (True) or as plain kernels (False, the default)."""
        self._check(self.lib.tfs_set_cooperative(self.h, int(bool(on))), "tfs_set_cooperative")

    def set_step_sleep(self, on):
        """Whether step() runs the sleep pass after the solve block (True, the default)."""
        self._check(self.lib.tfs_set_step_sleep(self.h, int(bool(on))), "tfs_set_step_sleep")

    def sleep(self, dt):
        """The sleep pass alone (updateSleep) over the resident bodies and contacts."""
        self._check(self.lib.tfs_sleep(self.h, np.float32(dt)), "tfs_sleep")

    def set_body_shapes(self, shapes):
        """One shape per resident body (SHAPE_DTYPE): every refresh then collides its pairs on the
        device with the reference's sphere/box/half-space narrowphase (collide3d.h), and no callback is
        called. None or an empty array turns it off."""
        if shapes is None or len(shapes) == 0:
            self._check(self.lib.tfs_set_body_shapes(self.h, 0, None), "tfs_set_body_shapes")
            return
        shapes = np.ascontiguousarray(shapes, SHAPE_DTYPE)
        self._check(self.lib.tfs_set_body_shapes(self.h, len(shapes), shapes.ctypes.data_as(_vp)), "tfs_set_body_shapes")

    def upload_joints(self, joints, world_motion=None):
        """The joints (JOINT_DTYPE, the reference's order) and the world body's velocities (v, w; None for
        zero); every step then solves them with the contacts (tfs_upload_joints). Upload the bodies
        first."""
        joints = np.ascontiguousarray(joints if joints is not None else np.zeros(0, JOINT_DTYPE), JOINT_DTYPE)
        wm = None if world_motion is None else np.ascontiguousarray(world_motion, np.float32)
        self._check(self.lib.tfs_upload_joints(self.h, joints.ctypes.data_as(_vp), len(joints),
                                               None if wm is None else wm.ctypes.data_as(_vp)), "tfs_upload_joints")
        self.nj = len(joints)

    def joints(self):
        """The resident joints and the world body's velocities (tfs_download_joints)."""
        n = getattr(self, "nj", 0)
        out = np.zeros(n, JOINT_DTYPE)
        wm = np.zeros(6, np.float32)
        self._check(self.lib.tfs_download_joints(self.h, out.ctypes.data_as(_vp), n, wm.ctypes.data_as(_vp)),
                    "tfs_download_joints")
        return out, wm

    def set_pass_shape(self, blocks, tail_rows, made_axes, fused=False):
        self._check(self.lib.tfs_set_pass_shape(self.h, blocks, tail_rows, int(bool(made_axes)), int(bool(fused))),
                    "tfs_set_pass_shape")

    def set_graph(self, on):
        """Launch each step's substep loop as one captured HIP graph (when no host refresh is needed)."""
        self._check(self.lib.tfs_set_graph(self.h, int(bool(on))), "tfs_set_graph")

    # ---- the fixed-contact model and its adjoint ----

    def fixed_begin(self, dt, max_steps):
        """Fixes the resident contacts for the fixed-contact model, with room for max_steps steps."""
        self._check(self.lib.tfs_fixed_begin(self.h, ctypes.c_float(dt), max_steps), "tfs_fixed_begin")

    def fixed_step(self):
        """Checkpoints the state, then takes one step of the fixed-contact model."""
        self._check(self.lib.tfs_fixed_step(self.h), "tfs_fixed_step")

    def adjoint_seed(self, dl_dcom):
        """d loss / d com after the last step (n x 3); clears the other adjoints."""
        g = np.ascontiguousarray(dl_dcom, np.float32).reshape(self.nb, 3)
        self._check(self.lib.tfs_adjoint_seed(self.h, _ptr(g)), "tfs_adjoint_seed")

    def adjoint_step(self):
        """Reverses the last step not yet reversed."""
        self._check(self.lib.tfs_adjoint_step(self.h), "tfs_adjoint_step")

    def adjoint_read(self):
        """(d loss / d v, d loss / d w) per body (n x 3 each) at the start of the earliest step reversed,
        and d loss / d friction per contact."""
        gv = np.zeros((self.nb, 3), np.float32)
        gw = np.zeros((self.nb, 3), np.float32)
        gm = np.zeros(max(self.nc, 1), np.float32)
        self._check(self.lib.tfs_adjoint_read(self.h, _ptr(gv), _ptr(gw), _ptr(gm)), "tfs_adjoint_read")
        return gv, gw, gm[:self.nc]

    def stream(self, pairs, manifolds):
        """The contact stream (tfs_stream_contacts): this step's candidate pairs (PAIR_DTYPE, key order) and
        their manifolds (MANIFOLD_DTYPE) against the resident contacts; the result stays resident."""
        p = np.ascontiguousarray(pairs, PAIR_DTYPE)
        m = np.ascontiguousarray(manifolds, MANIFOLD_DTYPE)
        n = ctypes.c_uint32()
        self._check(self.lib.tfs_stream_contacts(self.h, _ptr(p), len(p), _ptr(m), len(m), ctypes.byref(n)),
                    "tfs_stream_contacts")
        self.nc = n.value

    def upload_ext(self, ext):
        e = np.ascontiguousarray(ext, EXT_DTYPE)
        self._check(self.lib.tfs_upload_contact_ext(self.h, _ptr(e), len(e)), "tfs_upload_contact_ext")

    def ext(self):
        out = np.zeros(self.nc, EXT_DTYPE)
        self._check(self.lib.tfs_download_contact_ext(self.h, _ptr(out), self.nc), "tfs_download_contact_ext")
        return out

    def touches(self):
        """Touch events of the last step (reportTouches, src/world.cpp:1424-1468): events as
        (first contact of the pair << 1) | began, hits (HIT_DTYPE), and the touching contacts' count."""
        ne, nh, touching = ctypes.c_uint32(), ctypes.c_uint32(), ctypes.c_uint32()
        self._check(self.lib.tfs_download_touches(self.h, None, 0, ctypes.byref(ne), None, 0, ctypes.byref(nh),
                                                  ctypes.byref(touching)), "tfs_download_touches")
        events = np.zeros(ne.value, np.uint32)
        hits = np.zeros(nh.value, HIT_DTYPE)
        self._check(self.lib.tfs_download_touches(self.h, _ptr(events), ne.value, ctypes.byref(ne), _ptr(hits),
                                                  nh.value, ctypes.byref(nh), ctypes.byref(touching)),
                    "tfs_download_touches")
        return events, hits, touching.value

    def counts(self):
        """active contacts, their points, impact contacts, their points, colours used, rest contacts, and
        the dependency levels of the active and the impact rest list."""
        out = np.zeros(8, np.uint32)
        self._check(self.lib.tfs_last_step_counts(self.h, _ptr(out)), "tfs_last_step_counts")
        return dict(zip(("active", "points", "impact", "impact_points", "colors", "rest", "rest_levels",
                         "impact_rest_levels"), out.tolist()))
