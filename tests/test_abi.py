"""The C ABI library loads here (no GPU needed) and exports what include/twofold_solver.h declares,
with struct layouts equal to the numpy mirrors; no compute call is made."""
import ctypes
import os
import re

import numpy as np

from twofold_amd import _abi
from twofold_amd import solver

HEADER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include", "twofold_solver.h")


def declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:const\s+)?[a-z_0-9]+\s*\*?\s*(tfs_[a-z_0-9]+)\s*\(", text, flags=re.M)))


def test_library_exports_every_declared_function():
    lib = solver.load_library()
    names = declared()
    assert "tfs_step" in names and "tfs_create" in names
    for name in names:
        assert hasattr(lib, name), name
    assert set(solver.EXPORTS) <= set(names)


def test_layouts_match():
    lib = solver.load_library()
    assert lib.tfs_sizeof_body() == _abi.BODY_DTYPE.itemsize
    assert lib.tfs_sizeof_contact() == _abi.CONTACT_DTYPE.itemsize
    assert lib.tfs_sizeof_settings() == _abi.SETTINGS_DTYPE.itemsize


def test_default_settings_are_the_reference_defaults():
    lib = solver.load_library()
    s = np.zeros((), _abi.SETTINGS_DTYPE)
    lib.tfs_default_settings(s.ctypes.data)
    want = _abi.default_settings()
    for name in _abi.SETTINGS_DTYPE.names:
        np.testing.assert_array_equal(s[name], want[name], err_msg=name)


def test_default_settings_are_the_compiled_references():
    """tfs_default_settings against a default WorldSettings and the Tolerances a World derives from it
    (include/twofold/world.h:216-241, src/world.cpp:224-225), read from the reference built here."""
    import refworld

    if not refworld.ref_available():
        import pytest

        pytest.skip("reference checker not built (make -C oracle ref)")
    ref = refworld._load_ref()
    ref.tfr_default_settings.argtypes = [ctypes.c_void_p]
    want = np.zeros((), _abi.SETTINGS_DTYPE)
    ref.tfr_default_settings(want.ctypes.data)
    got = np.zeros((), _abi.SETTINGS_DTYPE)
    solver.load_library().tfs_default_settings(got.ctypes.data)
    for name in _abi.SETTINGS_DTYPE.names:
        np.testing.assert_array_equal(got[name], want[name], err_msg=name)


def test_argument_errors_without_a_device():
    lib = solver.load_library()
    assert lib.tfs_create(0, 1, 1, None) == -1  # TFS_E_ARG, checked before any device call
    assert lib.tfs_last_error(None) == b"no context"
    assert lib.tfs_abi_version() == solver.ABI_VERSION
