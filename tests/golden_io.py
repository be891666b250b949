"""Reads the golden fixtures under tests/golden/ (recorded from the reference by
tests/golden/make_fixtures.py): the solve block's input and output on selected steps."""
import glob
import os

import numpy as np

from twofold_amd import _abi

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
LISTS = ("active", "fast", "fast_solid", "impact", "rows", "rest", "impact_rows", "impact_rest")


def save(path, scene, step, rec, settings):
    arrays = dict(scene=np.array(scene), step=np.array(step), dt=np.array(rec.dt, np.float32), settings=settings,
                  pre_bodies=rec.pre_bodies, pre_contacts=rec.pre_contacts, post_bodies=rec.post_bodies,
                  post_contacts=rec.post_contacts, step_bodies=rec.step_bodies, colors=rec.colors,
                  n_refresh=np.array(len(rec.refreshes)))
    for k in LISTS:
        arrays["list_" + k] = rec.lists[k]
    for i, r in enumerate(rec.refreshes):
        arrays[f"refresh{i}_substep"] = np.array(r["substep"])
        arrays[f"refresh{i}_remaining"] = np.array(r["remaining"], np.float32)
        arrays[f"refresh{i}_again"] = r["again"]
        arrays[f"refresh{i}_fresh"] = r["fresh"]
    np.savez_compressed(path, **arrays)


def _settings(old):
    """Settings as the current layout has them: fields the fixture predates take their defaults."""
    s = _abi.default_settings()
    for name in old.dtype.names:
        s[name] = old[name]
    return s


def _contacts(a):
    """Contacts as the current layout has them. Fixtures recorded before the `touched` byte took one of
    the two pad bytes hold zero there: the same bytes, read with the current field names."""
    if a.dtype == _abi.CONTACT_DTYPE:
        return a
    assert a.dtype.itemsize == _abi.CONTACT_DTYPE.itemsize
    return np.frombuffer(np.ascontiguousarray(a).tobytes(), _abi.CONTACT_DTYPE).copy()


def _case(path):
    z = np.load(path, allow_pickle=False)
    case = dict(scene=str(z["scene"]), step=int(z["step"]), dt=np.float32(z["dt"]), settings=_settings(z["settings"]),
                pre_bodies=z["pre_bodies"], pre_contacts=_contacts(z["pre_contacts"]), post_bodies=z["post_bodies"],
                post_contacts=_contacts(z["post_contacts"]), step_bodies=z["step_bodies"], colors=z["colors"],
                lists={k: z["list_" + k] for k in LISTS}, refreshes=[])
    for i in range(int(z["n_refresh"])):
        case["refreshes"].append(dict(substep=int(z[f"refresh{i}_substep"]),
                                      remaining=np.float32(z[f"refresh{i}_remaining"]),
                                      again=z[f"refresh{i}_again"], fresh=_contacts(z[f"refresh{i}_fresh"])))
    return case


def names():
    return sorted({os.path.basename(p).rsplit("-step", 1)[0] for p in glob.glob(os.path.join(HERE, "*.npz"))})


def load(scene):
    paths = sorted(glob.glob(os.path.join(HERE, f"{scene}-step*.npz")))
    return [_case(p) for p in paths]


def all_cases():
    return [c for n in names() for c in load(n)]
