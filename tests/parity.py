"""Bit-level comparison of one solve step against what the reference recorded.

`want` holds the reference's record of the step (post_bodies: after the final speed limit;
step_bodies: after the whole step, sleep included; post_contacts; lists; colors). `got_*` is the
result of the path under test after one tfs_step / tfo_step on want's inputs: bodies after the
solve block and the sleep pass, contacts after the solve block.
"""
import numpy as np

LISTS = ("active", "fast", "fast_solid", "impact", "rows", "rest", "impact_rows", "impact_rest")
POSE = ("p", "q", "com", "dp", "dq", "kept")
MASS = ("inv_mass_world", "inv_inertia_world", "plane_inv_mass", "plane_inv_inertia")
STEP = ("v", "w", "island", "awake", "sleep_time")
CONTACT = ("normal", "tangent1", "tangent2", "rolling_impulse", "rolling_radius", "count", "planar", "laned")
POINT = ("anchor_a", "anchor_b", "adjusted_separation", "normal_impulse", "tangent_impulse", "max_normal_impulse",
         "relative_velocity", "local_a", "separation")


def _bits(a):
    a = np.ascontiguousarray(a)
    return a.view(np.uint32) if a.dtype == np.float32 else a


def _rows_differ(a, b):
    a, b = _bits(a), _bits(b)
    d = (a != b).reshape(len(a), -1).any(1) if len(a) else np.zeros(0, bool)
    return np.nonzero(d)[0]


def compare(want, got_bodies, got_contacts, got_lists):
    """Returns a list of mismatch descriptions (empty when everything is bit-identical)."""
    bad = []
    for k in LISTS:
        if not np.array_equal(np.asarray(want["lists"][k]), np.asarray(got_lists[k])):
            bad.append(f"list {k}: want {len(want['lists'][k])} entries, got {len(got_lists[k])}")
    act = want["lists"]["active"]
    if not np.array_equal(want["colors"][act], np.asarray(got_lists["colors"])[act]):
        bad.append("colours of active contacts")
    post, stepb = want["post_bodies"], want["step_bodies"]
    slept = (post["awake"] == 1) & (stepb["awake"] == 0)
    for f in POSE + MASS:
        ref = post[f].copy()
        if f in MASS:
            ref[slept] = stepb[f][slept]
        d = _rows_differ(ref, got_bodies[f])
        if len(d):
            bad.append(f"body {f} at {d[:5].tolist()}")
    for f in STEP:
        d = _rows_differ(stepb[f], got_bodies[f])
        if len(d):
            bad.append(f"body {f} after sleep at {d[:5].tolist()}")
    wc = want["post_contacts"]
    for f in CONTACT:
        d = _rows_differ(wc[f], got_contacts[f])
        if len(d):
            bad.append(f"contact {f} at {d[:5].tolist()}")
    for i in range(4):
        m = wc["count"] > i
        for f in POINT:
            d = _rows_differ(wc["points"][f][:, i][m], got_contacts["points"][f][:, i][m])
            if len(d):
                bad.append(f"point {i} {f} at {np.nonzero(m)[0][d][:5].tolist()}")
    return bad


def record_dict(rec):
    """A refworld.StepRecord as the dict compare() takes."""
    return dict(lists=rec.lists, colors=rec.colors, post_bodies=rec.post_bodies, step_bodies=rec.step_bodies,
                post_contacts=rec.post_contacts, pre_bodies=rec.pre_bodies, pre_contacts=rec.pre_contacts,
                refreshes=rec.refreshes, dt=rec.dt)


def invariants(pre_bodies, pre_contacts, contacts, bodies, lists):
    """Size-independent properties of a step's results, for any input size: the colouring is the
    sequential greedy one over moving bodies (src/lanes.cpp:356-374), the row order is colour-major
    with more points first (point counts as the step began; a refresh may change them later) and ties
    in list order (src/lanes.cpp:376-407), and every island is one connected component of touching
    awake dynamic bodies labelled by its lowest body index (src/world.cpp:1186-1237). `contacts` are
    the contacts after the step. Returns a list of violations."""
    bad = []
    moving = (pre_bodies["type"] == 2) & (pre_bodies["awake"] != 0) & (pre_bodies["alive"] != 0)
    colors = np.asarray(lists["colors"])
    act = np.asarray(lists["active"])
    taken = {}
    for ci in act:
        c = contacts[ci]
        if c["planar"] or c["rolling_resistance"] > 0:
            if colors[ci] < 32:
                bad.append(f"contact {ci}: planar or rolling but coloured")
            continue
        ends = [int(x) for x in (c["body_a"], c["body_b"]) if moving[x]]
        used = 0
        for e in ends:
            used |= taken.get(e, 0)
        want = next((k for k in range(32) if not used >> k & 1), 32)
        if colors[ci] != want:
            bad.append(f"contact {ci}: colour {colors[ci]}, greedy gives {want}")
            break
        if want < 32:
            for e in ends:
                taken[e] = taken.get(e, 0) | (1 << want)
    rows = np.asarray(lists["rows"])
    pos = {int(ci): k for k, ci in enumerate(act)}
    key = [(int(colors[ci]) if colors[ci] < 32 else 32, -int(pre_contacts[ci]["count"]) if colors[ci] < 32 else 0,
            pos[int(ci)]) for ci in rows]
    if key != sorted(key):
        bad.append("row order is not colour-major, points-descending, stable")
    n = len(bodies)
    awake = (bodies["type"] == 2) & (pre_bodies["awake"] != 0) & (bodies["alive"] != 0)
    parent = np.arange(n)

    def root(x):
        while parent[x] != x:
            parent[x] = parent[parent[x]]
            x = parent[x]
        return x

    for c in contacts:
        if c["count"] == 0 or not (awake[c["body_a"]] and awake[c["body_b"]]):
            continue
        ra, rb = root(int(c["body_a"])), root(int(c["body_b"]))
        if ra != rb:
            parent[max(ra, rb)] = min(ra, rb)
    labels = np.array([root(i) for i in range(n)])
    wrong = np.nonzero(awake & (bodies["island"] != labels))[0]
    if len(wrong):
        bad.append(f"island ids at {wrong[:5].tolist()}")
    return bad


def compare_touches(rec, got_contacts, events, hits, touching):
    """reportTouches (src/world.cpp:1424-1468) against the reference's step: every contact's touched flag
    after the step, the events (the pairs whose flag changed, in contact order, with their new state),
    the hits (contact, position, normal and speed, bit for bit; the point where the reference's position
    names one) and the touching contacts' count. Returns a list of mismatches."""
    bad = []
    pre, after = rec.pre_contacts, rec.after_contacts
    if not np.array_equal(after["touched"], np.asarray(got_contacts["touched"])):
        d = np.nonzero(after["touched"] != np.asarray(got_contacts["touched"]))[0]
        bad.append(f"touched at {d[:5].tolist()}")
    keys = pre["key"]
    head = np.ones(len(keys), bool)
    head[1:] = keys[1:] != keys[:-1]
    changed = np.nonzero(head & (pre["touched"] != after["touched"]))[0].astype(np.uint32)
    want = (changed << 1) | after["touched"][changed].astype(np.uint32)
    if not np.array_equal(want, np.asarray(events, np.uint32)):
        bad.append(f"events: want {len(want)}, got {len(events)}")
    w, g = rec.hits, np.asarray(hits)
    if len(w) != len(g):
        bad.append(f"hits: want {len(w)}, got {len(g)}")
    else:
        for f in ("contact", "position", "normal", "speed"):
            if _bits(w[f]).tobytes() != _bits(g[f]).tobytes():
                bad.append(f"hit {f}")
        known = w["point"] != 0xFFFFFFFF
        if not np.array_equal(w["point"][known], g["point"][known]):
            bad.append("hit point")
    if touching != rec.touching:
        bad.append(f"touching contacts: want {rec.touching}, got {touching}")
    return bad


def compare_stream(rec, contacts, ext):
    """The contact stream's result against the contacts and reuse data the reference's updateContacts
    left (src/world.cpp:880-989), field for field and bit for bit (pad bytes aside)."""
    bad = []
    want_c, want_e = rec.pre_contacts, rec.pre_ext
    if len(want_c) != len(contacts):
        return [f"contacts: want {len(want_c)}, got {len(contacts)}"]
    for f in want_c.dtype.names:
        if f in ("pad_", "points"):
            continue
        d = _rows_differ(want_c[f], contacts[f])
        if len(d):
            bad.append(f"contact {f} at {d[:5].tolist()}")
    for i in range(4):
        m = want_c["count"] > i
        for f in want_c["points"].dtype.names:
            d = _rows_differ(want_c["points"][f][:, i][m], contacts["points"][f][:, i][m])
            if len(d):
                bad.append(f"point {i} {f} at {np.nonzero(m)[0][d][:5].tolist()}")
    for f in want_e.dtype.names:
        if f == "pad_":
            continue
        w, g = want_e[f], ext[f]
        if f in ("surface_a", "surface_b"):
            for i in range(4):
                m = want_c["count"] > i
                d = _rows_differ(w[:, i][m], g[:, i][m])
                if len(d):
                    bad.append(f"{f} {i} at {np.nonzero(m)[0][d][:5].tolist()}")
            continue
        d = _rows_differ(w, g)
        if len(d):
            bad.append(f"ext {f} at {d[:5].tolist()}")
    return bad
