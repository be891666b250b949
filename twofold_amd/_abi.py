"""numpy mirrors of the POD structs in include/twofold_solver.h.

Layouts are checked at load time against the library's own ``tfs_sizeof_*`` exports, so a change
to the header that is not mirrored here fails loudly instead of corrupting memory.
"""
import numpy as np

F = np.float32
U8 = np.uint8
U32 = np.uint32

BODY_DTYPE = np.dtype(
    [
        ("p", F, 3), ("q", F, 4), ("com", F, 3), ("local_center", F, 3),
        ("v", F, 3), ("w", F, 3), ("force", F, 3), ("torque", F, 3),
        ("inv_mass", F), ("inertia", F, 9), ("lock_frame", F, 4),
        ("linear_damping", F), ("angular_damping", F), ("gravity_scale", F),
        ("extent", F), ("sleep_time", F), ("kept", F),
        ("inv_mass_world", F, 9), ("inv_inertia_world", F, 9),
        ("linear_proj", F, 9), ("angular_proj", F, 9),
        ("plane_inv_mass", F), ("plane_inv_inertia", F),
        ("dp", F, 3), ("dq", F, 4), ("q_base", F, 4), ("q0", F, 4), ("com0", F, 3),
        ("expected_travel", F), ("island", U32),
        ("type", U8), ("locks", U8), ("awake", U8), ("can_sleep", U8), ("alive", U8),
        ("pad_", U8, 3),
    ],
    align=True,
)

POINT_DTYPE = np.dtype(
    [
        ("local_a", F, 3), ("anchor_a", F, 3), ("anchor_b", F, 3),
        ("separation", F), ("adjusted_separation", F), ("normal_impulse", F),
        ("tangent_impulse", F, 3), ("max_normal_impulse", F), ("relative_velocity", F),
    ],
    align=True,
)

CONTACT_DTYPE = np.dtype(
    [
        ("key", np.uint64), ("body_a", U32), ("body_b", U32),
        ("normal", F, 3), ("tangent1", F, 3), ("tangent2", F, 3),
        ("friction", F), ("restitution", F), ("rolling_resistance", F), ("rolling_radius", F),
        ("rolling_impulse", F, 3), ("plane", F, 3), ("count", np.int32),
        ("planar", U8), ("laned", U8), ("touched", U8), ("pad_", U8),
        ("points", POINT_DTYPE, 4),
    ],
    align=True,
)

# tfs_contact_ext, tfs_pair, tfs_manifold: the contact stream (tfs_stream_contacts)
EXT_DTYPE = np.dtype([("local_normal", F, 3), ("base_offset", F, 3), ("base_turn", F, 4), ("surface_a", F, (4, 3)),
                      ("surface_b", F, (4, 3)), ("part", U32), ("reusable", U8), ("pad_", U8, 3)], align=True)
PAIR_DTYPE = np.dtype([("key", np.uint64), ("body_a", U32), ("body_b", U32), ("friction", F), ("restitution", F),
                       ("rolling_resistance", F), ("first", U32), ("count", U32), ("pad_", U32)], align=True)
MANIFOLD_DTYPE = np.dtype([("normal", F, 3), ("count", np.int32), ("position", F, (4, 3)), ("separation", F, 4)],
                          align=True)

# tfs_pose: a body's pose and velocities (tfs_download_poses)
POSE_DTYPE = np.dtype([("p", F, 3), ("q", F, 4), ("com", F, 3), ("v", F, 3), ("w", F, 3), ("type", U8), ("awake", U8),
                       ("pad_", U8, 6)], align=True)

# tfs_hit: a hit event of tfs_download_touches
HIT_DTYPE = np.dtype([("contact", U32), ("point", U32), ("position", F, 3), ("normal", F, 3), ("speed", F)])

# tfs_shape: one convex shape per body for the device refresh (kinds TFS_SHAPE_*)
SHAPE_DTYPE = np.dtype([("kind", np.uint32), ("half", np.float32, 3), ("radius", np.float32)])
SHAPE_SPHERE, SHAPE_BOX, SHAPE_HALF_SPACE = 0, 1, 4

# tfs_joint: a joint as the solve block reads and writes it (src/internal.h:508-548)
JOINT_DTYPE = np.dtype(
    [
        ("body_a", U32), ("body_b", U32), ("type", U8), ("alive", U8), ("broken", U8), ("limited", U8),
        ("local_a", F, 3), ("local_b", F, 3), ("axis_a", F, 3), ("axis_b", F, 3), ("ref_a", F, 3), ("ref_b", F, 3),
        ("rest_turn", F, 4), ("length", F), ("lower", F), ("upper", F), ("motor_speed", F), ("max_motor", F),
        ("hertz", F), ("damping_ratio", F), ("max_friction", F), ("frame_a", F, 4), ("frame_b", F, 4),
        ("swing_y", F), ("swing_z", F), ("break_force", F), ("break_torque", F),
        ("linear_impulse", F, 3), ("angular_impulse", F, 3), ("axial_impulse", F), ("motor_impulse", F),
        ("lower_impulse", F), ("upper_impulse", F), ("friction_impulse", F), ("swing_impulse", F),
        ("friction_turn", F, 3), ("row_force", F, (2, 3)), ("row_torque", F, (2, 3)), ("force", F, 3),
        ("torque", F, 3),
    ],
    align=True,
)
JOINT_POINT, JOINT_FIXED, JOINT_HINGE, JOINT_SLIDER, JOINT_DISTANCE, JOINT_SWING_TWIST = range(6)
WORLD_BODY = 0xFFFFFFFF

SETTINGS_DTYPE = np.dtype(
    [
        ("gravity", F, 3), ("length_unit", F), ("substeps", np.int32),
        ("contact_hertz", F), ("contact_damping_ratio", F), ("max_pushout_speed", F),
        ("max_linear_speed", F), ("max_angular_speed", F), ("restitution_threshold", F),
        ("sleep_threshold", F), ("time_to_sleep", F), ("speculative", F), ("hit_threshold", F),
        ("contact_reuse_distance", F), ("contact_reuse_angle", F), ("joint_hertz", F), ("joint_damping_ratio", F),
    ],
    align=True,
)

STATIC, KINEMATIC, DYNAMIC = 0, 1, 2
LOCK_PLANAR = 4 | 8 | 16
COLORS = 32

# tfs_download_list kinds
LIST_ACTIVE, LIST_FAST, LIST_FAST_SOLID, LIST_IMPACT, LIST_ROWS, LIST_IMPACT_ROWS, LIST_COLORS = range(7)


def default_settings(**overrides):
    """The reference defaults (include/twofold/world.h:216-241, src/internal.h:348-351)."""
    s = np.zeros((), SETTINGS_DTYPE)
    s["gravity"] = (0.0, -9.81, 0.0)
    s["length_unit"] = 1.0
    s["substeps"] = 4
    s["contact_hertz"] = 30.0
    s["contact_damping_ratio"] = 10.0
    s["max_pushout_speed"] = 3.0
    s["max_linear_speed"] = 1000.0
    s["max_angular_speed"] = 1000.0
    s["restitution_threshold"] = 1.0
    s["sleep_threshold"] = 0.05
    s["time_to_sleep"] = 0.5
    s["speculative"] = 0.02
    s["hit_threshold"] = 1.0
    s["contact_reuse_distance"] = 0.0005
    s["contact_reuse_angle"] = 0.035
    s["joint_hertz"] = 60.0
    s["joint_damping_ratio"] = 2.0
    for k, v in overrides.items():
        s[k] = v
    return s


def check_sizes(lib, prefix="tfs"):
    """Compare the numpy layouts with the C sizes a library reports."""
    import ctypes

    for name, dt in (("body", BODY_DTYPE), ("contact", CONTACT_DTYPE), ("joint", JOINT_DTYPE)):
        if not hasattr(lib, f"{prefix}_sizeof_{name}"):
            continue
        fn = getattr(lib, f"{prefix}_sizeof_{name}")
        fn.restype = ctypes.c_size_t
        size = fn()
        if size != dt.itemsize:
            raise RuntimeError(f"{prefix}_{name} is {size} bytes in C but {dt.itemsize} in {__name__}")
