/*
 * twofold_solver.h - C ABI of the MI355X soft-step contact solver.
 *
 * The solver replaces the solve block of the reference engine's step (src/world.cpp:1591-1691:
 * step state, active/fast lists, SolverContext, prepare, impact marking, colouring, row packing,
 * the substep loop with warm start, biased solve, position integration and relax, restitution and
 * the final speed limit) and the island/sleep pass that follows it (src/world.cpp:1174-1247).
 * Everything on that path runs as gfx950 HIP kernels on state that stays resident in HBM; the
 * host calls in through the functions below with plain pointers and sizes.
 *
 * Conventions (the reference has no FFI; these follow its C++ conventions where it has one):
 *  - Every function returns 0 on success or a negative TFS_E_* code. tfs_last_error() gives the
 *    message of the last failure on a context. No C++ exception crosses this boundary.
 *  - A context is one world. It is not thread safe. All work is ordered on the context's stream.
 *  - The library owns device memory; the caller owns host arrays.
 *  - A non-finite or non-positive dt is ignored, as World::step ignores it (src/world.cpp:1565).
 */
#ifndef TWOFOLD_SOLVER_H
#define TWOFOLD_SOLVER_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define TFS_ABI_VERSION 3

enum {
    TFS_OK = 0,
    TFS_E_ARG = -1,      /* bad argument: null pointer, size out of range, body index out of range */
    TFS_E_DEVICE = -2,   /* a HIP call failed */
    TFS_E_STATE = -3,    /* call out of order (e.g. step before upload) */
    TFS_E_CALLBACK = -4, /* the refresh callback returned an error */
    TFS_E_KERNEL = -5,   /* a kernel reported an internal error (e.g. a bounded wait ran out) */
    TFS_E_CAPACITY = -6  /* tfs_step_noalloc: a world larger than its fixed scratch holds */
};

/* Motion types, as twofold::MotionType (include/twofold/world.h:37). */
enum { TFS_STATIC = 0, TFS_KINEMATIC = 1, TFS_DYNAMIC = 2 };

/* Lock flags, as twofold::LockFlags (include/twofold/world.h:40-50). */
enum {
    TFS_LOCK_TX = 1, TFS_LOCK_TY = 2, TFS_LOCK_TZ = 4,
    TFS_LOCK_RX = 8, TFS_LOCK_RY = 16, TFS_LOCK_RZ = 32,
    TFS_LOCK_PLANAR = 4 | 8 | 16
};

/* Colour count; contacts that get none take TFS_COLORS (src/internal.h:24, 620). */
#define TFS_COLORS 32

/*
 * A body as the solver sees it: the fields of twofold::Body (src/internal.h:257-325) that the solve
 * block, the step-state set-up and the sleep pass read or write. Quaternions are (x, y, z, w);
 * matrices are row-major. Fields marked "out" are recomputed by tfs_step at step start and need not
 * be valid on upload.
 */
typedef struct tfs_body {
    float p[3];                 /* xf.p */
    float q[4];                 /* xf.q */
    float com[3];
    float local_center[3];
    float v[3];
    float w[3];
    float force[3];
    float torque[3];
    float inv_mass;
    float inertia[9];           /* local inertia about the centre of mass */
    float lock_frame[4];
    float linear_damping;
    float angular_damping;
    float gravity_scale;
    float extent;               /* reach of the shapes from the centre of mass */
    float sleep_time;
    float kept;
    float inv_mass_world[9];    /* out */
    float inv_inertia_world[9]; /* out */
    float linear_proj[9];       /* out */
    float angular_proj[9];      /* out */
    float plane_inv_mass;       /* out */
    float plane_inv_inertia;    /* out */
    float dp[3];                /* out */
    float dq[4];                /* out */
    float q_base[4];            /* out */
    float q0[4];                /* out */
    float com0[3];              /* out */
    float expected_travel;      /* out */
    uint32_t island;            /* out: lowest body index of the island, UINT32_MAX when not awake dynamic */
    uint8_t type;               /* TFS_STATIC / TFS_KINEMATIC / TFS_DYNAMIC */
    uint8_t locks;              /* TFS_LOCK_* */
    uint8_t awake;
    uint8_t can_sleep;
    uint8_t alive;
    uint8_t pad_[3];
} tfs_body;

/* A contact point, as twofold::ContactPoint (src/internal.h:428-440) less the surface points the
 * narrowphase keeps for itself. */
typedef struct tfs_point {
    float local_a[3];
    float anchor_a[3];
    float anchor_b[3];
    float separation;
    float adjusted_separation;  /* out */
    float normal_impulse;
    float tangent_impulse[3];
    float max_normal_impulse;   /* out */
    float relative_velocity;    /* out */
} tfs_point;

/* A contact manifold, as twofold::Contact (src/internal.h:463-503) less the narrowphase's own
 * reuse data. Contacts are in key order, as the reference keeps them (src/world.cpp:985-988). */
typedef struct tfs_contact {
    uint64_t key;
    uint32_t body_a;
    uint32_t body_b;
    float normal[3];
    float tangent1[3];
    float tangent2[3];
    float friction;
    float restitution;
    float rolling_resistance;
    float rolling_radius;       /* out */
    float rolling_impulse[3];
    float plane[3];             /* out */
    int32_t count;
    uint8_t planar;             /* out */
    uint8_t laned;              /* out */
    uint8_t touched;            /* the pair was touching after the last step, as contact events report it
                                   (Contact::touched, src/internal.h:492-493); tfs_step updates it */
    uint8_t pad_;
    tfs_point points[4];
} tfs_contact;

/* The subset of twofold::WorldSettings (include/twofold/world.h:216-241) and Tolerances
 * (src/internal.h:348-351) the path reads. tfs_default_settings() fills the reference defaults. */
typedef struct tfs_settings {
    float gravity[3];
    float length_unit;
    int32_t substeps;
    float contact_hertz;
    float contact_damping_ratio;
    float max_pushout_speed;
    float max_linear_speed;
    float max_angular_speed;
    float restitution_threshold;
    float sleep_threshold;
    float time_to_sleep;
    float speculative;          /* Tolerances::speculative (4 x Tolerances::linearSlop) */
    float hit_threshold;        /* approach speed above which a touching point is a hit */
    float contact_reuse_distance;  /* a pair moved less than this relative to itself keeps its points */
    float contact_reuse_angle;     /* ... and turned less than this (radians) */
    float joint_hertz;             /* WorldSettings::jointHertz (world.h:237) */
    float joint_damping_ratio;     /* WorldSettings::jointDampingRatio (world.h:238) */
} tfs_settings;

/* What a contact keeps for its reuse in the next step (src/internal.h:463-503: localNormal,
 * baseOffset, baseTurn, reusable, part, and per point surfaceA, surfaceB, src/internal.h:428-440), one
 * per contact beside tfs_contact. tfs_stream_contacts keeps it on the device. */
typedef struct tfs_contact_ext {
    float local_normal[3];      /* the normal in body A's frame when the points were collided */
    float base_offset[3];       /* body B's centre of mass from body A's, in body A's frame, then */
    float base_turn[4];         /* body B's rotation relative to body A's, then */
    float surface_a[4][3];      /* each point on body A's surface, in body A's frame */
    float surface_b[4][3];      /* and on body B's, in body B's frame */
    uint32_t part;              /* which of the pair's manifolds (a mesh pair has several) */
    uint8_t reusable;
    uint8_t pad_[3];
} tfs_contact_ext;

/* A pair of shapes the broadphase found this step whose bodies are not both resting: the candidates of
 * updateContacts (src/world.cpp:752-869), in key order, with the contact materials the pair's shapes
 * mix to (friction, restitution, rolling resistance, src/world.cpp:951-955) and the narrowphase's
 * manifolds for it, manifolds[first .. first + count) (count 0: nothing found). */
typedef struct tfs_pair {
    uint64_t key;
    uint32_t body_a;
    uint32_t body_b;
    float friction;
    float restitution;
    float rolling_resistance;
    uint32_t first;
    uint32_t count;
    uint32_t pad_;
} tfs_pair;

/* One manifold of the narrowphase (Manifold, src/internal.h:327-346): the normal from A to B and up to
 * four points, each the midpoint of the two surfaces with their separation along the normal. */
typedef struct tfs_manifold {
    float normal[3];
    int32_t count;
    float position[4][3];
    float separation[4];
} tfs_manifold;

/* A hit: the fastest approaching point of a pair that touches in a step, when it approached faster
 * than hit_threshold x length_unit (HitEvent, include/twofold/world.h:278-288). */
typedef struct tfs_hit {
    uint32_t contact;           /* index of the contact holding the point */
    uint32_t point;             /* the point's index in it */
    float position[3];          /* the point on body A: its local_a through body A's pose */
    float normal[3];            /* from body A to body B */
    float speed;                /* approach speed, -relative_velocity */
} tfs_hit;

/*
 * The refresh callback. Before substep i > 0 of a step, when the list `again` (fast contacts at
 * i == substeps/2, else fast contacts between bodies that are not planar) is not empty, the
 * reference collides those pairs again (src/world.cpp:1657-1668, 1057-1139). The solver does the
 * solver side of that (inverse mass update, rebase of the other active contacts, reset of dp/dq,
 * then prepare, nearest-point carry and impulse carry of the new points) and asks the host for the
 * new manifolds through this callback.
 *
 * On entry, for k < n_again, again[k] is the index (as uploaded) of the k-th contact to collide
 * again, again_contacts[k] that contact as it stands before the refresh, and again_bodies[2k] and
 * again_bodies[2k + 1] the current state of its bodies a and b (all three arrays are compact, in the
 * order of `again`). For each k the callback writes fresh[k]: normal, tangent1, tangent2, count and
 * per point local_a, separation, anchor_a, anchor_b (other fields are ignored); count 0 drops the
 * points. `remaining` is the time left in the step, (substeps - i) * h. Returns 0 or an error.
 */
typedef int (*tfs_refresh_fn)(void* user, int substep, float remaining, uint32_t n_again, const uint32_t* again,
                              const tfs_contact* again_contacts, const tfs_body* again_bodies, tfs_contact* fresh);

/*
 * One convex shape per body, centred on the body's centre of mass and turned with it, for the
 * device refresh (tfs_set_body_shapes): a sphere (radius), a box (half-extents) or a static
 * half-space (the body's local z is its outward normal, its plane through the centre). The kinds
 * are the host world's numbering (twofold_amd/csrc/scene.cpp).
 */
#define TFS_SHAPE_SPHERE 0
#define TFS_SHAPE_BOX 1
#define TFS_SHAPE_HALF_SPACE 4

typedef struct tfs_shape {
    uint32_t kind;
    float half[3];
    float radius;
} tfs_shape;

/* A joint as the solve block reads and writes it: src/internal.h:508-548 (Joint), in the reference's
 * own terms (anchors, axes and reference directions in each body's frame, the rest turn, limits, motor,
 * spring, friction, break limits, the impulses carried from step to step, the rows' force and torque
 * per unit impulse, and the force and torque of the last step). body_a may be TFS_WORLD_BODY (joined to
 * the world: a static body at the origin). The device solves the point, fixed, distance and slider
 * joints and the hinge without limits; a limited hinge and the swing-twist joint read angles through
 * atan2 and asin (src/joint.cpp:212-217, 239-287), whose float bits are the host C library's, and are
 * refused (TFS_E_ARG). */
#define TFS_JOINT_POINT 0
#define TFS_JOINT_FIXED 1
#define TFS_JOINT_HINGE 2
#define TFS_JOINT_SLIDER 3
#define TFS_JOINT_DISTANCE 4
#define TFS_JOINT_SWING_TWIST 5
#define TFS_WORLD_BODY 0xffffffffu

typedef struct tfs_joint {
    uint32_t body_a;
    uint32_t body_b;
    uint8_t type;               /* TFS_JOINT_* */
    uint8_t alive;
    uint8_t broken;
    uint8_t limited;
    float local_a[3];
    float local_b[3];
    float axis_a[3];
    float axis_b[3];
    float ref_a[3];
    float ref_b[3];
    float rest_turn[4];
    float length;
    float lower, upper;
    float motor_speed, max_motor;
    float hertz, damping_ratio;
    float max_friction;
    float frame_a[4], frame_b[4];
    float swing_y, swing_z;
    float break_force, break_torque;
    float linear_impulse[3];    /* in/out: carried from step to step */
    float angular_impulse[3];
    float axial_impulse, motor_impulse, lower_impulse, upper_impulse, friction_impulse, swing_impulse;
    float friction_turn[3];
    float row_force[2][3];      /* out */
    float row_torque[2][3];
    float force[3];             /* out: over the last step it was active */
    float torque[3];
} tfs_joint;

typedef struct tfs_ctx tfs_ctx;

/* Version of this ABI (TFS_ABI_VERSION) and struct sizes, for bindings to check their layouts. */
int tfs_abi_version(void);
size_t tfs_sizeof_body(void);
size_t tfs_sizeof_contact(void);
size_t tfs_sizeof_settings(void);
size_t tfs_sizeof_joint(void);

/* The reference's default settings (include/twofold/world.h:216-241; src/internal.h:348-351). */
void tfs_default_settings(tfs_settings* out);

/* Creates a context on HIP device `device` with room for the given counts (they can grow later by
 * uploading more, up to 268,435,455 contacts: TFS_E_ARG beyond). Replaces the construction of
 * World::Impl's solver state (src/world.cpp:145-165). */
int tfs_create(int device, uint32_t max_bodies, uint32_t max_contacts, tfs_ctx** out);
void tfs_destroy(tfs_ctx* ctx);
const char* tfs_last_error(const tfs_ctx* ctx);

int tfs_set_settings(tfs_ctx* ctx, const tfs_settings* settings);

/* Host arrays in, resident device state out. Bodies are indexed as the reference's bodies vector;
 * contacts in key order as the reference's contacts vector after its narrowphase
 * (src/world.cpp:752-993). A contact's body indices must be < n_bodies. */
int tfs_upload_bodies(tfs_ctx* ctx, const tfs_body* bodies, uint32_t n);
int tfs_upload_contacts(tfs_ctx* ctx, const tfs_contact* contacts, uint32_t n);

/*
 * The contact stream: the solver half of updateContacts (src/world.cpp:880-989) on the device, so that
 * the contacts stay resident from step to step. Before a step, over the resident bodies and the
 * resident contacts of the last step, for each pair (key order): when the pair had contacts and every
 * one still fits (stillFits, src/world.cpp:1002-1008: the bodies moved and turned relative to each
 * other less than the reuse distance and angle since the points were collided), the old contacts are
 * kept with their points measured again (measureAgain, 1011-1025); otherwise each manifold becomes a
 * contact (its points on both surfaces, anchors lifted apart when far, 927-943), remembers the pair's
 * pose (remember, 995-1000) and takes the impulses of the old contact with the closest normal, point by
 * point (closestNormal, carryImpulses, 543-549, 102-120), and the touched flag. The pair's materials go
 * to all its contacts. The contacts of resting pairs (no body dynamic and awake, or a moving kinematic
 * one) stay as they were (972-984), and all are merged in key order; they replace the resident
 * contacts. The caller collides every candidate pair (a kept pair's manifold is not read) and wakes
 * bodies as updateContacts does (957-958) before tfs_step. Contacts and their reuse data then never
 * cross PCIe. *n_contacts (may be NULL) receives the number of resident contacts. TFS_E_ARG for a pair
 * out of key order, a body out of range or a manifold range outside `manifolds`.
 */
int tfs_stream_contacts(tfs_ctx* ctx, const tfs_pair* pairs, uint32_t n_pairs, const tfs_manifold* manifolds,
                        uint32_t n_manifolds, uint32_t* n_contacts);
/* The reuse data of the resident contacts (tfs_contact_ext), one per contact in the same order. An upload
 * sets it for the contacts last uploaded (tfs_upload_contacts leaves every contact not reusable). */
int tfs_upload_contact_ext(tfs_ctx* ctx, const tfs_contact_ext* ext, uint32_t n);
int tfs_download_contact_ext(tfs_ctx* ctx, tfs_contact_ext* out, uint32_t n);

/*
 * Batched worlds (no counterpart in the reference, which steps one World at a time): declares that
 * the resident bodies form n_worlds independent worlds, world w owning bodies
 * [body_start[w], body_start[w + 1]) (body_start has n_worlds + 1 entries, from 0 to n_bodies).
 * Contacts are then uploaded world after world, each world's in its own key order, with the global
 * body indices of the concatenation; a contact may not join two worlds. tfs_step solves every world
 * as it would solve that world alone, to the bit: worlds share no body, so colours, row order,
 * islands and the serial rest order stay per world, and a mid-step refresh touches only the worlds
 * that have contacts to collide again (the refresh callback still sees global indices). Island ids
 * come back local to their world, as that world alone would have them. Call after
 * tfs_upload_bodies, which resets to one world; it drops the resident contacts, so upload them after.
 */
int tfs_set_worlds(tfs_ctx* ctx, uint32_t n_worlds, const uint32_t* body_start);

/*
 * One step of the path: src/world.cpp:1591-1691 followed by updateSleep (src/world.cpp:1174-1247)
 * over the resident bodies and contacts. `refresh` may be NULL when no refresh is needed (then a
 * step that would need one fails with TFS_E_CALLBACK before changing any state).
 */
int tfs_step(tfs_ctx* ctx, float dt, tfs_refresh_fn refresh, void* user);

/*
 * The sleep pass apart from the solve block. tfs_step runs updateSleep (src/world.cpp:1174-1247) right
 * after the solve block by default. The reference runs it later in World::step, after the continuous
 * collision pass, the sensors and the kinematic targets (src/world.cpp:1696-1725), which can move and
 * slow bodies; a binding that keeps that order turns the step's sleep pass off (on = 0) and calls
 * tfs_sleep where the reference calls updateSleep (1724), after uploading the bodies those passes
 * changed. tfs_sleep is the same pass over the resident bodies and contacts, with the same results as
 * inside tfs_step; dt is the step's time.
 */
int tfs_set_step_sleep(tfs_ctx* ctx, int on);
int tfs_sleep(tfs_ctx* ctx, float dt);

/*
 * The device refresh (the reference's refresh collides on the host: the narrowphase half of
 * refreshContacts, src/world.cpp:1076-1122). With one shape per resident body set here, at the body's
 * origin, every refresh of tfs_step collides its pairs on the device with the reference's own
 * narrowphase for spheres, boxes and half-spaces, restated operation for operation
 * (twofold_amd/csrc/collide3d.h: src/collide3d.cpp collide, collideHalfSpace, collideSpheres,
 * collideSpherePoly, the two-box collideConvex with its GJK gap test, reduceManifold), so the manifolds
 * are the reference's bit for bit; the refresh callback is then not called and may be NULL. This takes
 * the host round trip and the transfer of the pairs out of each refresh. Box half-extents and sphere
 * radii are the sizes the reference keeps after building the shape (buildShape, src/shape.cpp:149).
 * n = 0 turns it off. TFS_E_ARG for an unknown kind or a size that is not finite and positive;
 * TFS_E_STATE from tfs_step when n is not the number of resident bodies; TFS_E_KERNEL from tfs_step
 * when a pair is one the reference collides in 2D (every dynamic body of the pair locked to one plane,
 * usePlanar, src/world.cpp:737-750).
 */
int tfs_set_body_shapes(tfs_ctx* ctx, uint32_t n, const tfs_shape* shapes);

/* The joints (src/internal.h:508-548), in the reference's order, and the world body's velocities
 * (world_motion: v then w, NULL for zero): the reference keeps one static world body whose velocities
 * the joint pushes touch (src/joint.cpp:140-150). Each tfs_step then does the joints' part of the solve
 * block (src/world.cpp:1623-1632, 1647, 1670-1678, 1685-1691): the active joints (alive, not broken, a
 * dynamic awake end) scaled to the kept speed, warm started and solved before the contacts in every
 * pass, broken past their limits, settled, and the bodies' inverse masses made again; the sleep pass
 * joins their bodies' islands (1210-1223). The joint and world-body arrays stay resident (upload once,
 * download when needed). TFS_E_ARG for a body index out of range, a limited hinge or a swing-twist
 * joint; tfs_step fails with TFS_E_STATE while joints are resident beside several worlds
 * (tfs_set_worlds: each world has its own world body) or name bodies no longer resident. n = 0 removes
 * them. Replaces World::Impl::joints and worldBody for the solve block. */
int tfs_upload_joints(tfs_ctx* ctx, const tfs_joint* joints, uint32_t n, const float* world_motion);
int tfs_download_joints(tfs_ctx* ctx, tfs_joint* out, uint32_t n, float* world_motion);

/*
 * One step that allocates nothing. No counterpart in the reference: it is for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc). It is tfs_step (upload, step, download) over a world that the caller keeps in its own
 * device memory: `bodies` (n_bodies) and `contacts` (n_contacts, key order) are device pointers, read
 * at the call and overwritten with the step's results; `shapes` (device pointer, one per body, or NULL)
 * turns on the device refresh as tfs_set_body_shapes does, and without it a step that needs a refresh
 * fails with TFS_E_CALLBACK. Work is ordered on `stream` (a hipStream_t, NULL for the device's default
 * stream) and the call returns once the results are in the caller's arrays.
 *
 * Every scratch buffer the step uses (copies of the world, rows, lists, colouring and island scratch,
 * counters, the grid barrier) is carved from one fixed array of 28 MiB that the library holds on each
 * device it runs on until the process ends; a world of more than TFS_NOALLOC_MAX_BODIES bodies or
 * TFS_NOALLOC_MAX_CONTACTS contacts is refused with TFS_E_CAPACITY. The first call on a device sets
 * that device's array up (devices 0-15). All calls on one device share its array and its state, and
 * each call leaves its contents behind: calls that overlap on one device (from two threads, or on two
 * streams) race. Results are the same bits as tfs_step's on the same input. Errors: TFS_E_ARG for a
 * null pointer, substeps outside 1..64, or a contact or shape that names a body out of range or is
 * malformed (checked on the device before any step kernel runs); no message is kept.
 */
#define TFS_NOALLOC_MAX_BODIES 2048
#define TFS_NOALLOC_MAX_CONTACTS 8192
int tfs_step_noalloc(int device, tfs_body* bodies, uint32_t n_bodies, tfs_contact* contacts, uint32_t n_contacts,
                     const tfs_shape* shapes, const tfs_settings* settings, float dt, void* stream);

/* What a host's broadphase and narrowphase read of a body between steps, compact (72 bytes instead of
 * the 420 of tfs_body): for the contact stream's loop, whose contacts never leave the device. */
typedef struct tfs_pose {
    float p[3];
    float q[4];
    float com[3];
    float v[3];
    float w[3];
    uint8_t type;
    uint8_t awake;
    uint8_t pad_[6];
} tfs_pose;
int tfs_download_poses(tfs_ctx* ctx, tfs_pose* out, uint32_t n);

int tfs_download_bodies(tfs_ctx* ctx, tfs_body* out, uint32_t n);
int tfs_download_contacts(tfs_ctx* ctx, tfs_contact* out, uint32_t n);

/*
 * The integer results of the last step, for parity checks against the reference:
 *   kind 0: active contacts (src/world.cpp:1608-1621)            kind 1: fast contacts
 *   kind 2: fast contacts without a planar body                 kind 3: impact contacts (1029-1049)
 *   kind 4: active contacts in row order, colour by colour, then the rest (lanes.cpp:376-405)
 *   kind 5: impact contacts in row order
 *   kind 6: colour of each contact, one byte per contact index (TFS_COLORS: none; lanes.cpp:356-374)
 * Writes up to `cap` entries (bytes for kind 6, uint32 otherwise) and the full count to *count.
 */
int tfs_download_list(tfs_ctx* ctx, int kind, void* out, uint32_t cap, uint32_t* count);

/*
 * Touch events of the last step: reportTouches (src/world.cpp:1424-1468), which the reference runs
 * right after the solve block, run by tfs_step on the device at the same point (before the sleep pass)
 * over the contacts the solve left. A pair (the consecutive contacts of one key) touches when a point
 * carries a largest normal impulse above zero or lies closer than 0 (or, for a pair that touched
 * before, closer than the linear slop, speculative / 4); each contact's `touched` is updated, and a pair
 * whose state changed gives an event, events[k] = (index of its first contact << 1) | began, in
 * contact order. A pair solved this step (a body dynamic and awake) whose fastest approaching pushed
 * point approached faster than hit_threshold x length_unit gives a hit, in contact order. `touching`
 * is the number of contacts with a point overlapping or pushed (StepProfile::touchingContacts). Writes up
 * to the caps and the full counts; any pointer may be NULL.
 */
int tfs_download_touches(tfs_ctx* ctx, uint32_t* events, uint32_t events_cap, uint32_t* n_events, tfs_hit* hits,
                         uint32_t hits_cap, uint32_t* n_hits, uint32_t* touching);

/* Solve-stage timing of the last step in milliseconds, measured with HIP events on the stream,
 * as the reference's StepProfile::solver brackets it (src/world.cpp:1644-1692). */
int tfs_last_step_ms(tfs_ctx* ctx, float* solve_ms, float* sleep_ms);

/*
 * Measurement (no counterpart in the reference): with profiling on, every launch of the coloured-row
 * kernels is bracketed by HIP events on the context's stream, and per family (0 the substep start: velocities integrated and
 * the warm start of every row, coloured and rest, in one launch; 1 solve:
 * biased and relax passes including the impact repeats, 2 restitution) the library sums the launch
 * count, the event time in milliseconds and the algorithmic bytes of the rows launched (per contact
 * with n points: warm start 224 + 244n, solve 224 + 296n, restitution 224 + 96n). Two more families
 * time each refresh's narrowphase: through the callback, with the host clock, 3 the wait for the
 * contacts collided again to reach host memory and 4 the callback itself; on the device
 * (tfs_set_body_shapes), 4 the event time of the collision kernel and 3 nothing. Family 4's "bytes"
 * sum counts the pairs collided again. Family 5 is the biased and relax passes over the contacts outside the colours in the
 * active rest list (not the impact rest list's repeats), event-timed per pass, with bytes per contact of
 * 248 + 68n (planar) or 224 + 296n (3D). Family 6 is the reverse sweep's passes (tfs_adjoint_step), event-timed per
 * pass, with 440 + 348n bytes per row (the solve's row, its record, the velocity and impulse adjoints).
 * Family 7 is the fused substeps (tfs_set_pass_shape), event-timed per launch, with the warm start's and
 * both solve passes' bytes.
 * Turning it on clears the sums; turning it off keeps them for reading.
 */
int tfs_set_profiling(tfs_ctx* ctx, int on);
int tfs_read_profile(tfs_ctx* ctx, int family, uint64_t* launches, double* ms, double* bytes);

/*
 * Pass-shape autotuning (no counterpart in the reference). The coloured passes run as one
 * grid-barrier launch each; its shape is the number of blocks (one per compute unit at most; more
 * blocks spread a colour's rows wider but make each grid barrier dearer), the widest colour that the
 * one-block tail takes at the end of a pass (0: no tail), and whether each row loads its packed
 * axes or makes them from the bodies' inverse masses. Results are the same, bit for bit, for every
 * shape. tfs_autotune(ctx, rounds) makes the next rounds * C calls of tfs_step (C candidate shapes:
 * 64, 128 and 256 blocks as far as they stay resident, tail off / 64 / 128 rows, axes loaded / made)
 * each run one candidate, in rotating order, and then keeps the shape whose steps spent the least
 * time in their solve passes (coloured and rest, event-timed; median over its rounds); a step with
 * no solve pass does not count. rounds = 0 stops tuning and keeps the current shape. The fixed-contact
 * model's steps run the current shape and do not count. A candidate's time is that of the step's substep loop
 * (refreshes, substep starts, passes, positions), event-timed. The block count sets how many rows
 * each wavefront solves per colour (a colour of R rows over B blocks of 4 wavefronts: 64 x
 * ceil(R / 256B)). tfs_pass_shape reads shape[7]: blocks, tail rows (0: off), made axes (0/1), steps
 * of tuning left, candidates (0 when not tuning), fused substep (0/1), graph (0/1). tfs_set_pass_shape sets a shape and stops tuning
 * (TFS_E_ARG for blocks outside 1..the resident limit, beyond the fused launch's when fused, or
 * tail_rows above 256).
 */
int tfs_autotune(tfs_ctx* ctx, uint32_t rounds);
/* The batched form's launch (no counterpart in the reference). A substep with no rest list can run as
 * one persistent launch (the fused substep: velocities integrated and every coloured row's warm start,
 * the biased pass, the positions, the relax pass, with a grid barrier between the phases); it is a
 * fourth part of the pass shape, which tfs_autotune also tries (each block count with loaded and made
 * axes, tail 64). tfs_set_graph(ctx, 1) captures each step's substep loop (its refreshes, substep starts,
 * passes and positions) in a HIP graph and launches it as one, when no refresh of the step needs the
 * host callback and neither profiling nor tuning is on. Results are the same bits either way. */
int tfs_set_graph(tfs_ctx* ctx, int on);
/* The passes with a grid barrier are launched as plain kernels by default: their grid is at most half
 * of what the device holds at once and a stream runs one kernel at a time, so all their blocks are
 * resident together when nothing else runs on the device. tfs_set_cooperative(ctx, 1) launches them
 * with hipLaunchCooperativeKernel instead, which also holds when other work shares the device, at
 * ~12 us per launch (no counterpart in the reference; the results are the same either way). */
int tfs_set_cooperative(tfs_ctx* ctx, int on);
int tfs_pass_shape(tfs_ctx* ctx, uint32_t shape[7]);
int tfs_set_pass_shape(tfs_ctx* ctx, uint32_t blocks, uint32_t tail_rows, int made_axes, int fused);

/* Counts of the last step: active contacts, their points, impact contacts and their points,
 * colours used, rest contacts, and the dependency levels of the active and of the impact rest list
 * (the rest is solved level after level). Constraint-iterations per step follow from these. */
int tfs_last_step_counts(tfs_ctx* ctx, uint32_t counts[8]);

/*
 * The fixed-contact model and its adjoint. No counterpart in the reference (which has no
 * differentiation); it is the solve block of src/world.cpp:1644-1681 run for several steps over
 * the resident contacts, fixed at the first step:
 *  - tfs_fixed_begin runs the step's set-up (src/world.cpp:1591-1657: step state, prepare,
 *    colouring, row order, packing) once and keeps its rows: contact frames, arms and effective
 *    masses stay those of that step. Every active contact must get a colour (TFS_E_STATE if not).
 *    Room is made for `max_steps` checkpoints.
 *  - tfs_fixed_step checkpoints the state (velocities, centres of mass, stored impulses, adjusted
 *    separations), then takes one step: per substep integrateVelocities, the warm start, the biased
 *    pass, the centres of mass and travel (com, dp += v h), the relax pass; no impact pass, no
 *    restitution. Between steps each point's adjusted separation is rebased by the travel along the
 *    normal (rebaseContact, src/solver.cpp:293-306, with rotations left out of the point positions).
 *    Bodies can be downloaded after any step.
 *  - tfs_adjoint_seed takes d loss / d com of every body after the last step (3 floats per body)
 *    and clears the other adjoints.
 *  - tfs_adjoint_step reverses the last step not yet reversed: it restores that step's checkpoint,
 *    takes the step again recording what each row saw, then runs the reverse sweep. It leaves the
 *    bodies at that step's end state.
 *  - tfs_adjoint_read gives d loss / d v and d loss / d w at the start of the earliest step reversed
 *    (3 floats per body each) and d loss / d friction per contact index (0 for contacts not in a
 *    row). Any pointer may be NULL.
 * Any upload or tfs_step ends the model (TFS_E_STATE until tfs_fixed_begin again).
 */
int tfs_fixed_begin(tfs_ctx* ctx, float dt, uint32_t max_steps);
int tfs_fixed_step(tfs_ctx* ctx);
int tfs_adjoint_seed(tfs_ctx* ctx, const float* dl_dcom);
int tfs_adjoint_step(tfs_ctx* ctx);
int tfs_adjoint_read(tfs_ctx* ctx, float* dl_dv, float* dl_dw, float* dl_dfriction);

/* No counterpart in the reference: a check of the device's one-instruction max and min, which the
 * solves use where the reference calls maxf / minf (src/internal.h:28-45). For n pairs (a[i], b[i])
 * in host memory it writes out[2i] = max, out[2i + 1] = min as computed on `device`. They must equal
 * maxf(a, b) and minf(a, b) bit for bit unless a or b is a signalling NaN. */
int tfs_check_minmax(int device, const float* a, const float* b, uint32_t n, float* out);

#ifdef __cplusplus
}
#endif

#endif
