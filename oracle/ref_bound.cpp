// Test infrastructure: the reference engine with the drop-in binding compiled in. World::step runs as
// the reference wrote it, except that its solve block (src/world.cpp:1591-1691 after the step state and
// lists, from prepareContacts on) runs on the GPU through the C ABI (tfs_step, its sleep pass off) and
// its updateSleep (1724) runs on the GPU too, at the same place in the step (tfs_sleep), after the
// continuous collision pass and the sensors. The mid-step refresh collides its pairs with the
// reference's own narrowphase on the host through the refresh callback (the narrowphase half of
// refreshContacts, 1076-1122). Everything else (updateContacts, reportTouches, the continuous pass, the
// kinematic targets) is the reference's own code on the host.
//
// It is ref_capture.cpp compiled with TFB_BOUND, which hooks the solve block's free functions and the
// step's clock; this file adds what the hooks do. Built by oracle/Makefile (`make -C oracle bound`)
// into oracle/_ref/libtwofold_bound.so, linked with the HIP library; no reference source is copied.
//
// How the hooks move through a step: prepareContacts (the solve block's first call) uploads the
// bodies and contacts, runs tfs_step, writes the results back, empties the active and fast lists and
// sets the bodies aside, so that the rest of the solve block's code (member functions over the bodies
// and the lists) has nothing to do; the clock read that times the solve block (1692) puts the bodies
// back. Joints go up with the bodies, and come back into World::Impl::joints and the world body; while
// the rest of the solve block runs (breakJoints over the step's active list) the world holds stand-ins,
// and the real joints come back with the bodies, with the step's broken joints listed. The third clock
// read after it (the sleep pass's start, 1723) uploads the bodies as the
// continuous pass and the sensors left them, runs tfs_sleep, writes the results back and sets bodies,
// contacts and joints aside so that updateSleep has nothing to do; the next clock read (1725) puts them
// back.
#define TFB_BOUND 1
#include "ref_capture.cpp"

#include <cstdio>
#include <string>

namespace {

struct BoundState {
    tfs_ctx* ctx = nullptr;
    int phase = 0;  // 0 idle, 1 solve block done (bodies set aside), 2 after it, 3 sleep done (state set aside)
    int ticks = 0;
    std::string error;
    std::vector<Body> bodies;
    std::vector<Contact> contacts;
    std::vector<Joint> joints;
    std::vector<uint32_t> broken;  // the joints the step broke, in index order
};
BoundState g_b;

Vec3 v3(const float* p) { return Vec3{ p[0], p[1], p[2] }; }
Quat q4(const float* p) { return Quat{ p[0], p[1], p[2], p[3] }; }
Mat3 m9(const float* p) {
    Mat3 m;
    for (int r = 0; r < 3; ++r) { m.r[r] = v3(p + 3 * r); }
    return m;
}

// What the solve block and the sleep pass change of a body (the inverse of toBody for those fields).
void fromBody(Body& b, const tfs_body& o) {
    b.xf.p = v3(o.p);
    b.xf.q = q4(o.q);
    b.com = v3(o.com);
    b.v = v3(o.v);
    b.w = v3(o.w);
    b.sleepTime = o.sleep_time;
    b.kept = o.kept;
    b.invMassWorld = m9(o.inv_mass_world);
    b.invInertiaWorld = m9(o.inv_inertia_world);
    b.planeInvMass = o.plane_inv_mass;
    b.planeInvInertia = o.plane_inv_inertia;
    b.dp = v3(o.dp);
    b.dq = q4(o.dq);
    b.qBase = q4(o.q_base);
    b.q0 = q4(o.q0);
    b.com0 = v3(o.com0);
    b.expectedTravel = o.expected_travel;
    b.island = o.island;
    b.awake = o.awake != 0;
}

// What the solve block changes of a contact; `touched` stays the host's, for its reportTouches.
void fromContact(Contact& c, const tfs_contact& o) {
    c.normal = v3(o.normal);
    c.tangent1 = v3(o.tangent1);
    c.tangent2 = v3(o.tangent2);
    c.rollingRadius = o.rolling_radius;
    c.rollingImpulse = v3(o.rolling_impulse);
    c.plane = v3(o.plane);
    c.count = o.count;
    c.planar = o.planar != 0;
    c.laned = o.laned != 0;
    for (int i = 0; i < 4; ++i) {
        const tfs_point& q = o.points[i];
        ContactPoint& p = c.points[i];
        p.localA = v3(q.local_a);
        p.anchorA = v3(q.anchor_a);
        p.anchorB = v3(q.anchor_b);
        p.separation = q.separation;
        p.adjustedSeparation = q.adjusted_separation;
        p.normalImpulse = q.normal_impulse;
        p.tangentImpulse = v3(q.tangent_impulse);
        p.maxNormalImpulse = q.max_normal_impulse;
        p.relativeVelocity = q.relative_velocity;
    }
}

void fromExt(Contact& c, const tfs_contact_ext& o) {
    c.localNormal = v3(o.local_normal);
    c.baseOffset = v3(o.base_offset);
    c.baseTurn = q4(o.base_turn);
    for (int i = 0; i < 4; ++i) {
        c.points[i].surfaceA = v3(o.surface_a[i]);
        c.points[i].surfaceB = v3(o.surface_b[i]);
    }
    c.part = o.part;
    c.reusable = o.reusable != 0;
}

// What the solve block changes of a joint.
void fromJoint(Joint& j, const tfs_joint& o) {
    j.linearImpulse = v3(o.linear_impulse);
    j.angularImpulse = v3(o.angular_impulse);
    j.axialImpulse = o.axial_impulse;
    j.motorImpulse = o.motor_impulse;
    j.lowerImpulse = o.lower_impulse;
    j.upperImpulse = o.upper_impulse;
    j.frictionImpulse = o.friction_impulse;
    j.swingImpulse = o.swing_impulse;
    j.frictionTurn = v3(o.friction_turn);
    for (int k = 0; k < 2; ++k) {
        j.rowForce[k] = v3(o.row_force[k]);
        j.rowTorque[k] = v3(o.row_torque[k]);
    }
    j.force = v3(o.force);
    j.torque = v3(o.torque);
    j.broken = o.broken != 0;
}

tfs_settings settingsOf(const World::Impl& w) {
    tfs_settings o;
    std::memset(&o, 0, sizeof o);
    const WorldSettings& s = w.settings;
    put3(o.gravity, s.gravity);
    o.length_unit = s.lengthUnit;
    o.substeps = s.substeps;
    o.contact_hertz = s.contactHertz;
    o.contact_damping_ratio = s.contactDampingRatio;
    o.max_pushout_speed = s.maxPushoutSpeed;
    o.max_linear_speed = s.maxLinearSpeed;
    o.max_angular_speed = s.maxAngularSpeed;
    o.restitution_threshold = s.restitutionThreshold;
    o.sleep_threshold = s.sleepThreshold;
    o.time_to_sleep = s.timeToSleep;
    o.speculative = w.tol.speculative;
    o.hit_threshold = s.hitThreshold;
    o.contact_reuse_distance = s.contactReuseDistance;
    o.contact_reuse_angle = s.contactReuseAngle;
    o.joint_hertz = s.jointHertz;
    o.joint_damping_ratio = s.jointDampingRatio;
    return o;
}

bool check(int rc, const char* what) {
    if (rc == 0) { return true; }
    if (g_b.error.empty()) {
        char buf[512];
        std::snprintf(buf, sizeof buf, "%s: %d (%s)", what, rc, g_b.ctx ? tfs_last_error(g_b.ctx) : "no context");
        g_b.error = buf;
    }
    return false;
}

// The narrowphase half of refreshContacts (src/world.cpp:1076-1122) as the refresh callback: each pair
// collided again by the reference's collidePair at the bodies' current state (from the GPU), the new
// manifold with the normal closest to the contact's own, its points' local positions, separations and
// lifted anchors. The contact's feature is kept on the host (the ABI does not carry it), as the
// reference's refresh keeps it.
int refreshCallback(void*, int, float remaining, uint32_t n, const uint32_t* again, const tfs_contact* againContacts,
                    const tfs_body* againBodies, tfs_contact* fresh) {
    World::Impl& w = *g_impl;
    ++w.poseStamp;
    std::vector<Manifold> found;
    std::vector<uint8_t> taken;
    for (uint32_t k = 0; k < n; ++k) {
        const uint32_t ci = again[k];
        Contact& c = w.contacts[ci];
        fromBody(w.bodies[c.bodyA], againBodies[2 * k]);
        fromBody(w.bodies[c.bodyB], againBodies[2 * k + 1]);
        const Body& a = w.bodies[c.bodyA];
        const Body& b = w.bodies[c.bodyB];
        if (k == 0 || w.contacts[again[k - 1]].key != c.key) {
            auto run = std::lower_bound(w.boundRuns.begin(), w.boundRuns.end(), c.key,
                                        [](const World::Impl::BoundRun& r, uint64_t key) { return r.key < key; });
            TriangleBounds bounds;
            if (run != w.boundRuns.end() && run->key == c.key) {
                bounds.known = w.triangleBounds.data() + run->begin;
                bounds.knownCount = (int)(run->end - run->begin);
                bounds.moved = w.movedThisStep(a) + w.movedThisStep(b);
            }
            w.collidePair(w.shapes[c.shapeA], w.shapes[c.shapeB], remaining, found, &bounds,
                          w.turnedLittle(a, b) ? &c.feature : nullptr);
            taken.assign(found.size(), 0);
        }
        const Vec3 normal = v3(againContacts[k].normal);
        int pick = -1;
        for (size_t mi = 0; mi < found.size(); ++mi) {
            if (taken[mi]) { continue; }
            if (pick < 0 || dot(found[mi].normal, normal) > dot(found[(size_t)pick].normal, normal)) { pick = (int)mi; }
        }
        tfs_contact& f = fresh[k];
        std::memset(&f, 0, sizeof f);
        if (pick < 0) { continue; }
        taken[(size_t)pick] = 1;
        const Manifold& m = found[(size_t)pick];
        c.feature = m.feature;
        put3(f.normal, m.normal);
        const Vec3 t1 = perpendicular(m.normal);
        put3(f.tangent1, t1);
        put3(f.tangent2, cross(m.normal, t1));
        f.count = m.count;
        for (int i = 0; i < m.count; ++i) {
            tfs_point& p = f.points[i];
            put3(p.local_a, a.xf.applyInverse(m.points[i].position));
            p.separation = m.points[i].separation;
            const Vec3 up = w.lift(a, b, m.normal, p.separation);
            put3(p.anchor_a, m.points[i].position - up - a.com);
            put3(p.anchor_b, m.points[i].position + up - b.com);
        }
    }
    return 0;
}

bool upload(World::Impl& w) {
    const uint32_t nb = (uint32_t)w.bodies.size(), nc = (uint32_t)w.contacts.size();
    if (!g_b.ctx && !check(tfs_create(0, nb > 0 ? nb : 1, nc > 0 ? nc : 1, &g_b.ctx), "tfs_create")) { return false; }
    std::vector<tfs_body> bodies(nb);
    std::vector<tfs_contact> contacts(nc);
    std::vector<tfs_contact_ext> ext(nc);
    for (uint32_t i = 0; i < nb; ++i) { bodies[i] = toBody(w.bodies[i]); }
    for (uint32_t i = 0; i < nc; ++i) {
        contacts[i] = toContact(w.contacts[i]);
        ext[i] = toExt(w.contacts[i]);
    }
    const uint32_t nj = (uint32_t)w.joints.size();
    std::vector<tfs_joint> joints(nj);
    for (uint32_t i = 0; i < nj; ++i) { joints[i] = toJoint(w.joints[i]); }
    float world[6];
    put3(world, w.worldBody.v);
    put3(world + 3, w.worldBody.w);
    const tfs_settings s = settingsOf(w);
    return check(tfs_set_settings(g_b.ctx, &s), "tfs_set_settings") &&
           check(tfs_set_step_sleep(g_b.ctx, 0), "tfs_set_step_sleep") &&
           check(tfs_upload_bodies(g_b.ctx, bodies.data(), nb), "tfs_upload_bodies") &&
           check(tfs_upload_contacts(g_b.ctx, contacts.data(), nc), "tfs_upload_contacts") &&
           check(tfs_upload_contact_ext(g_b.ctx, ext.data(), nc), "tfs_upload_contact_ext") &&
           check(tfs_upload_joints(g_b.ctx, joints.data(), nj, world), "tfs_upload_joints");
}

// The joints and the world body's velocities after the solve block, and which joints it broke.
bool downloadJoints(World::Impl& w) {
    const uint32_t nj = (uint32_t)w.joints.size();
    std::vector<tfs_joint> joints(nj);
    float world[6];
    if (!check(tfs_download_joints(g_b.ctx, joints.data(), nj, world), "tfs_download_joints")) { return false; }
    g_b.broken.clear();
    for (uint32_t i = 0; i < nj; ++i) {
        if (joints[i].broken && !w.joints[i].broken) { g_b.broken.push_back(i); }
        fromJoint(w.joints[i], joints[i]);
    }
    w.worldBody.v = v3(world);
    w.worldBody.w = v3(world + 3);
    return true;
}

bool downloadBodies(World::Impl& w) {
    std::vector<tfs_body> bodies(w.bodies.size());
    if (!check(tfs_download_bodies(g_b.ctx, bodies.data(), (uint32_t)bodies.size()), "tfs_download_bodies")) { return false; }
    for (size_t i = 0; i < bodies.size(); ++i) { fromBody(w.bodies[i], bodies[i]); }
    return true;
}

void tick() {
    World::Impl* w = g_impl;
    if (!w) { return; }
    switch (g_b.phase) {
        case 1:  // the solve block's time is read (src/world.cpp:1692): its bodies and joints come back
            w->bodies.swap(g_b.bodies);
            g_b.bodies.clear();
            w->joints.swap(g_b.joints);
            g_b.joints.clear();
            w->brokenJoints = g_b.broken;
            g_b.phase = 2;
            g_b.ticks = 0;
            break;
        case 2:  // the continuous pass's start, its time, then the sleep pass's start (1719-1723)
            if (++g_b.ticks < 3) { break; }
            if (g_b.error.empty() && upload(*w) && check(tfs_sleep(g_b.ctx, g_dt), "tfs_sleep")) { downloadBodies(*w); }
            g_b.bodies.swap(w->bodies);
            g_b.contacts.swap(w->contacts);
            g_b.joints.swap(w->joints);
            g_b.phase = 3;
            break;
        case 3:  // the sleep pass's time (1725): bodies, contacts and joints come back
            w->bodies.swap(g_b.bodies);
            w->contacts.swap(g_b.contacts);
            w->joints.swap(g_b.joints);
            g_b.bodies.clear();
            g_b.contacts.clear();
            g_b.joints.clear();
            g_b.phase = 0;
            break;
        default:
            break;
    }
}

}  // namespace

namespace std {
namespace chrono {
twofold_step_clock::time_point twofold_step_clock::now() {
    tick();
    return steady_clock::now();
}
}  // namespace chrono
}  // namespace std

bool tfbBound() { return true; }

void tfbSolveBlock() {
    World::Impl& w = *g_impl;
    if (g_b.error.empty() && upload(w) &&
        check(tfs_step(g_b.ctx, g_dt, refreshCallback, nullptr), "tfs_step") && downloadBodies(w)) {
        const uint32_t nc = (uint32_t)w.contacts.size();
        std::vector<tfs_contact> contacts(nc);
        std::vector<tfs_contact_ext> ext(nc);
        if (check(tfs_download_contacts(g_b.ctx, contacts.data(), nc), "tfs_download_contacts") &&
            check(tfs_download_contact_ext(g_b.ctx, ext.data(), nc), "tfs_download_contact_ext")) {
            for (uint32_t i = 0; i < nc; ++i) {
                fromContact(w.contacts[i], contacts[i]);
                fromExt(w.contacts[i], ext[i]);
            }
        }
        downloadJoints(w);
    }
    w.activeContacts.clear();
    w.fastContacts.clear();
    w.fastSolidContacts.clear();
    g_b.bodies.swap(w.bodies);
    g_b.joints.swap(w.joints);
    w.joints.assign(g_b.joints.size(), Joint{});
    g_b.phase = 1;
}

extern "C" {

// The first error of a GPU call in a bound step (an empty string when none), and the GPU context's
// teardown (before the process exits).
const char* tfb_error(void) { return g_b.error.c_str(); }
void tfb_release(void) {
    if (g_b.ctx) { tfs_destroy(g_b.ctx); }
    g_b.ctx = nullptr;
}

}
