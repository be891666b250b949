// Test infrastructure: records the inputs and outputs of the reference engine's solve block, step by
// step, so the oracle and the HIP path can be checked against the reference itself.
//
// This file is compiled only by oracle/Makefile, from the reference sources where they lie. It
// includes the reference's src/world.cpp into this translation unit with the solver's free
// functions renamed (one macro each), so World::Impl::step calls the recorders below, which record
// and then call the reference's own functions. No reference source is copied or changed.
//
// Hooked calls in World::Impl::step (src/world.cpp): prepareContacts 1649 (solve-block input),
// colorContacts 1651, batchContacts 1656-1657 (colours, row order, impact list), warmStartBatches
// 1672 (substep counter), prepareContact 1126 (the fresh manifold of a refreshed fast contact),
// prepareLanes 1666 (end of a refresh), applyRestitution 1683 + limitSpeed 1157 (end of the solve
// block, after the final limitSpeeds at 1684), settleJoints 1687 (with active joints, the end of the
// solve block is after it and the inverse masses of 1688-1690).

#include <algorithm>
#include <cfloat>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <iterator>
#include <limits>
#include <map>
#include <memory>
#include <utility>
#include <variant>
#include <vector>

// World keeps its state behind a private pimpl; the recorder reads it directly.
#define private public
#include "twofold/world.h"
#undef private

#ifdef TFB_BOUND
// The bound build (oracle/ref_bound.cpp): world.cpp's Clock (src/world.cpp:33) becomes this clock,
// whose now() also moves the binding through the step (the end of the solve block, the sleep pass),
// and the rest of the solve block's free functions are hooked too.
namespace std {
namespace chrono {
struct twofold_step_clock {
    using duration = steady_clock::duration;
    using rep = steady_clock::rep;
    using period = steady_clock::period;
    using time_point = steady_clock::time_point;
    static constexpr bool is_steady = true;
    static time_point now();
};
}  // namespace chrono
}  // namespace std
#define steady_clock twofold_step_clock
#define warmStart tfcap_warmStart
#define solveBatches tfcap_solveBatches
#define restituteBatches tfcap_restituteBatches
#define makeContactAxes tfcap_makeContactAxes
#define scaleToKeptSpeed tfcap_scaleToKeptSpeed
#define warmStartJoints tfcap_warmStartJoints
#define solveJoints tfcap_solveJoints
#endif

#define prepareContacts tfcap_prepareContacts
#define colorContacts tfcap_colorContacts
#define batchContacts tfcap_batchContacts
#define prepareLanes tfcap_prepareLanes
#define prepareContact tfcap_prepareContact
#define warmStartBatches tfcap_warmStartBatches
#define applyRestitution tfcap_applyRestitution
#define limitSpeed tfcap_limitSpeed
#define settleJoints tfcap_settleJoints
#include "world.cpp"
#undef prepareContacts
#undef colorContacts
#undef batchContacts
#undef prepareLanes
#undef prepareContact
#undef warmStartBatches
#undef applyRestitution
#undef limitSpeed
#undef settleJoints
#ifdef TFB_BOUND
#undef steady_clock
#undef warmStart
#undef solveBatches
#undef restituteBatches
#undef makeContactAxes
#undef scaleToKeptSpeed
#undef warmStartJoints
#undef solveJoints
#endif

#include "../include/twofold_solver.h"

namespace twofold {
// The reference's solver entry points under their own names (src/internal.h:627-647); the macros
// above renamed the declarations world.cpp saw.
void prepareContacts(SolverData& s, const std::vector<uint32_t>& active);
void colorContacts(SolverData& s, const std::vector<uint32_t>& list, ContactColors& colors);
void batchContacts(SolverData& s, const std::vector<uint32_t>& list, const ContactColors& colors,
                   const SolverContext& ctx, ContactBatches& out);
void prepareLanes(SolverData& s, const SolverContext& ctx, ContactBatches& batches);
void prepareContact(const Body& a, const Body& b, Contact& c, PlanarAxes* planar);
void warmStartBatches(SolverData& s, const ContactBatches& batches);
void applyRestitution(SolverData& s, const std::vector<uint32_t>& active, const SolverContext& ctx);
void limitSpeed(Body& b, float maxSpeed, float maxSpin);
void settleJoints(SolverData& s, const std::vector<uint32_t>& active, int iterations);
#ifdef TFB_BOUND
void warmStart(SolverData& s, const std::vector<uint32_t>& active);
void solveBatches(SolverData& s, const ContactBatches& batches, const SolverContext& ctx, bool useBias);
void restituteBatches(SolverData& s, const ContactBatches& batches, const SolverContext& ctx);
void makeContactAxes(const Body& a, const Body& b, const Contact& c, PointAxes* axes);
void scaleToKeptSpeed(SolverData& s, const std::vector<uint32_t>& active);
void warmStartJoints(SolverData& s, const std::vector<uint32_t>& active);
void solveJoints(SolverData& s, const std::vector<uint32_t>& active, const SolverContext& ctx, bool useBias);
#endif
}

#ifdef TFB_BOUND
// The binding's part of the step (oracle/ref_bound.cpp): the solve block on the GPU.
bool tfbBound();
void tfbSolveBlock();
#else
bool tfbBound() { return false; }
void tfbSolveBlock() {}
#endif

using namespace twofold;

namespace {

void put3(float* d, const Vec3& v) { d[0] = v.x; d[1] = v.y; d[2] = v.z; }
void put4(float* d, const Quat& q) { d[0] = q.x; d[1] = q.y; d[2] = q.z; d[3] = q.w; }
void put9(float* d, const Mat3& m) {
    for (int r = 0; r < 3; ++r) { put3(d + 3 * r, m.r[r]); }
}

tfs_body toBody(const Body& b) {
    tfs_body o;
    std::memset(&o, 0, sizeof o);
    put3(o.p, b.xf.p);
    put4(o.q, b.xf.q);
    put3(o.com, b.com);
    put3(o.local_center, b.localCenter);
    put3(o.v, b.v);
    put3(o.w, b.w);
    put3(o.force, b.force);
    put3(o.torque, b.torque);
    o.inv_mass = b.invMass;
    put9(o.inertia, b.inertia);
    put4(o.lock_frame, b.lockFrame);
    o.linear_damping = b.linearDamping;
    o.angular_damping = b.angularDamping;
    o.gravity_scale = b.gravityScale;
    o.extent = b.extent;
    o.sleep_time = b.sleepTime;
    o.kept = b.kept;
    put9(o.inv_mass_world, b.invMassWorld);
    put9(o.inv_inertia_world, b.invInertiaWorld);
    put9(o.linear_proj, b.linearProj);
    put9(o.angular_proj, b.angularProj);
    o.plane_inv_mass = b.planeInvMass;
    o.plane_inv_inertia = b.planeInvInertia;
    put3(o.dp, b.dp);
    put4(o.dq, b.dq);
    put4(o.q_base, b.qBase);
    put4(o.q0, b.q0);
    put3(o.com0, b.com0);
    o.expected_travel = b.expectedTravel;
    o.island = b.island;
    o.type = (uint8_t)b.type;
    o.locks = b.locks;
    o.awake = b.awake ? 1 : 0;
    o.can_sleep = b.canSleep ? 1 : 0;
    o.alive = b.alive ? 1 : 0;
    return o;
}

tfs_contact toContact(const Contact& c) {
    tfs_contact o;
    std::memset(&o, 0, sizeof o);
    o.key = c.key;
    o.body_a = c.bodyA;
    o.body_b = c.bodyB;
    put3(o.normal, c.normal);
    put3(o.tangent1, c.tangent1);
    put3(o.tangent2, c.tangent2);
    o.friction = c.friction;
    o.restitution = c.restitution;
    o.rolling_resistance = c.rollingResistance;
    o.rolling_radius = c.rollingRadius;
    put3(o.rolling_impulse, c.rollingImpulse);
    put3(o.plane, c.plane);
    o.count = c.count;
    o.planar = c.planar ? 1 : 0;
    o.laned = c.laned ? 1 : 0;
    o.touched = c.touched ? 1 : 0;
    for (int i = 0; i < 4; ++i) {
        const ContactPoint& p = c.points[i];
        tfs_point& q = o.points[i];
        put3(q.local_a, p.localA);
        put3(q.anchor_a, p.anchorA);
        put3(q.anchor_b, p.anchorB);
        q.separation = p.separation;
        q.adjusted_separation = p.adjustedSeparation;
        q.normal_impulse = p.normalImpulse;
        put3(q.tangent_impulse, p.tangentImpulse);
        q.max_normal_impulse = p.maxNormalImpulse;
        q.relative_velocity = p.relativeVelocity;
    }
    return o;
}

tfs_joint toJoint(const Joint& j) {
    tfs_joint o;
    std::memset(&o, 0, sizeof o);
    o.body_a = j.bodyA;
    o.body_b = j.bodyB;
    o.type = (uint8_t)j.type;
    o.alive = j.alive ? 1 : 0;
    o.broken = j.broken ? 1 : 0;
    o.limited = j.limited ? 1 : 0;
    put3(o.local_a, j.localA);
    put3(o.local_b, j.localB);
    put3(o.axis_a, j.axisA);
    put3(o.axis_b, j.axisB);
    put3(o.ref_a, j.refA);
    put3(o.ref_b, j.refB);
    put4(o.rest_turn, j.restTurn);
    o.length = j.length;
    o.lower = j.lower;
    o.upper = j.upper;
    o.motor_speed = j.motorSpeed;
    o.max_motor = j.maxMotor;
    o.hertz = j.hertz;
    o.damping_ratio = j.dampingRatio;
    o.max_friction = j.maxFriction;
    put4(o.frame_a, j.frameA);
    put4(o.frame_b, j.frameB);
    o.swing_y = j.swingY;
    o.swing_z = j.swingZ;
    o.break_force = j.breakForce;
    o.break_torque = j.breakTorque;
    put3(o.linear_impulse, j.linearImpulse);
    put3(o.angular_impulse, j.angularImpulse);
    o.axial_impulse = j.axialImpulse;
    o.motor_impulse = j.motorImpulse;
    o.lower_impulse = j.lowerImpulse;
    o.upper_impulse = j.upperImpulse;
    o.friction_impulse = j.frictionImpulse;
    o.swing_impulse = j.swingImpulse;
    put3(o.friction_turn, j.frictionTurn);
    for (int k = 0; k < 2; ++k) {
        put3(o.row_force[k], j.rowForce[k]);
        put3(o.row_torque[k], j.rowTorque[k]);
    }
    put3(o.force, j.force);
    put3(o.torque, j.torque);
    return o;
}

tfs_contact_ext toExt(const Contact& c) {
    tfs_contact_ext o;
    std::memset(&o, 0, sizeof o);
    put3(o.local_normal, c.localNormal);
    put3(o.base_offset, c.baseOffset);
    put4(o.base_turn, c.baseTurn);
    for (int i = 0; i < 4; ++i) {
        put3(o.surface_a[i], c.points[i].surfaceA);
        put3(o.surface_b[i], c.points[i].surfaceB);
    }
    o.part = c.part;
    o.reusable = c.reusable ? 1 : 0;
    return o;
}

struct Refresh {
    int substep = 0;
    float remaining = 0.0f;
    std::vector<uint32_t> again;
    std::vector<tfs_contact> fresh;
    std::map<uint32_t, tfs_contact> seen;
};

struct Capture {
    bool solved = false;  // the step reached the solve block
    std::vector<tfs_body> preBodies, postBodies, stepBodies;
    std::vector<tfs_contact> preContacts, postContacts;
    std::vector<uint32_t> active, fast, fastSolid, impact;
    std::vector<uint32_t> byColor, rest, impactByColor, impactRest;
    std::vector<uint8_t> colors;
    std::vector<Refresh> refreshes;
    float ctx[12] = {};
    // updateContacts (src/world.cpp:752-989): the state it started from, its candidate pairs with their
    // mixed materials and the narrowphase's manifolds for each, and the reuse data it left
    std::vector<tfs_body> oldBodies;
    std::vector<tfs_contact> oldContacts;
    std::vector<tfs_contact_ext> oldExt, preExt;
    std::vector<tfs_pair> pairs;
    std::vector<tfs_manifold> manifolds;
    // the joints and the world body's velocities (v, w) at the solve block's start and end
    std::vector<tfs_joint> preJoints, postJoints;
    float preWorld[6] = {}, postWorld[6] = {};
};

struct Handle {
    World world;
    Capture cap;
    uint32_t counts[4] = {};
    explicit Handle(const WorldSettings& s) : world(s) {}
};

Capture* g_cap = nullptr;
// Point counts of the last step (active contacts, their points, impact contacts, their points),
// gathered in the batchContacts hook for the CPU timing leg.
uint32_t* g_counts = nullptr;
World::Impl* g_impl = nullptr;
const Contact* g_contactBase = nullptr;
int g_substep = 0;
int g_batchCalls = 0;
int g_lanesCalls = 0;
bool g_final = false;
size_t g_limitCount = 0, g_alive = 0;
float g_dt = 0.0f;

}

namespace twofold {

void recordJoints(std::vector<tfs_joint>& out, float* world) {
    out.clear();
    for (const Joint& j : g_impl->joints) { out.push_back(toJoint(j)); }
    put3(world, g_impl->worldBody.v);
    put3(world + 3, g_impl->worldBody.w);
}

void tfcap_prepareContacts(SolverData& s, const std::vector<uint32_t>& active) {
    if (tfbBound()) {
        tfbSolveBlock();
        return;
    }
    if (g_cap) {
        Capture& c = *g_cap;
        c.solved = true;
        c.preBodies.clear();
        for (const Body& b : s.bodies) { c.preBodies.push_back(toBody(b)); }
        c.preContacts.clear();
        for (const Contact& x : s.contacts) { c.preContacts.push_back(toContact(x)); }
        c.active = active;
        recordJoints(c.preJoints, c.preWorld);
        c.fast = g_impl->fastContacts;
        c.fastSolid = g_impl->fastSolidContacts;
        c.preExt.clear();
        for (const Contact& x : s.contacts) { c.preExt.push_back(toExt(x)); }
        // the candidates of this step's updateContacts, each collided again now: collidePair depends on the
        // bodies' poses and velocities only, which nothing has changed since
        World::Impl& w = *g_impl;
        c.pairs.clear();
        c.manifolds.clear();
        std::vector<Manifold> found;
        for (uint64_t key : w.candidatePairs) {
            const uint32_t ia = (uint32_t)(key >> 32), ib = (uint32_t)(key & 0xffffffffu);
            const Shape& sa = w.shapes[ia];
            const Shape& sb = w.shapes[ib];
            const Body& ba = w.bodies[sa.body];
            const Body& bb = w.bodies[sb.body];
            tfs_pair p;
            std::memset(&p, 0, sizeof p);
            p.key = key;
            p.body_a = sa.body;
            p.body_b = sb.body;
            p.friction = w.frictionMix({ sa.friction, ba.type }, { sb.friction, bb.type });
            p.restitution = w.restitutionMix({ sa.restitution, ba.type }, { sb.restitution, bb.type });
            p.rolling_resistance = maxf(sa.rollingResistance, sb.rollingResistance);
            w.collidePair(sa, sb, g_dt, found);
            p.first = (uint32_t)c.manifolds.size();
            p.count = (uint32_t)found.size();
            for (const Manifold& m : found) {
                tfs_manifold x;
                std::memset(&x, 0, sizeof x);
                put3(x.normal, m.normal);
                x.count = m.count;
                for (int i = 0; i < m.count; ++i) {
                    put3(x.position[i], m.points[i].position);
                    x.separation[i] = m.points[i].separation;
                }
                c.manifolds.push_back(x);
            }
            c.pairs.push_back(p);
        }
        g_contactBase = s.contacts.data();
        g_alive = 0;
        for (const Body& b : s.bodies) { g_alive += b.alive ? 1 : 0; }
    }
    prepareContacts(s, active);
}

void tfcap_colorContacts(SolverData& s, const std::vector<uint32_t>& list, ContactColors& colors) {
    if (tfbBound()) { return; }
    colorContacts(s, list, colors);
    if (g_cap) { g_cap->colors = colors.of; }
}

void tfcap_batchContacts(SolverData& s, const std::vector<uint32_t>& list, const ContactColors& colors,
                         const SolverContext& ctx, ContactBatches& out) {
    if (tfbBound()) { return; }
    batchContacts(s, list, colors, ctx, out);
    if (g_counts) {
        uint32_t points = 0;
        for (uint32_t ci : list) { points += (uint32_t)s.contacts[ci].count; }
        g_counts[g_batchCalls == 0 ? 0 : 2] = (uint32_t)list.size();
        g_counts[g_batchCalls == 0 ? 1 : 3] = points;
    }
    if (!g_cap) {
        ++g_batchCalls;
        return;
    }
    Capture& c = *g_cap;
    if (g_batchCalls == 0) {
        c.byColor = out.byColor;
        c.rest = out.rest;
        const float v[12] = { ctx.h, ctx.invH, ctx.contactSoft.biasRate, ctx.contactSoft.massScale,
                              ctx.contactSoft.impulseScale, ctx.staticSoft.biasRate, ctx.staticSoft.massScale,
                              ctx.staticSoft.impulseScale, ctx.maxPushout, ctx.restitutionThreshold, 0.0f, 0.0f };
        std::memcpy(c.ctx, v, sizeof v);
    } else {
        c.impact = list;
        c.impactByColor = out.byColor;
        c.impactRest = out.rest;
    }
    ++g_batchCalls;
}

void tfcap_warmStartBatches(SolverData& s, const ContactBatches& batches) {
    if (tfbBound()) { return; }
    warmStartBatches(s, batches);
    ++g_substep;
}

void tfcap_prepareContact(const Body& a, const Body& b, Contact& c, PlanarAxes* planar) {
    if (g_cap) {
        Capture& cap = *g_cap;
        if (cap.refreshes.empty() || cap.refreshes.back().substep != g_substep) {
            Refresh r;
            r.substep = g_substep;
            cap.refreshes.push_back(r);
        }
        const uint32_t ci = (uint32_t)(&c - g_contactBase);
        cap.refreshes.back().seen[ci] = toContact(c);
    }
    prepareContact(a, b, c, planar);
}

void tfcap_prepareLanes(SolverData& s, const SolverContext& ctx, ContactBatches& batches) {
    if (tfbBound()) { return; }
    prepareLanes(s, ctx, batches);
    if (!g_cap) { return; }
    if (g_lanesCalls % 2 == 0) {
        Capture& cap = *g_cap;
        if (cap.refreshes.empty() || cap.refreshes.back().substep != g_substep) {
            Refresh r;
            r.substep = g_substep;
            cap.refreshes.push_back(r);
        }
        Refresh& r = cap.refreshes.back();
        const int n = std::max(1, g_impl->settings.substeps);
        r.again = g_substep == n / 2 ? g_impl->fastContacts : g_impl->fastSolidContacts;
        r.remaining = (float)(n - g_substep) * ctx.h;
        for (uint32_t ci : r.again) {
            auto it = r.seen.find(ci);
            if (it != r.seen.end()) {
                r.fresh.push_back(it->second);
            } else {
                // No manifold was found: the contact keeps its fields and loses its points.
                tfs_contact x = toContact(s.contacts[ci]);
                x.count = 0;
                r.fresh.push_back(x);
            }
        }
        r.seen.clear();
    }
    ++g_lanesCalls;
}

void tfcap_applyRestitution(SolverData& s, const std::vector<uint32_t>& active, const SolverContext& ctx) {
    if (tfbBound()) { return; }
    applyRestitution(s, active, ctx);
    g_final = true;
    g_limitCount = 0;
}

void tfcap_limitSpeed(Body& b, float maxSpeed, float maxSpin) {
    limitSpeed(b, maxSpeed, maxSpin);
    if (!g_final) { return; }
    if (++g_limitCount < g_alive) { return; }
    g_final = false;
    if (!g_cap) { return; }
    Capture& c = *g_cap;
    c.postBodies.clear();
    for (const Body& x : g_impl->bodies) { c.postBodies.push_back(toBody(x)); }
    c.postContacts.clear();
    for (const Contact& x : g_impl->contacts) { c.postContacts.push_back(toContact(x)); }
    recordJoints(c.postJoints, c.postWorld);
}

// With active joints the solve block ends after settleJoints and the inverse masses made again
// (src/world.cpp:1687-1690); they are made here too (the same values again) to record that end.
void tfcap_settleJoints(SolverData& s, const std::vector<uint32_t>& active, int iterations) {
    if (tfbBound()) { return; }
    settleJoints(s, active, iterations);
    if (!g_cap || tfbBound()) { return; }
    Capture& c = *g_cap;
    c.postBodies.clear();
    for (Body& b : g_impl->bodies) {
        if (b.alive && b.type == MotionType::Dynamic && b.awake) { updateInverseMass(b); }
        c.postBodies.push_back(toBody(b));
    }
    recordJoints(c.postJoints, c.postWorld);
}

#ifdef TFB_BOUND
// The rest of the solve block in the bound build: done on the GPU by tfbSolveBlock.
void tfcap_warmStart(SolverData& s, const std::vector<uint32_t>& active) {
    if (!tfbBound()) { warmStart(s, active); }
}
void tfcap_solveBatches(SolverData& s, const ContactBatches& batches, const SolverContext& ctx, bool useBias) {
    if (!tfbBound()) { solveBatches(s, batches, ctx, useBias); }
}
void tfcap_restituteBatches(SolverData& s, const ContactBatches& batches, const SolverContext& ctx) {
    if (!tfbBound()) { restituteBatches(s, batches, ctx); }
}
void tfcap_makeContactAxes(const Body& a, const Body& b, const Contact& c, PointAxes* axes) {
    if (!tfbBound()) { makeContactAxes(a, b, c, axes); }
}
void tfcap_scaleToKeptSpeed(SolverData& s, const std::vector<uint32_t>& active) {
    if (!tfbBound()) { scaleToKeptSpeed(s, active); }
}
void tfcap_warmStartJoints(SolverData& s, const std::vector<uint32_t>& active) {
    if (!tfbBound()) { warmStartJoints(s, active); }
}
void tfcap_solveJoints(SolverData& s, const std::vector<uint32_t>& active, const SolverContext& ctx, bool useBias) {
    if (!tfbBound()) { solveJoints(s, active, ctx, useBias); }
}
#endif

}

namespace {

Quat quat4(const float* q) { return q ? Quat{ q[0], q[1], q[2], q[3] } : Quat{}; }
Vec3 vec3(const float* v) { return v ? Vec3{ v[0], v[1], v[2] } : Vec3{}; }

template <class T>
uint32_t copyOut(const std::vector<T>& v, T* out, uint32_t cap) {
    if (out) { std::memcpy(out, v.data(), sizeof(T) * std::min<size_t>(cap, v.size())); }
    return (uint32_t)v.size();
}

}

extern "C" {

// Creates a world with the reference defaults overridden by `s` (NULL keeps them).
void* tfr_create(const tfs_settings* s) {
    WorldSettings ws;
    if (s) {
        ws.gravity = vec3(s->gravity);
        ws.lengthUnit = s->length_unit;
        ws.substeps = s->substeps;
        ws.contactHertz = s->contact_hertz;
        ws.contactDampingRatio = s->contact_damping_ratio;
        ws.maxPushoutSpeed = s->max_pushout_speed;
        ws.maxLinearSpeed = s->max_linear_speed;
        ws.maxAngularSpeed = s->max_angular_speed;
        ws.restitutionThreshold = s->restitution_threshold;
        ws.sleepThreshold = s->sleep_threshold;
        ws.timeToSleep = s->time_to_sleep;
        ws.hitThreshold = s->hit_threshold;
        ws.contactReuseDistance = s->contact_reuse_distance;
        ws.contactReuseAngle = s->contact_reuse_angle;
        ws.jointHertz = s->joint_hertz;
        ws.jointDampingRatio = s->joint_damping_ratio;
    }
    return new Handle(ws);
}

void tfr_destroy(void* h) { delete (Handle*)h; }

// A default WorldSettings and the Tolerances a World derives from it, as tfs_settings.
void tfr_default_settings(tfs_settings* o) {
    std::memset(o, 0, sizeof *o);
    const World w{ WorldSettings{} };
    const WorldSettings& s = w.settings();
    put3(o->gravity, s.gravity);
    o->length_unit = s.lengthUnit;
    o->substeps = s.substeps;
    o->contact_hertz = s.contactHertz;
    o->contact_damping_ratio = s.contactDampingRatio;
    o->max_pushout_speed = s.maxPushoutSpeed;
    o->max_linear_speed = s.maxLinearSpeed;
    o->max_angular_speed = s.maxAngularSpeed;
    o->restitution_threshold = s.restitutionThreshold;
    o->sleep_threshold = s.sleepThreshold;
    o->time_to_sleep = s.timeToSleep;
    o->speculative = w.impl_->tol.speculative;
    o->hit_threshold = s.hitThreshold;
    o->contact_reuse_distance = s.contactReuseDistance;
    o->contact_reuse_angle = s.contactReuseAngle;
    o->joint_hertz = s.jointHertz;
    o->joint_damping_ratio = s.jointDampingRatio;
}

// Makes `h` a copy of `from` (World's copy is a full snapshot, src/world.cpp:1732-1737).
void tfr_copy(void* h, const void* from) { ((Handle*)h)->world = ((const Handle*)from)->world; }

void tfr_set_continuous(void* h, int on) {
    WorldSettings s = ((Handle*)h)->world.settings();
    s.continuous = on != 0;
    ((Handle*)h)->world.setSettings(s);
}

// type: 0 static, 1 kinematic, 2 dynamic. Returns the body index.
int tfr_add_body(void* h, int type, const float* pos, const float* rot, const float* v, const float* w,
                 int locks, const float* lockFrame, int awake, float linearDamping, float angularDamping,
                 float gravityScale) {
    BodyDesc d;
    d.type = (MotionType)type;
    d.position = vec3(pos);
    d.rotation = quat4(rot);
    d.linearVelocity = vec3(v);
    d.angularVelocity = vec3(w);
    d.locks = (uint8_t)locks;
    d.lockFrame = quat4(lockFrame);
    d.awake = awake != 0;
    d.linearDamping = linearDamping;
    d.angularDamping = angularDamping;
    d.gravityScale = gravityScale;
    return (int)((Handle*)h)->world.createBody(d).index;
}

// kind: 0 sphere {r}, 1 box {hx, hy, hz}, 2 cylinder {r, half height}, 3 prism {zMin, zMax, x0, y0, ...},
// 4 half-space {}. Returns the shape index or -1.
int tfr_add_shape(void* h, int body, int kind, const float* dims, int ndims, const float* pos, const float* rot,
                  float density, float friction, float restitution, float rolling) {
    World& wd = ((Handle*)h)->world;
    ShapeDesc sd;
    switch (kind) {
        case 0: sd.geometry = Sphere{ dims[0] }; break;
        case 1: sd.geometry = Box{ { dims[0], dims[1], dims[2] } }; break;
        case 2: sd.geometry = Cylinder{ dims[0], dims[1] }; break;
        case 3: {
            Prism p;
            p.zMin = dims[0];
            p.zMax = dims[1];
            for (int i = 2; i + 1 < ndims; i += 2) { p.polygon.push_back({ dims[i], dims[i + 1] }); }
            sd.geometry = p;
            break;
        }
        case 4: sd.geometry = HalfSpace{}; break;
        default: return -1;
    }
    sd.position = vec3(pos);
    sd.rotation = quat4(rot);
    sd.density = density;
    sd.friction = friction;
    sd.restitution = restitution;
    sd.rollingResistance = rolling;
    if (body < 0 || (size_t)body >= wd.impl_->bodies.size()) { return -1; }
    const BodyId id{ (uint32_t)body, wd.impl_->bodies[(size_t)body].generation };
    return (int)wd.addShape(id, sd).index;
}

// Adds a force and a torque to a body for the next step through the public API (World::addForce,
// World::addTorque, src/world.cpp:2064-2088), without waking it: a sleeping body ignores them.
void tfr_add_force(void* h, int body, const float* force, const float* torque) {
    World& wd = ((Handle*)h)->world;
    if (body < 0 || (size_t)body >= wd.impl_->bodies.size()) { return; }
    const BodyId id{ (uint32_t)body, wd.impl_->bodies[(size_t)body].generation };
    wd.addForce(id, vec3(force), false);
    wd.addTorque(id, vec3(torque), false);
}

// Steps the world, recording the solve block when `record` is set. Returns 1 when the step reached
// the solve block, 0 otherwise.
int tfr_step(void* h, float dt, int record) {
    Handle& hd = *(Handle*)h;
    hd.cap = Capture{};
    g_cap = record == 1 ? &hd.cap : nullptr;
    g_dt = dt;
    if (g_cap) {
        for (const Body& b : hd.world.impl_->bodies) { g_cap->oldBodies.push_back(toBody(b)); }
        for (const Contact& x : hd.world.impl_->contacts) {
            g_cap->oldContacts.push_back(toContact(x));
            g_cap->oldExt.push_back(toExt(x));
        }
    }
    std::memset(hd.counts, 0, sizeof hd.counts);
    g_counts = hd.counts;
    g_impl = hd.world.impl_.get();
    g_substep = 0;
    g_batchCalls = 0;
    g_lanesCalls = 0;
    g_final = false;
    hd.world.step(dt);
    if (record == 1) {
        for (const Body& b : g_impl->bodies) { hd.cap.stepBodies.push_back(toBody(b)); }
    }
    g_cap = nullptr;
    g_counts = nullptr;
    g_impl = nullptr;
    return hd.cap.solved ? 1 : 0;
}

// Active contacts, their points, impact contacts, their points of the last step.
void tfr_counts(void* h, uint32_t* out) { std::memcpy(out, ((Handle*)h)->counts, sizeof ((Handle*)h)->counts); }

// Moves every body that is not static to the pose and velocities of `b` (indexed as the world's
// bodies), through the public API; bodies asleep in `b` are put to sleep.
void tfr_set_bodies(void* h, const tfs_body* b, uint32_t n) {
    World& w = ((Handle*)h)->world;
    for (uint32_t i = 0; i < n && i < w.impl_->bodies.size(); ++i) {
        const Body& x = w.impl_->bodies[i];
        if (!x.alive || x.type == MotionType::Static) { continue; }
        const BodyId id{ i, x.generation };
        w.setTransform(id, vec3(b[i].p), quat4(b[i].q));
        w.setLinearVelocity(id, vec3(b[i].v));
        w.setAngularVelocity(id, vec3(b[i].w));
        if (!b[i].awake) { w.sleep(id); }
    }
}

// StepProfile of the last step: total, broadphase, narrowphase, solver, continuous, sleep (ms), and
// touching contacts.
void tfr_profile(void* h, float* out) {
    const StepProfile& p = ((Handle*)h)->world.profile();
    out[0] = p.total;
    out[1] = p.broadphase;
    out[2] = p.narrowphase;
    out[3] = p.solver;
    out[4] = p.continuous;
    out[5] = p.sleep;
    out[6] = (float)p.touchingContacts;
}

// The hit events of the last step (reportTouches, src/world.cpp:1424-1468) as tfs_hit: the contact is
// the first of the pair's key (the scenes here have one contact per pair), the point is the one whose
// position the event holds (UINT32_MAX if none matches).
uint32_t tfr_hits(void* h, tfs_hit* out, uint32_t cap) {
    World::Impl& w = *((Handle*)h)->world.impl_;
    std::vector<tfs_hit> v;
    for (const HitEvent& e : w.hitEvents) {
        tfs_hit x;
        std::memset(&x, 0, sizeof x);
        const uint64_t key = ((uint64_t)e.shapeA.index << 32) | e.shapeB.index;
        auto it = std::lower_bound(w.contacts.begin(), w.contacts.end(), key,
                                   [](const Contact& c, uint64_t k) { return c.key < k; });
        x.contact = it != w.contacts.end() && it->key == key ? (uint32_t)(it - w.contacts.begin()) : UINT32_MAX;
        x.point = UINT32_MAX;
        if (x.contact != UINT32_MAX) {
            const Body& a = w.bodies[it->bodyA];
            for (int i = 0; i < it->count; ++i) {
                const Vec3 p = a.xf.apply(it->points[i].localA);
                if (p.x == e.point.x && p.y == e.point.y && p.z == e.point.z) {
                    x.point = (uint32_t)i;
                    break;
                }
            }
        }
        put3(x.position, e.point);
        put3(x.normal, e.normal);
        x.speed = e.speed;
        v.push_back(x);
    }
    return copyOut(v, out, cap);
}

uint32_t tfr_num_bodies(void* h) { return (uint32_t)((Handle*)h)->world.impl_->bodies.size(); }
uint32_t tfr_num_contacts(void* h) { return (uint32_t)((Handle*)h)->world.impl_->contacts.size(); }

// Current state, outside any step.
uint32_t tfr_bodies(void* h, tfs_body* out, uint32_t cap) {
    std::vector<tfs_body> v;
    for (const Body& b : ((Handle*)h)->world.impl_->bodies) { v.push_back(toBody(b)); }
    return copyOut(v, out, cap);
}
uint32_t tfr_contacts(void* h, tfs_contact* out, uint32_t cap) {
    std::vector<tfs_contact> v;
    for (const Contact& c : ((Handle*)h)->world.impl_->contacts) { v.push_back(toContact(c)); }
    return copyOut(v, out, cap);
}

// Recorded bodies: which 0 = at the start of the solve block, 1 = after its final limitSpeeds,
// 2 = after the whole step (sleep included).
uint32_t tfr_cap_bodies(void* h, int which, tfs_body* out, uint32_t cap) {
    const Capture& c = ((Handle*)h)->cap;
    return copyOut(which == 0 ? c.preBodies : which == 1 ? c.postBodies : c.stepBodies, out, cap);
}

// Recorded contacts: which 0 = at the start of the solve block, 1 = after it, 2 = before the step.
uint32_t tfr_cap_contacts(void* h, int which, tfs_contact* out, uint32_t cap) {
    const Capture& c = ((Handle*)h)->cap;
    return copyOut(which == 0 ? c.preContacts : which == 1 ? c.postContacts : c.oldContacts, out, cap);
}

// Recorded reuse data: which 0 = at the start of the solve block, 2 = before the step.
uint32_t tfr_cap_ext(void* h, int which, tfs_contact_ext* out, uint32_t cap) {
    const Capture& c = ((Handle*)h)->cap;
    return copyOut(which == 0 ? c.preExt : c.oldExt, out, cap);
}

// Bodies before the step.
uint32_t tfr_cap_old_bodies(void* h, tfs_body* out, uint32_t cap) { return copyOut(((Handle*)h)->cap.oldBodies, out, cap); }

// The recorded joints (which: 0 at the solve block's start, 1 at its end) and the world body's v, w.
uint32_t tfr_cap_joints(void* h, int which, tfs_joint* out, uint32_t cap, float* world) {
    const Capture& c = ((Handle*)h)->cap;
    if (world) { std::memcpy(world, which ? c.postWorld : c.preWorld, sizeof c.preWorld); }
    return copyOut(which ? c.postJoints : c.preJoints, out, cap);
}

size_t tfr_sizeof_joint(void) { return sizeof(tfs_joint); }

// A joint as World::createJoint makes it (src/world.cpp:2221-2300): kind TFS_JOINT_*, body_a -1 for the
// world, and p[] = anchor A (3), anchor B (3, the distance joint's), axis (3), limited, lower, upper,
// motor speed, max motor force or torque, friction, length, hertz, damping ratio, break force, break
// torque, collide connected. Returns the joint's index, or -1.
int tfr_add_joint(void* h, int kind, int bodyA, int bodyB, const float* p) {
    World& w = ((Handle*)h)->world;
    JointDesc d;
    if (bodyA >= 0) { d.bodyA = BodyId{ (uint32_t)bodyA, w.impl_->bodies[bodyA].generation }; }
    d.bodyB = BodyId{ (uint32_t)bodyB, w.impl_->bodies[bodyB].generation };
    const Vec3 anchor{ p[0], p[1], p[2] }, anchorB{ p[3], p[4], p[5] }, axis{ p[6], p[7], p[8] };
    const bool limited = p[9] != 0.0f;
    switch (kind) {
        case TFS_JOINT_POINT: d.kind = PointJoint{ anchor }; break;
        case TFS_JOINT_FIXED: d.kind = FixedJoint{ anchor }; break;
        case TFS_JOINT_HINGE: d.kind = HingeJoint{ anchor, axis, limited, p[10], p[11], p[12], p[13], p[14] }; break;
        case TFS_JOINT_SLIDER: d.kind = SliderJoint{ anchor, axis, limited, p[10], p[11], p[12], p[13], p[14] }; break;
        case TFS_JOINT_DISTANCE: d.kind = DistanceJoint{ anchor, anchorB, p[15], p[16], p[17], limited, p[10], p[11] }; break;
        default: return -1;
    }
    d.breakForce = p[18];
    d.breakTorque = p[19];
    d.collideConnected = p[20] != 0.0f;
    const JointId id = w.createJoint(d);
    return id.index == UINT32_MAX ? -1 : (int)id.index;
}

uint32_t tfr_joints(void* h, tfs_joint* out, uint32_t cap) {
    std::vector<tfs_joint> v;
    for (const Joint& j : ((Handle*)h)->world.impl_->joints) { v.push_back(toJoint(j)); }
    return copyOut(v, out, cap);
}

// The step's candidate pairs and their manifolds.
uint32_t tfr_cap_pairs(void* h, tfs_pair* out, uint32_t cap) { return copyOut(((Handle*)h)->cap.pairs, out, cap); }
uint32_t tfr_cap_manifolds(void* h, tfs_manifold* out, uint32_t cap) { return copyOut(((Handle*)h)->cap.manifolds, out, cap); }

// Recorded lists: 0 active, 1 fast, 2 fastSolid, 3 impact, 4 byColor, 5 rest, 6 impact byColor,
// 7 impact rest.
uint32_t tfr_cap_list(void* h, int which, uint32_t* out, uint32_t cap) {
    const Capture& c = ((Handle*)h)->cap;
    const std::vector<uint32_t>* lists[8] = { &c.active, &c.fast, &c.fastSolid, &c.impact,
                                              &c.byColor, &c.rest, &c.impactByColor, &c.impactRest };
    if (which < 0 || which > 7) { return 0; }
    return copyOut(*lists[which], out, cap);
}

// Colour of each contact index (kColors: none). Entries of inactive contacts are 255.
uint32_t tfr_cap_colors(void* h, uint8_t* out, uint32_t cap) {
    Capture& c = ((Handle*)h)->cap;
    std::vector<uint8_t> v(c.preContacts.size(), 255);
    for (uint32_t ci : c.active) { v[ci] = c.colors[ci]; }
    return copyOut(v, out, cap);
}

// The SolverContext of the step: h, invH, contact soft (3), static soft (3), maxPushout, restitution threshold.
void tfr_cap_ctx(void* h, float* out) { std::memcpy(out, ((Handle*)h)->cap.ctx, sizeof(float) * 10); }

uint32_t tfr_cap_num_refreshes(void* h) { return (uint32_t)((Handle*)h)->cap.refreshes.size(); }

// Refresh k: its substep and remaining time, and the again list with the fresh manifold of each entry.
uint32_t tfr_cap_refresh(void* h, uint32_t k, int* substep, float* remaining, uint32_t* again,
                         tfs_contact* fresh, uint32_t cap) {
    const Refresh& r = ((Handle*)h)->cap.refreshes[k];
    if (substep) { *substep = r.substep; }
    if (remaining) { *remaining = r.remaining; }
    copyOut(r.again, again, cap);
    return copyOut(r.fresh, fresh, cap);
}

size_t tfr_sizeof_body(void) { return sizeof(tfs_body); }
size_t tfr_sizeof_contact(void) { return sizeof(tfs_contact); }

// Each body's shape as tfs_shape (include/twofold_solver.h) for the device refresh: the reference's
// sizes after buildShape, for bodies with exactly one shape at the body's origin. Kinds: 0 sphere,
// 1 box, 4 half-space; 255 for anything else (another geometry, several shapes, a shape off the origin).
uint32_t tfr_body_shapes(void* h, tfs_shape* out, uint32_t cap) {
    const World::Impl& w = *((Handle*)h)->world.impl_;
    const uint32_t n = (uint32_t)w.bodies.size();
    for (uint32_t i = 0; i < n && i < cap; ++i) {
        tfs_shape& o = out[i];
        std::memset(&o, 0, sizeof o);
        o.kind = 255;
        const Body& b = w.bodies[i];
        if (b.shapes.size() != 1) { continue; }
        const Shape& s = w.shapes[b.shapes[0]];
        const bool atOrigin = s.local.p.x == 0.0f && s.local.p.y == 0.0f && s.local.p.z == 0.0f && s.local.q.x == 0.0f &&
                              s.local.q.y == 0.0f && s.local.q.z == 0.0f && s.local.q.w == 1.0f;
        if (!atOrigin) { continue; }
        if (s.kind == ShapeKind::Sphere) {
            o.kind = 0;
            o.radius = s.radius;
        } else if (const Box* bx = std::get_if<Box>(&s.geometry)) {
            o.kind = 1;
            put3(o.half, bx->halfExtents);
        } else if (s.kind == ShapeKind::HalfSpace) {
            o.kind = 4;
        }
    }
    return n;
}

// The reference's own narrowphase on one pair of shapes, for checking a restatement of it: kind 0 sphere
// (dims[0] = radius), 1 box (dims = half-extents), 4 half-space; pa, pb the shapes' world transforms
// (p[3], q[4]); motion = v, w and centre of A, then of B (18 floats), over dt, fast as Motion::fast.
// The shapes go through buildShape (src/shape.cpp:149), so their sizes are the ones the reference keeps;
// solid[0..3] / solid[4..7] receive them (half-extents, radius). out = normal[3], count, then per point
// position[3], separation, predicted. Returns 0, or -1 for another kind.
int tfr_collide(int ka, const float* da, const float* pa, int kb, const float* db, const float* pb, float margin,
                float linearSlop, const float* motion, float dt, int fast, float* solid, float* out) {
    Shape s[2];
    const int kinds[2] = { ka, kb };
    const float* dims[2] = { da, db };
    for (int i = 0; i < 2; ++i) {
        if (kinds[i] == 0) {
            s[i].geometry = Sphere{ dims[i][0] };
        } else if (kinds[i] == 1) {
            s[i].geometry = Box{ Vec3{ dims[i][0], dims[i][1], dims[i][2] } };
        } else if (kinds[i] == 4) {
            s[i].geometry = HalfSpace{};
        } else {
            return -1;
        }
        buildShape(s[i]);
        float* o = solid + 4 * i;
        o[0] = o[1] = o[2] = 0.0f;
        o[3] = s[i].radius;
        if (const Box* b = std::get_if<Box>(&s[i].geometry)) { put3(o, b->halfExtents); }
    }
    const Transform xa{ vec3(pa), quat4(pa + 3) }, xb{ vec3(pb), quat4(pb + 3) };
    Tolerances tol;
    tol.linearSlop = linearSlop;
    Motion mo{ vec3(motion), vec3(motion + 3), vec3(motion + 6), vec3(motion + 9), vec3(motion + 12),
               vec3(motion + 15), dt, fast != 0 };
    Manifold m;
    collide(s[0], xa, s[1], xb, margin, tol, mo, m);
    put3(out, m.normal);
    out[3] = (float)m.count;
    for (int i = 0; i < 4; ++i) {
        put3(out + 4 + 5 * i, m.points[i].position);
        out[4 + 5 * i + 3] = m.points[i].separation;
        out[4 + 5 * i + 4] = m.points[i].predicted;
    }
    return 0;
}

}
