// stages.h - the numerics of each stage of the solve path, per body or per contact.
//
// The kernels in solver.hip load a body's or a row's data into the structs below, call one of these
// functions and store the result. Each function takes the float operations of the reference
// function it cites, in the same order. Point axes (src/solver.cpp:5-17) are rebuilt where they are
// used from the same body matrices and anchors the reference builds them from once per step; that
// gives the same bits and reads far fewer bytes than storing them.
#pragma once

#include "../../include/twofold_solver.h"
#include "fmath.h"

namespace tf {

// The points of a contact, i = 0 .. n-1 of at most 4, as a loop unrolled over the 4 slots: every
// index is known at compile time, so a contact held in registers stays there. A `continue` in the
// body skips to the next point.
#define TF_POINTS(i, n) _Pragma("unroll") for (int i = 0; i < 4; ++i) if (i < (n))

// Row flags: point count, which bodies move, whether the contact is solved with the static
// softness (src/lanes.cpp:77-84).
enum : uint32_t { kCountMask = 7u, kMoveA = 8u, kMoveB = 16u };

// A body's motion during the step, as the solve passes read and write it (src/lanes.cpp:112-131).
struct Motion {
    V3 v, w, dp;
    Q4 dq;
};

// A body's inverse mass data, fixed between inverse-mass updates (src/world.cpp:62-92).
struct Mass {
    M3 imw, iiw;
    float invMass, pim, pii;
};

struct Soft {
    float biasRate, massScale, impulseScale;
};

struct Ctx {
    float h, invH;
    Soft contactSoft, staticSoft;
    float maxPushout, restitutionThreshold, maxSpeed, maxSpin;
    V3 gravity;
};

TF_HD bool isPlanarLocks(uint32_t locks) { return (locks & TFS_LOCK_PLANAR) == TFS_LOCK_PLANAR; }
TF_HD bool movingBody(const tfs_body& b) { return b.type == TFS_DYNAMIC && b.awake; }

// makeSoft: src/solver.cpp:255-262.
TF_HD Soft makeSoft(float hertz, float zeta, float h) {
    if (hertz == 0.0f) { return Soft{ 0.0f, 1.0f, 0.0f }; }
    const float omega = 2.0f * 3.14159265358979f * hertz;
    const float a1 = 2.0f * zeta + h * omega;
    const float a2 = h * omega * a1;
    const float a3 = 1.0f / (1.0f + a2);
    return Soft{ omega / a1, a2 * a3, a3 };
}

// SolverContext: src/world.cpp:1634-1642, plus the speed limits of src/world.cpp:1142, 1154-1155.
inline Ctx makeCtx(const tfs_settings& s, float dt) {
    Ctx c;
    const int n = s.substeps > 1 ? s.substeps : 1;
    c.h = dt / (float)n;
    c.invH = 1.0f / c.h;
    const float hertz = minf(s.contact_hertz, 0.25f * c.invH);
    c.contactSoft = makeSoft(hertz, s.contact_damping_ratio, c.h);
    c.staticSoft = makeSoft(2.0f * hertz, s.contact_damping_ratio, c.h);
    c.maxPushout = s.max_pushout_speed * s.length_unit;
    c.restitutionThreshold = s.restitution_threshold * s.length_unit;
    c.maxSpeed = s.max_linear_speed * s.length_unit;
    c.maxSpin = s.max_angular_speed;
    c.gravity = ld3(s.gravity);
    return c;
}

// ---- bodies --------------------------------------------------------------------------------------

// invertSubspace: src/world.cpp:44-60.
TF_HD M3 invertSubspace(const M3& m, const V3* axes, int k) {
    if (k == 0) { return mzero(); }
    if (k == 3) { return inverse(m); }
    if (k == 1) {
        const float d = dot(axes[0], m * axes[0]);
        return d > 0.0f ? outer(axes[0], axes[0]) * (1.0f / d) : mzero();
    }
    const float a = dot(axes[0], m * axes[0]);
    const float b = dot(axes[0], m * axes[1]);
    const float c = dot(axes[1], m * axes[0]);
    const float d = dot(axes[1], m * axes[1]);
    float det = a * d - b * c;
    if (det == 0.0f) { return mzero(); }
    det = 1.0f / det;
    return outer(axes[0], axes[0]) * (d * det) - outer(axes[0], axes[1]) * (b * det) -
           outer(axes[1], axes[0]) * (c * det) + outer(axes[1], axes[1]) * (a * det);
}

// updateInverseMass: src/world.cpp:62-92.
TF_HD void updateInverseMass(tfs_body& b) {
    const M3 lf = toMat3(ld4(b.lock_frame));
    V3 ang[3];
    int na = 0;
    M3 lp = mzero(), ap = mzero();
    for (int i = 0; i < 3; ++i) {
        const V3 axis = column(lf, i);
        if (!(b.locks & (TFS_LOCK_TX << i))) { lp = lp + outer(axis, axis); }
        if (!(b.locks & (TFS_LOCK_RX << i))) {
            ang[na++] = axis;
            ap = ap + outer(axis, axis);
        }
    }
    st9(b.linear_proj, lp);
    st9(b.angular_proj, ap);
    if (b.type != TFS_DYNAMIC || !b.awake) {
        st9(b.inv_mass_world, mzero());
        st9(b.inv_inertia_world, mzero());
        b.plane_inv_mass = 0.0f;
        b.plane_inv_inertia = 0.0f;
        return;
    }
    st9(b.inv_mass_world, lp * b.inv_mass);
    const M3 r = toMat3(ld4(b.q));
    const M3 iw = mul(mul(r, ld9(b.inertia)), transpose(r));
    const M3 inv = invertSubspace(iw, ang, na);
    st9(b.inv_inertia_world, inv);
    if (isPlanarLocks(b.locks)) {
        const V3 n = rotate(ld4(b.lock_frame), vec(0, 0, 1));
        b.plane_inv_mass = b.inv_mass;
        b.plane_inv_inertia = dot(n, inv * n);
    }
}

TF_HD Mass massOf(const tfs_body& b) {
    return Mass{ ld9(b.inv_mass_world), ld9(b.inv_inertia_world), b.inv_mass, b.plane_inv_mass, b.plane_inv_inertia };
}

// limitSpeed: src/solver.cpp:30-41.
TF_HD void limitSpeed(const tfs_body& b, V3& v, V3& w, float& kept, float maxSpeed, float maxSpin) {
    if (b.type != TFS_DYNAMIC || !b.awake) { return; }
    const float speed = length(v), spin = length(w);
    if (speed > maxSpeed) {
        v = v * (maxSpeed / speed);
        kept = minf(kept, maxSpeed / speed);
    }
    if (spin > maxSpin) {
        w = w * (maxSpin / spin);
        kept = minf(kept, maxSpin / spin);
    }
}

// integrateVelocities, one body: src/world.cpp:1141-1151.
TF_HD void integrateVelocity(const tfs_body& b, const Ctx& c, V3& v, V3& w, float& kept) {
    v = v + (c.gravity * b.gravity_scale + ld3(b.force) * b.inv_mass) * c.h;
    v = ld9(b.linear_proj) * (v * (1.0f / (1.0f + c.h * b.linear_damping)));
    w = w + (ld9(b.inv_inertia_world) * ld3(b.torque)) * c.h;
    w = ld9(b.angular_proj) * (w * (1.0f / (1.0f + c.h * b.angular_damping)));
    limitSpeed(b, v, w, kept, c.maxSpeed, c.maxSpin);
}

// integratePositions, one body, after limitSpeeds: src/world.cpp:1161-1172.
TF_HD void integratePosition(tfs_body& b, Motion& m, float h) {
    const V3 vh = m.v * h;
    const V3 com = ld3(b.com) + vh;
    st3(b.com, com);
    m.dp = m.dp + vh;
    const Q4 q = integrate(ld4(b.q), m.w, h);
    st4(b.q, q);
    m.dq = q * conjugate(ld4(b.q_base));
    st3(b.p, com - rotate(q, ld3(b.local_center)));
}

// relativeTravel: src/world.cpp:484-486.
TF_HD float relativeTravel(const tfs_body& a, const tfs_body& b, float dt) {
    return (length(ld3(b.v) - ld3(a.v)) + length(ld3(a.w)) * a.extent + length(ld3(b.w)) * b.extent) * dt;
}

// movingAwake: src/world.cpp:94-98.
TF_HD bool movingAwake(const tfs_body& b, V3 v, V3 w) {
    if (b.type == TFS_DYNAMIC) { return b.awake; }
    if (b.type == TFS_KINEMATIC) { return dot(v, v) > 0.0f || dot(w, w) > 0.0f; }
    return false;
}

// ---- contact axes -------------------------------------------------------------------------------

struct Axis {
    V3 armA, armB, linearA, angularA, linearB, angularB;
    float mass;
};

// makeAxis: src/solver.cpp:5-17 (makeDirection, src/lanes.cpp:38-49, is the same per lane).
TF_HD Axis makeAxis(const Mass& a, const Mass& b, V3 ra, V3 rb, V3 d) {
    Axis x;
    x.armA = cross(ra, d);
    x.armB = cross(rb, d);
    x.linearA = a.imw * d;
    x.angularA = a.iiw * x.armA;
    x.linearB = b.imw * d;
    x.angularB = b.iiw * x.armB;
    const float k = dot(d, x.linearA) + dot(d, x.linearB) + dot(x.armA, x.angularA) + dot(x.armB, x.angularB);
    x.mass = k > 1e-6f * (a.invMass + b.invMass) ? 1.0f / k : 0.0f;
    return x;
}

// speedAlong: src/solver.cpp:19-21.
TF_HD float speedAlong(const Motion& a, const Motion& b, V3 d, const Axis& x) {
    return dot(d, b.v - a.v) + dot(x.armB, b.w) - dot(x.armA, a.w);
}

// push / pushMoving / the masked lane push: src/solver.cpp:23-28, 212-221; src/lanes.cpp:161-166.
TF_HD void push(Motion& a, Motion& b, const Axis& x, float impulse, bool moveA, bool moveB) {
    if (moveA) {
        a.v = a.v - x.linearA * impulse;
        a.w = a.w - x.angularA * impulse;
    }
    if (moveB) {
        b.v = b.v + x.linearB * impulse;
        b.w = b.w + x.angularB * impulse;
    }
}

// turned: src/solver.cpp:57-59 (src/lanes.cpp:168).
TF_HD V3 turned(const Motion& m, V3 r) { return m.dq.x == 0.0f && m.dq.y == 0.0f && m.dq.z == 0.0f ? r : rotate(m.dq, r); }

// ---- a coloured ("laned") contact row -----------------------------------------------------------

struct RowPoint {
    V3 ra, rb;
    float adj;       // adjusted separation
    float rv;        // relative velocity at prepare
    float ni;        // normal impulse
    V3 ti;           // tangent impulse
    float nimax;     // largest normal impulse this step
    float sep;       // the current separation, when the pass computed it beforehand (PRESEP)
};

struct Row {
    uint32_t flags;
    V3 n, t1, t2;
    float friction, restitution;
};

// The row functions unroll over N = 4 point slots, so every point index is known at compile time and
// the points stay in registers, and solve the first `count` of them: a wave whose rows have different
// point counts runs one body of code, masked per lane, rather than one per count (up to 1 + 2 + 3 + 4
// point solves). They also take an axis source: ax(i, d) gives point i's axis
// along the normal (d = 0), tangent1 (1) or tangent2 (2). The kernels read axes that the packing
// kernel built once per step, as prepareBatch builds them (src/lanes.cpp:98-109); the serial path
// builds them on the spot from the same masses and anchors (makeAxis), which gives the same bits.

// The axes of a row built from the bodies' masses.
struct MadeAxes {
    const Mass& ma;
    const Mass& mb;
    const Row& r;
    const RowPoint* p;
    TF_HD Axis operator()(int i, int d) const {
        return makeAxis(ma, mb, p[i].ra, p[i].rb, d == 0 ? r.n : d == 1 ? r.t1 : r.t2);
    }
};

// warmStartLanes, one lane: src/lanes.cpp:195-217.
template <int N, class AX>
TF_HD void warmRowN(const Row& r, const RowPoint* p, const AX& ax, Motion& a, Motion& b, int count) {
    const bool moveA = r.flags & kMoveA, moveB = r.flags & kMoveB;
#pragma unroll
    for (int i = 0; i < N; ++i) {
        if (i >= count) { continue; }
        push(a, b, ax(i, 0), p[i].ni, moveA, moveB);
        push(a, b, ax(i, 1), dot(p[i].ti, r.t1), moveA, moveB);
        push(a, b, ax(i, 2), dot(p[i].ti, r.t2), moveA, moveB);
    }
}

// The current separation of a point (solveLanes, src/lanes.cpp:236-240): its anchors turned by the
// bodies' rotation this step, plus their travel dp = b.dp - a.dp, along the normal.
TF_HD float pointSeparation(const Motion& a, const Motion& b, V3 dp, const RowPoint& p, V3 n) {
    const V3 d = dp + turned(b, p.rb) - turned(a, p.ra);
    return dot(d, n) + p.adj;
}

// What the reverse sweep needs of one point's normal solve: the scales its bias branch picked, d bias
// / d sep, and whether the clamp at zero let the impulse through.
struct NormalTrace {
    float massScale, impulseScale, dbias;
    bool active;
};

// One point's normal row (src/lanes.cpp:241-262).
template <bool BIAS>
TF_HD NormalTrace normalPoint(RowPoint& p, float sep, const Axis& x, V3 n, Motion& a, Motion& b, const Ctx& c,
                              const Soft& soft, bool moveA, bool moveB) {
    float bias = 0.0f, massScale = 1.0f, impulseScale = 0.0f, dbias = 0.0f;
    if (sep > 0.0f) {
        bias = sep * c.invH;
        dbias = c.invH;
    } else if (BIAS) {
        bias = maxf(soft.biasRate * sep, -c.maxPushout);
        massScale = soft.massScale;
        impulseScale = soft.impulseScale;
        dbias = soft.biasRate * sep > -c.maxPushout ? soft.biasRate : 0.0f;
    }
    const float vn = speedAlong(a, b, n, x);
    float impulse = -x.mass * massScale * (vn + bias) - impulseScale * p.ni;
    const float sum = p.ni + impulse;
    const float total = maxq(sum, 0.0f);
    impulse = total - p.ni;
    p.ni = total;
    p.nimax = maxq(p.nimax, impulse);
    push(a, b, x, impulse, moveA, moveB);
    return NormalTrace{ massScale, impulseScale, dbias, sum > 0.0f };
}

// What the reverse sweep needs of one point's friction solve: the unclamped forces, their length,
// the cone's radius and whether the forces were scaled onto it.
struct FrictionTrace {
    float f1, f2, mag, limit;
    bool scaled;
};

// One point's friction cone (src/lanes.cpp:263-304).
TF_HD FrictionTrace frictionPoint(RowPoint& p, const Axis& x1, const Axis& x2, const Row& r, Motion& a, Motion& b,
                                  bool moveA, bool moveB) {
    const float old1 = dot(p.ti, r.t1), old2 = dot(p.ti, r.t2);
    float f1 = old1 - x1.mass * speedAlong(a, b, r.t1, x1);
    float f2 = old2 - x2.mass * speedAlong(a, b, r.t2, x2);
    const float limit = r.friction * p.ni;
    const float mag = fsqrt(f1 * f1 + f2 * f2);
    const FrictionTrace t{ f1, f2, mag, limit, mag > limit };
    if (mag > limit) {
        const float scale = mag > 0.0f ? limit / mag : 0.0f;
        f1 *= scale;
        f2 *= scale;
    }
    p.ti = r.t1 * f1 + r.t2 * f2;
    push(a, b, x1, f1 - old1, moveA, moveB);
    push(a, b, x2, f2 - old2, moveA, moveB);
    return t;
}

// The 3D contact solve: solveLanes for one lane (src/lanes.cpp:220-315), which is also the 3D part
// of solveContacts (src/solver.cpp:370-417). PRESEP: each point's separation (pointSeparation of the
// same bodies' dp and dq, which do not change during a pass) was computed before the pass, in p[i].sep.
template <int N, bool BIAS, class AX, bool PRESEP = false>
TF_HD void solveRowN(const Row& r, RowPoint* p, const AX& ax, Motion& a, Motion& b, const Ctx& c, int count) {
    const bool moveA = r.flags & kMoveA, moveB = r.flags & kMoveB;
    const Soft soft = (moveA && moveB) ? c.contactSoft : c.staticSoft;
    const V3 dp = b.dp - a.dp;
#pragma unroll
    for (int i = 0; i < N; ++i) {
        if (i >= count) { continue; }
        const float sep = PRESEP ? p[i].sep : pointSeparation(a, b, dp, p[i], r.n);
        normalPoint<BIAS>(p[i], sep, ax(i, 0), r.n, a, b, c, soft, moveA, moveB);
    }
#pragma unroll
    for (int j = 0; j < N; ++j) {
        const int i = BIAS ? j : N - 1 - j;
        if (i >= count) { continue; }
        frictionPoint(p[i], ax(i, 1), ax(i, 2), r, a, b, moveA, moveB);
    }
}

// ---- the adjoint of a row ----------------------------------------------------------------------

// Adjoints of the two bodies' velocities.
struct MotionBar {
    V3 va, wa, vb, wb;
};

// d push / d impulse, dotted with the velocity adjoints: how the loss moves with the impulse a push
// applies (push, src/solver.cpp:23-28, masked as the lanes mask it).
TF_HD float pushBar(const Axis& x, const MotionBar& g, bool moveA, bool moveB) {
    float s = 0.0f;
    if (moveA) { s = s - dot(x.linearA, g.va) - dot(x.angularA, g.wa); }
    if (moveB) { s = s + dot(x.linearB, g.vb) + dot(x.angularB, g.wb); }
    return s;
}

// speedAlong (src/solver.cpp:19-21) is linear in the velocities: its adjoint s adds J^T s.
TF_HD void speedBar(const Axis& x, V3 d, float s, MotionBar& g) {
    g.va = g.va - d * s;
    g.wa = g.wa - x.armA * s;
    g.vb = g.vb + d * s;
    g.wb = g.wb + x.armB * s;
}

// The reverse sweep of solveRowN for one row. Inputs are what the forward solve saw: the bodies'
// velocities a, b, the points' impulses p (ni, ti) and separations sep. The row is solved again from
// them (the same operations, so the same clamps), then the adjoints of its outputs, g (velocities),
// niB and tiB (impulses), are turned into those of its inputs, in place; sepB receives d loss /
// d separation per point, and muB adds d loss / d friction. The friction cone's projection
// f * limit / |f| has the derivative (limit / |f|) (I - f^ f^T) in f and f^ in the limit; with a zero
// limit the cone is a point and the forces do not depend on f at all.
template <bool BIAS, class AX>
TF_HD void solveRowAdjoint(const Row& r, RowPoint* p, const float* sep, const AX& ax, Motion a, Motion b, const Ctx& c,
                           int count, MotionBar& g, float* niB, V3* tiB, float* sepB, float& muB) {
    const bool moveA = r.flags & kMoveA, moveB = r.flags & kMoveB;
    const Soft soft = (moveA && moveB) ? c.contactSoft : c.staticSoft;
    NormalTrace nt[4];
    FrictionTrace ft[4];
    float lamF[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        if (i >= count) { continue; }
        nt[i] = normalPoint<BIAS>(p[i], sep[i], ax(i, 0), r.n, a, b, c, soft, moveA, moveB);
    }
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const int i = BIAS ? j : 3 - j;
        if (i >= count) { continue; }
        lamF[i] = p[i].ni;
        ft[i] = frictionPoint(p[i], ax(i, 1), ax(i, 2), r, a, b, moveA, moveB);
    }
#pragma unroll
    for (int j = 3; j >= 0; --j) {
        const int i = BIAS ? j : 3 - j;
        if (i >= count) { continue; }
        const Axis x1 = ax(i, 1);
        const Axis x2 = ax(i, 2);
        const float d1 = pushBar(x1, g, moveA, moveB), d2 = pushBar(x2, g, moveA, moveB);
        const float g1 = d1 + dot(r.t1, tiB[i]), g2 = d2 + dot(r.t2, tiB[i]);
        const FrictionTrace& t = ft[i];
        float f1b = g1, f2b = g2, limB = 0.0f;
        if (t.scaled) {
            const float h1 = t.mag > 0.0f ? t.f1 / t.mag : 0.0f, h2 = t.mag > 0.0f ? t.f2 / t.mag : 0.0f;
            const float along = h1 * g1 + h2 * g2;
            limB = along;
            const float s = t.mag > 0.0f ? t.limit / t.mag : 0.0f;
            f1b = s * (g1 - h1 * along);
            f2b = s * (g2 - h2 * along);
        }
        if (!(t.limit > 0.0f)) {
            f1b = 0.0f;
            f2b = 0.0f;
        }
        muB += lamF[i] * limB;
        niB[i] += r.friction * limB;
        speedBar(x1, r.t1, -x1.mass * f1b, g);
        speedBar(x2, r.t2, -x2.mass * f2b, g);
        tiB[i] = r.t1 * (f1b - d1) + r.t2 * (f2b - d2);
    }
#pragma unroll
    for (int i = 3; i >= 0; --i) {
        if (i >= count) { continue; }
        const Axis x = ax(i, 0);
        const float d = pushBar(x, g, moveA, moveB);
        const float totalB = niB[i] + d;
        float inB = -d, impulseB = 0.0f;
        if (nt[i].active) {
            inB += totalB;
            impulseB = totalB;
        }
        const float vnB = -x.mass * nt[i].massScale * impulseB;
        inB -= nt[i].impulseScale * impulseB;
        speedBar(x, r.n, vnB, g);
        sepB[i] = nt[i].dbias * vnB;
        niB[i] = inB;
    }
}

// restituteLanes, one lane: src/lanes.cpp:318-352.
template <int N, class AX>
TF_HD void restituteRowN(const Row& r, RowPoint* p, const AX& ax, Motion& a, Motion& b, const Ctx& c, int count) {
    const bool moveA = r.flags & kMoveA, moveB = r.flags & kMoveB;
    const float e = r.restitution;
#pragma unroll
    for (int i = 0; i < N; ++i) {
        if (i >= count) { continue; }
        const bool on = !(e == 0.0f) && !(p[i].rv > -c.restitutionThreshold) && !(p[i].nimax == 0.0f);
        const Axis x = ax(i, 0);
        const float vn = speedAlong(a, b, r.n, x);
        float impulse = -x.mass * (vn + e * p[i].rv);
        const float total = maxq(p[i].ni + impulse, 0.0f);
        impulse = total - p[i].ni;
        if (on) {
            p[i].ni = total;
            p[i].nimax = maxq(p[i].nimax, impulse);
        }
        push(a, b, x, impulse, on && moveA, on && moveB);
    }
}

enum RowMode { kWarm = 0, kBias = 1, kRelax = 2, kRestitute = 3 };

// One row in the given mode: its first (flags & kCountMask) points, at most 4. In relax mode the
// friction half runs from the last point down, so skipping the empty slots 3..count gives the order
// count-1..0 that a solve over exactly `count` points has.
template <int MODE, class AX, bool PRESEP = false>
TF_HD void rows(const Row& r, RowPoint* p, const AX& ax, Motion& a, Motion& b, const Ctx& c) {
    const int count = (int)(r.flags & kCountMask);
    if (MODE == kWarm) {
        warmRowN<4>(r, p, ax, a, b, count);
    } else if (MODE == kRestitute) {
        restituteRowN<4>(r, p, ax, a, b, c, count);
    } else {
        solveRowN<4, MODE == kBias, AX, PRESEP>(r, p, ax, a, b, c, count);
    }
}

// ---- contacts outside the colours (planar, rolling resistance, no colour left): one at a time -----

// A body as prepareContact and the serial stages read it.
struct BodyView {
    uint8_t type, awake, planar;
    V3 planeNormal;
};

TF_HD BodyView viewOf(const tfs_body& b) {
    return BodyView{ b.type, b.awake, (uint8_t)(isPlanarLocks(b.locks) ? 1 : 0),
                     rotate(ld4(b.lock_frame), vec(0, 0, 1)) };
}
TF_HD bool anchoredView(const BodyView& b) { return !(b.type == TFS_DYNAMIC && b.awake); }

// planarContact: src/solver.cpp:66-79; n is written as the reference writes it.
TF_HD bool planarContact(const BodyView& a, const BodyView& b, V3 normal, V3& n) {
    const BodyView* first = nullptr;
    const BodyView* xs[2] = { &a, &b };
    for (int k = 0; k < 2; ++k) {
        const BodyView* x = xs[k];
        if (x->type == TFS_STATIC) { continue; }
        if (x->type == TFS_KINEMATIC || !x->planar) { return false; }
        if (!first) {
            first = x;
            n = x->planeNormal;
        } else if (fabs_(dot(n, x->planeNormal)) < 0.9999f) {
            return false;
        }
    }
    return first && fabs_(dot(normal, n)) < 1e-3f;
}

// planarMass: src/solver.cpp:81-84.
TF_HD float planarMass(const Mass& a, const Mass& b, float armA, float armB) {
    const float k = a.pim + b.pim + a.pii * armA * armA + b.pii * armB * armB;
    return k > 1e-6f * (a.invMass + b.invMass) ? 1.0f / k : 0.0f;
}

struct PlanarAxes {
    float normalA, normalB, normalMass, tangentA, tangentB, tangentMass;
};

// makePlanarAxes, point i: src/solver.cpp:86-98.
TF_HD PlanarAxes planarAxes(const Mass& a, const Mass& b, const tfs_contact& c, int i) {
    const V3 n = ld3(c.plane);
    const tfs_point& p = c.points[i];
    PlanarAxes x;
    x.normalA = dot(n, cross(ld3(p.anchor_a), ld3(c.normal)));
    x.normalB = dot(n, cross(ld3(p.anchor_b), ld3(c.normal)));
    x.normalMass = planarMass(a, b, x.normalA, x.normalB);
    x.tangentA = dot(n, cross(ld3(p.anchor_a), ld3(c.tangent1)));
    x.tangentB = dot(n, cross(ld3(p.anchor_b), ld3(c.tangent1)));
    x.tangentMass = planarMass(a, b, x.tangentA, x.tangentB);
    return x;
}

// PlanarState: src/solver.cpp:101-129.
struct PlanarState {
    V3 va, vb;
    float wa, wb, ma, mb, ia, ib;

    TF_HD PlanarState(const Motion& a, const Motion& b, const Mass& x, const Mass& y, V3 n)
        : va(a.v), vb(b.v), wa(dot(a.w, n)), wb(dot(b.w, n)), ma(x.pim), mb(y.pim), ia(x.pii), ib(y.pii) {}
    TF_HD float speed(V3 d, float armA, float armB) const { return dot(d, vb - va) + armB * wb - armA * wa; }
    TF_HD void push(V3 d, float armA, float armB, float impulse) {
        va = va - d * (ma * impulse);
        wa -= ia * armA * impulse;
        vb = vb + d * (mb * impulse);
        wb += ib * armB * impulse;
    }
    TF_HD void store(Motion& a, Motion& b, V3 n, bool moveA, bool moveB) const {
        if (moveA) {
            a.v = va;
            a.w = n * wa;
        }
        if (moveB) {
            b.v = vb;
            b.w = n * wb;
        }
    }
};

// prepareContact: src/solver.cpp:264-291.
TF_HD void prepareContact(const BodyView& va, const BodyView& vb, const tfs_body& ta, const tfs_body& tb,
                          const Motion& a, const Motion& b, const Mass& ma, const Mass& mb, tfs_contact& c) {
    V3 plane = ld3(c.plane);
    c.planar = planarContact(va, vb, ld3(c.normal), plane) ? 1 : 0;
    st3(c.plane, plane);
    if (c.planar) {
        const V3 n = normalize(ld3(c.normal) - plane * dot(ld3(c.normal), plane));
        st3(c.normal, n);
        st3(c.tangent1, cross(plane, n));
        st3(c.tangent2, plane);
    }
    const PlanarState st(a, b, ma, mb, plane);
    const V3 normal = ld3(c.normal);
    TF_POINTS(i, c.count) {
        tfs_point& p = c.points[i];
        const V3 ra = ld3(p.anchor_a), rb = ld3(p.anchor_b);
        p.adjusted_separation = p.separation - dot(rb - ra, normal);
        if (c.planar) {
            const PlanarAxes x = planarAxes(ma, mb, c, i);
            p.relative_velocity = st.speed(normal, x.normalA, x.normalB);
        } else {
            p.relative_velocity = dot(normal, b.v - a.v) + dot(cross(rb, normal), b.w) - dot(cross(ra, normal), a.w);
        }
        p.max_normal_impulse = 0.0f;
    }
    if (c.rolling_resistance > 0.0f && c.count > 0) {
        float ha = 0.0f, hb = 0.0f;
        TF_POINTS(i, c.count) {
            ha += fabs_(dot(ld3(c.points[i].anchor_a), normal));
            hb += fabs_(dot(ld3(c.points[i].anchor_b), normal));
        }
        if (ta.type != TFS_DYNAMIC) { ha = 0.0f; }
        if (tb.type != TFS_DYNAMIC) { hb = 0.0f; }
        c.rolling_radius = maxf(ha, hb) / (float)c.count;
    }
}

// rebaseContact: src/solver.cpp:293-306.
TF_HD void rebaseContact(const Motion& a, const Motion& b, tfs_contact& c) {
    const float shift = dot(b.dp - a.dp, ld3(c.normal));
    TF_POINTS(i, c.count) {
        tfs_point& p = c.points[i];
        st3(p.anchor_a, rotate(a.dq, ld3(p.anchor_a)));
        st3(p.anchor_b, rotate(b.dq, ld3(p.anchor_b)));
        p.adjusted_separation += shift;
    }
}

// planarTurn: src/solver.cpp:132-135.
TF_HD void planarTurn(const Motion& m, V3 n, float& c, float& s) {
    const Q4 q = m.dq;
    c = q.w * q.w - (q.x * q.x + q.y * q.y + q.z * q.z);
    s = 2.0f * q.w * dot(n, vec(q.x, q.y, q.z));
}

// warmStart's effect (src/solver.cpp:319-358) of one contact that is not laned on one of its bodies
// (sideB: body B), in the order warmStart applies it to that body: a planar contact changes a body
// only when something is held and the body moves (PlanarState::push, ::store; the arms are the
// contact's planar axes), a 3D contact pushes both bodies (push with both flags set) and turns them by
// the rolling impulse.
TF_HD void warmRestSide(const tfs_contact& c, bool sideB, bool moves, const Mass& m, Motion& x, const PlanarAxes* ax) {
    const V3 n = ld3(c.normal), t1 = ld3(c.tangent1), t2 = ld3(c.tangent2);
    if (c.planar) {
        bool held = c.rolling_resistance > 0.0f;
        TF_POINTS(i, c.count) {
            const V3 ti = ld3(c.points[i].tangent_impulse);
            held = held || c.points[i].normal_impulse != 0.0f || dot(ti, ti) != 0.0f;
        }
        if (!held || !moves) { return; }
        const V3 plane = ld3(c.plane);
        V3 v = x.v;
        float w = dot(x.w, plane);
        TF_POINTS(i, c.count) {
            const tfs_point& p = c.points[i];
            const float armN = sideB ? ax[i].normalB : ax[i].normalA, armT = sideB ? ax[i].tangentB : ax[i].tangentA;
            const float in = p.normal_impulse, it = dot(ld3(p.tangent_impulse), t1);
            if (sideB) {
                v = v + n * (m.pim * in);
                w += m.pii * armN * in;
                v = v + t1 * (m.pim * it);
                w += m.pii * armT * it;
            } else {
                v = v - n * (m.pim * in);
                w -= m.pii * armN * in;
                v = v - t1 * (m.pim * it);
                w -= m.pii * armT * it;
            }
        }
        if (c.rolling_resistance > 0.0f) {
            const float r = dot(ld3(c.rolling_impulse), plane);
            if (sideB) {
                w += m.pii * r;
            } else {
                w -= m.pii * r;
            }
        }
        x.v = v;
        x.w = plane * w;
        return;
    }
    TF_POINTS(i, c.count) {
        const tfs_point& p = c.points[i];
        const V3 r = sideB ? ld3(p.anchor_b) : ld3(p.anchor_a);
        const V3 dirs[3] = { n, t1, t2 };
        const float imp[3] = { p.normal_impulse, dot(ld3(p.tangent_impulse), t1), dot(ld3(p.tangent_impulse), t2) };
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            const V3 lin = m.imw * dirs[d];
            const V3 ang = m.iiw * cross(r, dirs[d]);
            if (sideB) {
                x.v = x.v + lin * imp[d];
                x.w = x.w + ang * imp[d];
            } else {
                x.v = x.v - lin * imp[d];
                x.w = x.w - ang * imp[d];
            }
        }
    }
    if (c.rolling_resistance > 0.0f) {
        if (sideB) {
            x.w = x.w + m.iiw * ld3(c.rolling_impulse);
        } else {
            x.w = x.w - m.iiw * ld3(c.rolling_impulse);
        }
    }
}

// What a planar contact's normal rows take from the bodies' travel this pass (dp, dq), which a velocity
// pass does not change (src/solver.cpp:143-166): per point the bias of its separation, and the factors
// of its impulse, scale = -normalMass * massScale (the first product of the impulse's expression,
// evaluated left to right) and impulseScale.
struct PlanarRowPre {
    float scale[4], bias[4], impulseScale[4];
};

TF_HD void planarPre(const tfs_contact& c, bool moveA, bool moveB, const Motion& a, const Motion& b, const Ctx& ctx,
                     const PlanarAxes* axes, bool useBias, PlanarRowPre& out) {
    const Soft soft = (!moveA || !moveB) ? ctx.staticSoft : ctx.contactSoft;
    const V3 normal = ld3(c.normal), plane = ld3(c.plane);
    const float shift = dot(b.dp - a.dp, normal);
    float tax, tay, tbx, tby;
    planarTurn(a, plane, tax, tay);
    planarTurn(b, plane, tbx, tby);
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const PlanarAxes& x = axes[i];
        const float reachA = tax * x.tangentA + tay * x.normalA;
        const float reachB = tbx * x.tangentB + tby * x.normalB;
        const float sep = shift + reachB - reachA + c.points[i].adjusted_separation;
        float bias = 0.0f, massScale = 1.0f, impulseScale = 0.0f;
        if (sep > 0.0f) {
            bias = sep * ctx.invH;
        } else if (useBias) {
            bias = maxf(soft.biasRate * sep, -ctx.maxPushout);
            massScale = soft.massScale;
            impulseScale = soft.impulseScale;
        }
        out.scale[i] = -x.normalMass * massScale;
        out.bias[i] = bias;
        out.impulseScale[i] = impulseScale;
    }
}

// solvePlanar: src/solver.cpp:137-209. PRE: the normal rows' biases and factors come from planarPre of
// the same pass (the same operations, done before the pass).
template <bool useBias, bool PRE = false>
TF_HD void solvePlanar(tfs_contact& c, const BodyView& va, const BodyView& vb, const Mass& ma, const Mass& mb,
                       Motion& a, Motion& b, const Ctx& ctx, const PlanarAxes* axes, const PlanarRowPre* pre = nullptr) {
    const bool moveA = !anchoredView(va), moveB = !anchoredView(vb);
    const V3 normal = ld3(c.normal), tangent1 = ld3(c.tangent1), plane = ld3(c.plane);
    PlanarRowPre local;
    if (!PRE) { planarPre(c, moveA, moveB, a, b, ctx, axes, useBias, local); }
    const PlanarRowPre& q = PRE ? *pre : local;
    PlanarState st(a, b, ma, mb, plane);
    bool changed = false;
    TF_POINTS(i, c.count) {
        tfs_point& p = c.points[i];
        const PlanarAxes& x = axes[i];
        const float vn = st.speed(normal, x.normalA, x.normalB);
        float impulse = q.scale[i] * (vn + q.bias[i]) - q.impulseScale[i] * p.normal_impulse;
        const float total = maxq(p.normal_impulse + impulse, 0.0f);
        impulse = total - p.normal_impulse;
        p.normal_impulse = total;
        p.max_normal_impulse = maxq(p.max_normal_impulse, impulse);
        if (impulse != 0.0f) {
            st.push(normal, x.normalA, x.normalB, impulse);
            changed = true;
        }
    }
    // relax: from the last point down; the unrolled slots keep every index known at compile time
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const int i = useBias ? j : 3 - j;
        if (i >= c.count) { continue; }
        tfs_point& p = c.points[i];
        const PlanarAxes& x = axes[i];
        const float old = dot(ld3(p.tangent_impulse), tangent1);
        if (p.normal_impulse == 0.0f && old == 0.0f) { continue; }
        changed = true;
        const float limit = c.friction * p.normal_impulse;
        const float t = minq(maxq(old - x.tangentMass * st.speed(tangent1, x.tangentA, x.tangentB), -limit), limit);
        st3(p.tangent_impulse, tangent1 * t);
        st.push(tangent1, x.tangentA, x.tangentB, t - old);
    }
    if (c.rolling_resistance > 0.0f) {
        changed = true;
        float total = dot(ld3(c.rolling_impulse), plane);
        const float spin = st.wb - st.wa;
        const float k = st.ia + st.ib;
        if (spin != 0.0f && k > 0.0f) {
            const float lambda = -spin / k;
            total += lambda;
            st.wa -= st.ia * lambda;
            st.wb += st.ib * lambda;
        }
        float normalSum = 0.0f;
        TF_POINTS(i, c.count) { normalSum += c.points[i].normal_impulse; }
        const float limit = c.rolling_resistance * c.rolling_radius * normalSum;
        const float held = minq(maxq(total, -limit), limit);
        st.wa -= st.ia * (held - total);
        st.wb += st.ib * (held - total);
        st3(c.rolling_impulse, plane * held);
    }
    if (changed) { st.store(a, b, plane, moveA, moveB); }
}

// resistTurning: src/solver.cpp:225-251.
TF_HD void resistTurning(tfs_contact& c, const Mass& ma, const Mass& mb, Motion& a, Motion& b) {
    const V3 spin = b.w - a.w;
    const float speed = length(spin);
    V3 total = ld3(c.rolling_impulse);
    if (speed > 0.0f) {
        const V3 d = spin * (1.0f / speed);
        const V3 ia = ma.iiw * d, ib = mb.iiw * d;
        const float k = dot(d, ia) + dot(d, ib);
        if (k > 0.0f) {
            const float lambda = -speed / k;
            total = total + d * lambda;
            a.w = a.w - ia * lambda;
            b.w = b.w + ib * lambda;
        }
    }
    float normalSum = 0.0f;
    TF_POINTS(i, c.count) { normalSum += c.points[i].normal_impulse; }
    const float limit = c.rolling_resistance * c.rolling_radius * normalSum;
    const float mag = length(total);
    if (mag > limit) {
        const V3 fix = total * (limit / mag) - total;
        a.w = a.w - ma.iiw * fix;
        b.w = b.w + mb.iiw * fix;
        total = total + fix;
    }
    st3(c.rolling_impulse, total);
}

// Loads the points of a contact into row form and back.
TF_HD void pointsOf(const tfs_contact& c, RowPoint* p) {
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        if (i >= c.count) { break; }
        const tfs_point& q = c.points[i];
        p[i].ra = ld3(q.anchor_a);
        p[i].rb = ld3(q.anchor_b);
        p[i].adj = q.adjusted_separation;
        p[i].rv = q.relative_velocity;
        p[i].ni = q.normal_impulse;
        p[i].ti = ld3(q.tangent_impulse);
        p[i].nimax = q.max_normal_impulse;
    }
}
TF_HD void storeImpulses(tfs_contact& c, const RowPoint* p) {
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        if (i >= c.count) { break; }
        c.points[i].normal_impulse = p[i].ni;
        st3(c.points[i].tangent_impulse, p[i].ti);
        c.points[i].max_normal_impulse = p[i].nimax;
    }
}
TF_HD Row rowOf(const tfs_contact& c, bool moveA, bool moveB) {
    return Row{ (uint32_t)c.count | (moveA ? kMoveA : 0u) | (moveB ? kMoveB : 0u), ld3(c.normal), ld3(c.tangent1),
                ld3(c.tangent2), c.friction, c.restitution };
}

// solveContacts, one contact of the rest: src/solver.cpp:360-421.
template <bool useBias>
TF_HD void solveRest(tfs_contact& c, const BodyView& va, const BodyView& vb, const Mass& ma, const Mass& mb, Motion& a,
                     Motion& b, const Ctx& ctx, const PlanarAxes* axes) {
    if (c.planar) {
        solvePlanar<useBias>(c, va, vb, ma, mb, a, b, ctx, axes);
        return;
    }
    RowPoint p[4];
    pointsOf(c, p);
    const Row row = rowOf(c, !anchoredView(va), !anchoredView(vb));
    const MadeAxes ax{ ma, mb, row, p };
    if (useBias) {
        rows<kBias>(row, p, ax, a, b, ctx);
    } else {
        rows<kRelax>(row, p, ax, a, b, ctx);
    }
    storeImpulses(c, p);
    if (c.rolling_resistance > 0.0f) { resistTurning(c, ma, mb, a, b); }
}

// applyRestitution, one contact that is not laned: src/solver.cpp:423-460.
TF_HD void restituteRest(tfs_contact& c, const BodyView& va, const BodyView& vb, const Mass& ma, const Mass& mb,
                         Motion& a, Motion& b, const Ctx& ctx, const PlanarAxes* axes) {
    if (c.restitution == 0.0f || c.laned) { return; }
    const V3 n = ld3(c.normal);
    if (c.planar) {
        PlanarState st(a, b, ma, mb, ld3(c.plane));
        TF_POINTS(i, c.count) {
            tfs_point& p = c.points[i];
            if (p.relative_velocity > -ctx.restitutionThreshold || p.max_normal_impulse == 0.0f) { continue; }
            const PlanarAxes& x = axes[i];
            const float vn = st.speed(n, x.normalA, x.normalB);
            float impulse = -x.normalMass * (vn + c.restitution * p.relative_velocity);
            const float total = maxq(p.normal_impulse + impulse, 0.0f);
            impulse = total - p.normal_impulse;
            p.normal_impulse = total;
            p.max_normal_impulse = maxq(p.max_normal_impulse, impulse);
            st.push(n, x.normalA, x.normalB, impulse);
        }
        st.store(a, b, ld3(c.plane), !anchoredView(va), !anchoredView(vb));
        return;
    }
    TF_POINTS(i, c.count) {
        tfs_point& p = c.points[i];
        if (p.relative_velocity > -ctx.restitutionThreshold || p.max_normal_impulse == 0.0f) { continue; }
        const Axis x = makeAxis(ma, mb, ld3(p.anchor_a), ld3(p.anchor_b), n);
        const float vn = speedAlong(a, b, n, x);
        float impulse = -x.mass * (vn + c.restitution * p.relative_velocity);
        const float total = maxq(p.normal_impulse + impulse, 0.0f);
        impulse = total - p.normal_impulse;
        p.normal_impulse = total;
        p.max_normal_impulse = maxq(p.max_normal_impulse, impulse);
        push(a, b, x, impulse, true, true);
    }
}

// nearestPoint: src/world.cpp:122-133.
TF_HD int nearestPoint(const tfs_contact& old, V3 localA) {
    int best = -1;
    float bestSq = 3.40282347e+38f;
    TF_POINTS(k, old.count) {
        const V3 d = ld3(old.points[k].local_a) - localA;
        const float dsq = dot(d, d);
        if (dsq < bestSq) {
            best = k;
            bestSq = dsq;
        }
    }
    return best;
}

// carryImpulses: src/world.cpp:102-120.
TF_HD void carryImpulses(const tfs_contact& old, tfs_contact& c, float matchSq) {
    int from[4] = { -1, -1, -1, -1 };
    TF_POINTS(i, c.count) {
        float bestSq = matchSq;
        TF_POINTS(k, old.count) {
            const V3 d = ld3(old.points[k].local_a) - ld3(c.points[i].local_a);
            const float dsq = dot(d, d);
            if (dsq <= bestSq) {
                from[i] = k;
                bestSq = dsq;
            }
        }
    }
    const V3 n = ld3(c.normal);
    TF_POINTS(i, c.count) {
        if (from[i] < 0) { continue; }
        // The number of new points that take their impulse from the same old point.
        int claims = 0;
        TF_POINTS(j, c.count) { claims += from[j] == from[i] ? 1 : 0; }
        const tfs_point& o = old.points[from[i]];
        const float share = 1.0f / (float)claims;
        c.points[i].normal_impulse = o.normal_impulse * share;
        const V3 t = ld3(o.tangent_impulse);
        st3(c.points[i].tangent_impulse, (t - n * dot(n, t)) * share);
    }
    c.rolling_impulse[0] = old.rolling_impulse[0];
    c.rolling_impulse[1] = old.rolling_impulse[1];
    c.rolling_impulse[2] = old.rolling_impulse[2];
}

// ---- the contact stream: the solver half of updateContacts (src/world.cpp:880-1025) ----------------

// Quat::unrotate: include/twofold/math.h:176. Quaternion dot: math.h:189.
TF_HD V3 unrotate(Q4 q, V3 v) { return rotate(conjugate(q), v); }
TF_HD float qdot(Q4 a, Q4 b) { return a.x * b.x + a.y * b.y + a.z * b.z + a.w * b.w; }
// perpendicular: include/twofold/math.h:70-73.
TF_HD V3 perpendicular(V3 n) {
    if (fabs_(n.x) > 0.57735f) { return normalize(vec(n.y, -n.x, 0.0f)); }
    return normalize(vec(0.0f, n.z, -n.y));
}
// Transform::apply / applyInverse: include/twofold/math.h:236-237, for a body's pose.
TF_HD V3 applyPose(const tfs_body& b, V3 v) { return ld3(b.p) + rotate(ld4(b.q), v); }
TF_HD V3 applyInversePose(const tfs_body& b, V3 v) { return unrotate(ld4(b.q), v - ld3(b.p)); }

// lift: src/world.cpp:537-540. Points further apart than the speculative distance anchor on each surface.
TF_HD V3 lift(const tfs_body& a, const tfs_body& b, V3 normal, float separation, float speculative) {
    if (isPlanarLocks(a.locks) || isPlanarLocks(b.locks)) { return vec(0.0f, 0.0f, 0.0f); }
    return separation > speculative ? normal * (0.5f * separation) : vec(0.0f, 0.0f, 0.0f);
}

// stillFits: src/world.cpp:1002-1008. `reach` = contactReuseDistance x lengthUnit, `reuseCos` =
// cos(contactReuseAngle / 2) as the host's cosf gives it.
TF_HD bool stillFits(const tfs_contact_ext& c, const tfs_body& a, const tfs_body& b, float reach, float reuseCos) {
    if (!c.reusable) { return false; }
    const V3 off = unrotate(ld4(a.q), ld3(b.com) - ld3(a.com)) - ld3(c.base_offset);
    if (dot(off, off) > reach * reach) { return false; }
    const Q4 turn = conjugate(ld4(a.q)) * ld4(b.q);
    return fabs_(qdot(turn, ld4(c.base_turn))) >= reuseCos;
}

// remember: src/world.cpp:995-1000.
TF_HD void remember(const tfs_body& a, const tfs_body& b, const tfs_contact& c, tfs_contact_ext& e) {
    const Q4 qa = ld4(a.q);
    st3(e.local_normal, unrotate(qa, ld3(c.normal)));
    st3(e.base_offset, unrotate(qa, ld3(b.com) - ld3(a.com)));
    st4(e.base_turn, conjugate(qa) * ld4(b.q));
    e.reusable = 1;
}

// measureAgain: src/world.cpp:1011-1025. Kept points stay where they were on each body; the normal
// turns with body A.
TF_HD void measureAgain(const tfs_body& a, const tfs_body& b, tfs_contact& c, const tfs_contact_ext& e, float speculative) {
    const V3 n = rotate(ld4(a.q), ld3(e.local_normal));
    const V3 t1 = perpendicular(n);
    st3(c.normal, n);
    st3(c.tangent1, t1);
    st3(c.tangent2, cross(n, t1));
    TF_POINTS(i, c.count) {
        tfs_point& p = c.points[i];
        const V3 pa = applyPose(a, ld3(e.surface_a[i])), pb = applyPose(b, ld3(e.surface_b[i]));
        const V3 mid = (pa + pb) * 0.5f;
        p.separation = dot(pb - pa, n);
        const V3 up = lift(a, b, n, p.separation, speculative);
        st3(p.anchor_a, mid - up - ld3(a.com));
        st3(p.anchor_b, mid + up - ld3(b.com));
        const V3 t = ld3(p.tangent_impulse);
        st3(p.tangent_impulse, t - n * dot(n, t));
    }
}

// A new contact from a manifold: src/world.cpp:919-944 (the fields the solver path reads), then remember.
TF_HD void contactFromManifold(const tfs_body& a, const tfs_body& b, const tfs_manifold& m, tfs_contact& c,
                               tfs_contact_ext& e, float speculative) {
    const V3 n = ld3(m.normal);
    const V3 t1 = perpendicular(n);
    st3(c.normal, n);
    st3(c.tangent1, t1);
    st3(c.tangent2, cross(n, t1));
    c.count = m.count;
    TF_POINTS(i, m.count) {
        tfs_point& p = c.points[i];
        const V3 pos = ld3(m.position[i]);
        const float sep = m.separation[i];
        const V3 half = n * (0.5f * sep);
        st3(e.surface_a[i], applyInversePose(a, pos - half));
        st3(e.surface_b[i], applyInversePose(b, pos + half));
        st3(p.local_a, applyInversePose(a, pos));
        p.separation = sep;
        const V3 up = lift(a, b, n, sep, speculative);
        st3(p.anchor_a, pos - up - ld3(a.com));
        st3(p.anchor_b, pos + up - ld3(b.com));
    }
    remember(a, b, c, e);
}

}  // namespace tf
