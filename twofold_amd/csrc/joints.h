// joints.h - the reference's joint rows (src/joint.cpp), restated for the device: the point, fixed,
// distance and slider joints and the hinge without limits. Each function performs the reference's float
// operations in its order (with -ffp-contract=off and IEEE division and square root, the same bits):
// pseudoInverse 22-70, arm 74, flipped 86-96, pointResponse 98-102, anchoredTurn 104-118, lockedSlide
// 120-131, reachable 134-138, pushPoint / pushTurn 140-150, solvePoint / solveTurn / solveEqual /
// solveLimit / solveMotor 155-204, turnError 206-210, across 219, warmStartJoint 289-335, solveJoint
// 337-432, shift / movePoint / moveTurn / moveAlong 434-459, settleJoint 461-542, scaleToKeptSpeed
// 555-575, addReaction 603-627; makeAxis / speedAlong / push from src/solver.cpp:5-28 and makeSoft
// 255-262 (stages.h). A limited hinge and the swing-twist joint read angles through atan2 and asin, whose bits are
// the host C library's; they are refused before any step.
#pragma once

#include "../../include/twofold_solver.h"
#include "fmath.h"
#include "stages.h"

namespace tfj {

using tf::M3;
using tf::makeSoft;
using tf::Q4;
using tf::Soft;
using tf::V3;
using tf::cross;
using tf::dot;
using tf::length;
using tf::maxf;
using tf::minf;
using tf::mul;
using tf::quat;
using tf::rotate;
using tf::vec;

// A joint's end as the rows read it: a body of the world, or the world body (static, at the origin,
// zero inverse mass, with velocities of its own that the pushes touch).
struct End {
    V3 v, w;
    V3 com, lc, p;
    Q4 q;
    M3 iMW, iIW;
    float invMass;
    uint32_t locks;
    uint8_t type, awake;
};

struct Axis {
    V3 armA, armB, linearA, angularA, linearB, angularB;
    float mass;
};

struct JCtx {
    float h, invH;
    Soft jointSoft;
};

TF_HD M3 identity() { return M3{ { vec(1.0f, 0.0f, 0.0f), vec(0.0f, 1.0f, 0.0f), vec(0.0f, 0.0f, 1.0f) } }; }
TF_HD V3 neg(V3 a) { return vec(-a.x, -a.y, -a.z); }
TF_HD V3 vdiv(V3 a, float s) { return vec(a.x / s, a.y / s, a.z / s); }
TF_HD float comp(V3 v, int k) { return k == 0 ? v.x : k == 1 ? v.y : v.z; }

TF_HD M3 skew(V3 v) { return M3{ { vec(0.0f, -v.z, v.y), vec(v.z, 0.0f, -v.x), vec(-v.y, v.x, 0.0f) } }; }
TF_HD float trace(const M3& m) { return m.r[0].x + m.r[1].y + m.r[2].z; }

// Inverse within the directions m acts in: the plain inverse when m is well conditioned, else Jacobi
// rotations of its symmetric part and the reciprocals of the eigenvalues that are not negligible.
TF_HD M3 pseudoInverse(const M3& m) {
    const float tr = trace(m);
    if (tr <= 0.0f) { return tf::mzero(); }
    const float third = tr / 3.0f;
    if (dot(m.r[0], cross(m.r[1], m.r[2])) > 1e-3f * third * third * third) { return tf::inverse(m); }
    float a[3][3], v[3][3];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            a[i][k] = 0.5f * (comp(m.r[i], k) + comp(m.r[k], i));
            v[i][k] = i == k ? 1.0f : 0.0f;
        }
    }
    for (int sweep = 0; sweep < 8; ++sweep) {
        if (a[0][1] * a[0][1] + a[0][2] * a[0][2] + a[1][2] * a[1][2] <= 1e-14f * tr * tr) { break; }
#pragma unroll
        for (int p = 0; p < 2; ++p) {
#pragma unroll
            for (int q = p + 1; q < 3; ++q) {
                if (a[p][q] == 0.0f) { continue; }
                const float theta = (a[q][q] - a[p][p]) / (2.0f * a[p][q]);
                const float t = (theta >= 0.0f ? 1.0f : -1.0f) / (fabsf(theta) + sqrtf(theta * theta + 1.0f));
                const float c = 1.0f / sqrtf(t * t + 1.0f), s = t * c;
#pragma unroll
                for (int k = 0; k < 3; ++k) {
                    const float kp = a[k][p], kq = a[k][q];
                    a[k][p] = c * kp - s * kq;
                    a[k][q] = s * kp + c * kq;
                }
#pragma unroll
                for (int k = 0; k < 3; ++k) {
                    const float pk = a[p][k], qk = a[q][k];
                    a[p][k] = c * pk - s * qk;
                    a[q][k] = s * pk + c * qk;
                }
#pragma unroll
                for (int k = 0; k < 3; ++k) {
                    const float kp = v[k][p], kq = v[k][q];
                    v[k][p] = c * kp - s * kq;
                    v[k][q] = s * kp + c * kq;
                }
            }
        }
    }
    const float top = maxf(a[0][0], maxf(a[1][1], a[2][2]));
    M3 out = tf::mzero();
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        if (a[i][i] <= 1e-6f * top) { continue; }
        const V3 e = vec(v[0][i], v[1][i], v[2][i]);
        out = out + tf::outer(e, e) * (1.0f / a[i][i]);
    }
    return out;
}

TF_HD V3 arm(const End& b, V3 local) { return rotate(b.q, local - b.lc); }

// src/solver.cpp:5-28
TF_HD Axis makeAxis(const End& a, const End& b, V3 ra, V3 rb, V3 d) {
    Axis x;
    x.armA = cross(ra, d);
    x.armB = cross(rb, d);
    x.linearA = a.iMW * d;
    x.angularA = a.iIW * x.armA;
    x.linearB = b.iMW * d;
    x.angularB = b.iIW * x.armB;
    const float k = dot(d, x.linearA) + dot(d, x.linearB) + dot(x.armA, x.angularA) + dot(x.armB, x.angularB);
    x.mass = k > 1e-6f * (a.invMass + b.invMass) ? 1.0f / k : 0.0f;
    return x;
}
TF_HD float speedAlong(const End& a, const End& b, V3 d, const Axis& x) {
    return dot(d, b.v - a.v) + dot(x.armB, b.w) - dot(x.armA, a.w);
}
TF_HD void push(End& a, End& b, const Axis& x, float impulse) {
    a.v = a.v - x.linearA * impulse;
    a.w = a.w - x.angularA * impulse;
    b.v = b.v + x.linearB * impulse;
    b.w = b.w + x.angularB * impulse;
}

TF_HD Axis flipped(const Axis& x) {
    Axis f = x;
    f.armA = neg(x.armA);
    f.armB = neg(x.armB);
    f.linearA = neg(x.linearA);
    f.angularA = neg(x.angularA);
    f.linearB = neg(x.linearB);
    f.angularB = neg(x.angularB);
    return f;
}

TF_HD M3 pointResponse(const End& b, V3 r) {
    const M3 s = skew(r);
    return b.iMW - mul(mul(s, b.iIW), s);
}

TF_HD Axis anchoredTurn(const End& a, const End& b, V3 ra, V3 rb, V3 axis, V3& pointImpulse) {
    const V3 drift = cross(b.iIW * axis, rb) + cross(a.iIW * axis, ra);
    pointImpulse = pseudoInverse(pointResponse(a, ra) + pointResponse(b, rb)) * neg(drift);
    Axis x;
    x.armA = axis + cross(ra, pointImpulse);
    x.armB = axis + cross(rb, pointImpulse);
    x.linearA = a.iMW * pointImpulse;
    x.angularA = a.iIW * x.armA;
    x.linearB = b.iMW * pointImpulse;
    x.angularB = b.iIW * x.armB;
    const float k = dot(pointImpulse, x.linearA + x.linearB) + dot(x.armA, x.angularA) + dot(x.armB, x.angularB);
    x.mass = k > 1e-6f * (trace(a.iIW) + trace(b.iIW)) ? 1.0f / k : 0.0f;
    return x;
}

TF_HD Axis lockedSlide(const End& a, const End& b, V3 ra, V3 rb, V3 axis) {
    Axis x = makeAxis(a, b, ra, rb, axis);
    const V3 twist = pseudoInverse(a.iIW + b.iIW) * neg(x.angularA + x.angularB);
    x.armA = x.armA + twist;
    x.armB = x.armB + twist;
    x.angularA = a.iIW * x.armA;
    x.angularB = b.iIW * x.armB;
    const float k = dot(axis, x.linearA + x.linearB) + dot(x.armA, x.angularA) + dot(x.armB, x.angularB);
    x.mass = k > 1e-6f * (a.invMass + b.invMass) ? 1.0f / k : 0.0f;
    return x;
}

TF_HD float reachable(const End& a, const End& b, const Axis& x) {
    const float invMass = a.invMass + b.invMass;
    if ((!a.locks && !b.locks) || x.mass <= 0.0f || invMass <= 0.0f) { return 1.0f; }
    return minf(1.0f, 1.0f / (x.mass * invMass));
}

TF_HD void pushPoint(End& a, End& b, V3 ra, V3 rb, V3 impulse) {
    a.v = a.v - a.iMW * impulse;
    a.w = a.w - a.iIW * cross(ra, impulse);
    b.v = b.v + b.iMW * impulse;
    b.w = b.w + b.iIW * cross(rb, impulse);
}
TF_HD void pushTurn(End& a, End& b, V3 impulse) {
    a.w = a.w - a.iIW * impulse;
    b.w = b.w + b.iIW * impulse;
}

TF_HD Soft softFor(const JCtx& ctx, bool useBias) { return useBias ? ctx.jointSoft : Soft{ 0.0f, 1.0f, 0.0f }; }

TF_HD void solvePoint(End& a, End& b, V3 ra, V3 rb, const M3& proj, V3 error, V3& total, const Soft& soft) {
    const M3 k = mul(mul(proj, pointResponse(a, ra) + pointResponse(b, rb)), proj);
    const V3 speed = proj * (b.v + cross(b.w, rb) - a.v - cross(a.w, ra));
    const V3 old = proj * total;
    const V3 impulse = (pseudoInverse(k) * (speed + error * soft.biasRate)) * -soft.massScale - old * soft.impulseScale;
    total = old + impulse;
    pushPoint(a, b, ra, rb, impulse);
}

TF_HD void solveTurn(End& a, End& b, const M3& proj, V3 error, V3& total, const Soft& soft) {
    const M3 k = mul(mul(proj, a.iIW + b.iIW), proj);
    const V3 speed = proj * (b.w - a.w);
    const V3 old = proj * total;
    const V3 impulse = (pseudoInverse(k) * (speed + error * soft.biasRate)) * -soft.massScale - old * soft.impulseScale;
    total = old + impulse;
    pushTurn(a, b, impulse);
}

TF_HD void solveEqual(End& a, End& b, const Axis& x, V3 d, float error, float& total, const Soft& soft) {
    const float impulse = -x.mass * soft.massScale * (speedAlong(a, b, d, x) + soft.biasRate * error) - soft.impulseScale * total;
    total += impulse;
    push(a, b, x, impulse);
}

TF_HD void solveLimit(End& a, End& b, const Axis& x, V3 d, float error, float& total, const JCtx& ctx, bool useBias) {
    float bias = 0.0f, massScale = 1.0f, impulseScale = 0.0f;
    if (error > 0.0f) {
        bias = error * ctx.invH;
    } else if (useBias) {
        bias = ctx.jointSoft.biasRate * error;
        massScale = ctx.jointSoft.massScale;
        impulseScale = ctx.jointSoft.impulseScale;
    }
    float impulse = -x.mass * massScale * (speedAlong(a, b, d, x) + bias) - impulseScale * total;
    const float next = maxf(total + impulse, 0.0f);
    impulse = next - total;
    total = next;
    push(a, b, x, impulse);
}

TF_HD void solveMotor(End& a, End& b, const Axis& x, V3 d, float speed, float maxImpulse, float& total) {
    const float old = total;
    total = minf(maxf(old - x.mass * (speedAlong(a, b, d, x) - speed), -maxImpulse), maxImpulse);
    push(a, b, x, total - old);
}

TF_HD V3 turnError(const tfs_joint& j, const End& a, const End& b) {
    Q4 e = tf::operator*(b.q, tf::conjugate(tf::operator*(a.q, tf::ld4(j.rest_turn))));
    if (e.w < 0.0f) { e = quat(-e.x, -e.y, -e.z, -e.w); }
    return vec(e.x, e.y, e.z) * 2.0f;
}

TF_HD M3 across(V3 axis) { return identity() - tf::outer(axis, axis); }

TF_HD void warmStartJoint(End& a, End& b, const tfs_joint& j) {
    const V3 ra = arm(a, tf::ld3(j.local_a)), rb = arm(b, tf::ld3(j.local_b));
    const V3 d = (b.com + rb) - (a.com + ra);
    float axial = j.axial_impulse + j.motor_impulse + j.lower_impulse - j.upper_impulse;
    if (j.max_friction > 0.0f) { axial += j.friction_impulse; }
    const V3 linear = tf::ld3(j.linear_impulse), angular = tf::ld3(j.angular_impulse);
    switch (j.type) {
        case TFS_JOINT_POINT: pushPoint(a, b, ra, rb, linear); break;
        case TFS_JOINT_FIXED:
            pushPoint(a, b, ra, rb, linear);
            pushTurn(a, b, angular);
            break;
        case TFS_JOINT_HINGE: {
            const V3 axis = rotate(a.q, tf::ld3(j.axis_a));
            V3 along;
            pushPoint(a, b, ra, rb, linear);
            pushTurn(a, b, across(axis) * angular);
            push(a, b, anchoredTurn(a, b, ra, rb, axis, along), axial);
            break;
        }
        case TFS_JOINT_SLIDER: {
            const V3 axis = rotate(a.q, tf::ld3(j.axis_a));
            pushPoint(a, b, ra + d, rb, across(axis) * linear);
            pushTurn(a, b, angular);
            push(a, b, lockedSlide(a, b, ra + d, rb, axis), axial);
            break;
        }
        case TFS_JOINT_DISTANCE: {
            const float len = length(d);
            if (len > 0.0f) { push(a, b, makeAxis(a, b, ra, rb, vdiv(d, len)), axial); }
            break;
        }
        default: break;
    }
}

TF_HD void solveJoint(End& a, End& b, tfs_joint& j, const JCtx& ctx, bool useBias) {
    const Soft soft = softFor(ctx, useBias);
    const V3 ra = arm(a, tf::ld3(j.local_a)), rb = arm(b, tf::ld3(j.local_b));
    const V3 d = (b.com + rb) - (a.com + ra);
    V3 linear = tf::ld3(j.linear_impulse), angular = tf::ld3(j.angular_impulse);
    switch (j.type) {
        case TFS_JOINT_POINT: solvePoint(a, b, ra, rb, identity(), d, linear, soft); break;
        case TFS_JOINT_FIXED:
            solveTurn(a, b, identity(), turnError(j, a, b), angular, soft);
            solvePoint(a, b, ra, rb, identity(), d, linear, soft);
            break;
        case TFS_JOINT_HINGE: {
            const V3 axis = rotate(a.q, tf::ld3(j.axis_a));
            V3 along;
            const Axis turn = anchoredTurn(a, b, ra, rb, axis, along);
            tf::st3(j.row_force[0], along);
            tf::st3(j.row_torque[0], axis);
            if (j.max_friction > 0.0f) { solveMotor(a, b, turn, along, 0.0f, j.max_friction * ctx.h, j.friction_impulse); }
            if (j.max_motor > 0.0f) { solveMotor(a, b, turn, along, j.motor_speed, j.max_motor * ctx.h, j.motor_impulse); }
            solveTurn(a, b, across(axis), cross(axis, rotate(b.q, tf::ld3(j.axis_b))), angular, soft);
            solvePoint(a, b, ra, rb, identity(), d, linear, soft);
            break;
        }
        case TFS_JOINT_SLIDER: {
            const V3 axis = rotate(a.q, tf::ld3(j.axis_a));
            const Axis slide = lockedSlide(a, b, ra + d, rb, axis);
            tf::st3(j.row_force[0], axis);
            tf::st3(j.row_torque[0], slide.armB - cross(rb, axis));
            if (j.max_friction > 0.0f) { solveMotor(a, b, slide, axis, 0.0f, j.max_friction * ctx.h, j.friction_impulse); }
            if (j.max_motor > 0.0f) { solveMotor(a, b, slide, axis, j.motor_speed, j.max_motor * ctx.h, j.motor_impulse); }
            if (j.limited) {
                const float t = dot(d, axis);
                solveLimit(a, b, slide, axis, t - j.lower, j.lower_impulse, ctx, useBias);
                solveLimit(a, b, flipped(slide), neg(axis), j.upper - t, j.upper_impulse, ctx, useBias);
            }
            solveTurn(a, b, identity(), turnError(j, a, b), angular, soft);
            const M3 side = across(axis);
            solvePoint(a, b, ra + d, rb, side, side * d, linear, soft);
            break;
        }
        case TFS_JOINT_DISTANCE: {
            const float len = length(d);
            if (len <= 0.0f) { break; }
            const V3 u = vdiv(d, len);
            const Axis x = makeAxis(a, b, ra, rb, u);
            const float reach = reachable(a, b, x);
            tf::st3(j.row_force[0], u);
            if (j.limited ? j.lower >= j.upper : j.hertz == 0.0f) {
                solveEqual(a, b, x, u, (len - (j.limited ? j.lower : j.length)) * reach, j.axial_impulse, soft);
                break;
            }
            if (j.hertz > 0.0f) {
                const Soft spring = makeSoft(j.hertz, j.damping_ratio, ctx.h);
                solveEqual(a, b, x, u, (len - j.length) * reach, j.axial_impulse, spring);
            }
            if (j.limited) {
                const float lower = len - j.lower, upper = j.upper - len;
                solveLimit(a, b, x, u, maxf(lower, lower * reach), j.lower_impulse, ctx, useBias);
                solveLimit(a, b, flipped(x), neg(u), maxf(upper, upper * reach), j.upper_impulse, ctx, useBias);
            }
            break;
        }
        default: break;
    }
    tf::st3(j.linear_impulse, linear);
    tf::st3(j.angular_impulse, angular);
}

TF_HD void addReaction(tfs_joint& j) {
    V3 force = tf::ld3(j.linear_impulse), torque = vec(0.0f, 0.0f, 0.0f);
    const float axial = j.axial_impulse + j.motor_impulse + j.lower_impulse - j.upper_impulse + j.friction_impulse;
    switch (j.type) {
        case TFS_JOINT_FIXED: torque = tf::ld3(j.angular_impulse); break;
        case TFS_JOINT_HINGE:
        case TFS_JOINT_SLIDER:
            force = force + tf::ld3(j.row_force[0]) * axial;
            torque = tf::ld3(j.angular_impulse) + tf::ld3(j.row_torque[0]) * axial;
            break;
        case TFS_JOINT_DISTANCE: force = tf::ld3(j.row_force[0]) * axial; break;
        default: break;
    }
    tf::st3(j.force, tf::ld3(j.force) + force);
    tf::st3(j.torque, tf::ld3(j.torque) + torque);
}

// ---- settling: position steps after the solve (src/joint.cpp:434-542) --------------------------------

TF_HD void shift(End& b, V3 move, V3 turn) {
    if (b.type != TFS_DYNAMIC || !b.awake) { return; }
    b.com = b.com + move;
    b.q = tf::integrate(b.q, turn, 1.0f);
    b.p = b.com - rotate(b.q, b.lc);
}

TF_HD void movePoint(End& a, End& b, V3 ra, V3 rb, const M3& proj, V3 error) {
    const M3 k = mul(mul(proj, pointResponse(a, ra) + pointResponse(b, rb)), proj);
    const V3 p = (pseudoInverse(k) * (proj * error)) * -1.0f;
    shift(a, neg(a.iMW * p), neg(a.iIW * cross(ra, p)));
    shift(b, b.iMW * p, b.iIW * cross(rb, p));
}

TF_HD void moveTurn(End& a, End& b, const M3& proj, V3 error) {
    const M3 k = mul(mul(proj, a.iIW + b.iIW), proj);
    const V3 l = (pseudoInverse(k) * (proj * error)) * -1.0f;
    shift(a, vec(0.0f, 0.0f, 0.0f), neg(a.iIW * l));
    shift(b, vec(0.0f, 0.0f, 0.0f), b.iIW * l);
}

TF_HD void moveAlong(End& a, End& b, const Axis& x, float error) {
    const float l = -x.mass * error;
    shift(a, neg(x.linearA * l), neg(x.angularA * l));
    shift(b, x.linearB * l, x.angularB * l);
}

TF_HD void settleJoint(End& a, End& b, const tfs_joint& j) {
    const V3 la = tf::ld3(j.local_a), lb = tf::ld3(j.local_b);
    V3 ra = arm(a, la), rb = arm(b, lb);
    V3 d = (b.com + rb) - (a.com + ra);
    auto measure = [&]() {
        ra = arm(a, la);
        rb = arm(b, lb);
        d = (b.com + rb) - (a.com + ra);
    };
    switch (j.type) {
        case TFS_JOINT_POINT: movePoint(a, b, ra, rb, identity(), d); break;
        case TFS_JOINT_FIXED:
            moveTurn(a, b, identity(), turnError(j, a, b));
            measure();
            movePoint(a, b, ra, rb, identity(), d);
            break;
        case TFS_JOINT_HINGE: {
            const V3 axis = rotate(a.q, tf::ld3(j.axis_a));
            moveTurn(a, b, across(axis), cross(axis, rotate(b.q, tf::ld3(j.axis_b))));
            measure();
            movePoint(a, b, ra, rb, identity(), d);
            break;
        }
        case TFS_JOINT_SLIDER: {
            moveTurn(a, b, identity(), turnError(j, a, b));
            measure();
            const V3 axis = rotate(a.q, tf::ld3(j.axis_a));
            if (j.limited) {
                const float t = dot(d, axis);
                const Axis slide = lockedSlide(a, b, ra + d, rb, axis);
                if (t < j.lower) { moveAlong(a, b, slide, t - j.lower); }
                if (t > j.upper) { moveAlong(a, b, slide, t - j.upper); }
                measure();
            }
            const M3 side = across(axis);
            movePoint(a, b, ra + d, rb, side, d);
            break;
        }
        case TFS_JOINT_DISTANCE: {
            const float len = length(d);
            if (len <= 0.0f) { break; }
            const V3 u = vdiv(d, len);
            const Axis x = makeAxis(a, b, ra, rb, u);
            const float reach = reachable(a, b, x);
            if (j.limited ? j.lower >= j.upper : j.hertz == 0.0f) {
                moveAlong(a, b, x, (len - (j.limited ? j.lower : j.length)) * reach);
            } else if (j.limited && len < j.lower) {
                moveAlong(a, b, x, (len - j.lower) * reach);
            } else if (j.limited && len > j.upper) {
                moveAlong(a, b, x, (len - j.upper) * reach);
            }
            break;
        }
        default: break;
    }
}

// scaleToKeptSpeed's scaling of one joint's impulses (src/joint.cpp:555-569).
TF_HD void scaleImpulses(tfs_joint& j, float f) {
    tf::st3(j.linear_impulse, tf::ld3(j.linear_impulse) * f);
    tf::st3(j.angular_impulse, tf::ld3(j.angular_impulse) * f);
    tf::st3(j.friction_turn, tf::ld3(j.friction_turn) * f);
    j.axial_impulse *= f;
    j.motor_impulse *= f;
    j.lower_impulse *= f;
    j.upper_impulse *= f;
    j.friction_impulse *= f;
    j.swing_impulse *= f;
}

// The kinds the device solves: every joint but a limited hinge and the swing-twist joint.
TF_HD bool supported(const tfs_joint& j) {
    return j.type <= TFS_JOINT_DISTANCE && !(j.type == TFS_JOINT_HINGE && j.limited);
}

}  // namespace tfj
