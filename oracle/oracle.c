/*
 * oracle.c - CPU restatement of the reference's soft-step contact solve path, in plain C.
 *
 * TEST INFRASTRUCTURE ONLY (the checker): tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may call it; the product never does.
 *
 * It follows the reference one function at a time, with the same floating-point operations in the
 * same order (build with -ffp-contract=off, as the reference's CMakeLists.txt:32-37 does), so that
 * its results are bit-identical to the reference's. The reference solves coloured contacts four
 * lanes at a time (src/lanes.cpp); lanes never share a moving body and every lane takes exactly the
 * scalar operations (src/lanes.h:448-449), so this restatement solves them one contact at a time in
 * row order. Point axes (src/solver.cpp:5-17) are rebuilt where they are used from the same body
 * matrices and anchors the reference builds them from, which gives the same bits.
 */
#include "oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

typedef struct { float x, y, z; } V3;
typedef struct { float x, y, z, w; } Q4;
typedef struct { V3 r[3]; } M3;

/* ---- include/twofold/math.h and src/internal.h:28-45 --------------------------------------- */

/* maxf / minf: src/internal.h:28-45 (std::fmax / std::fmin with NaN and signed zeros). */
float tfo_maxf(float a, float b) {
    if (a == b || a != a || b != b) {
        if (a != a) { return b; }
        if (b != b) { return a; }
        return a == 0.0f ? a + b : a;
    }
    return a > b ? a : b;
}

float tfo_minf(float a, float b) {
    if (a == b || a != a || b != b) {
        if (a != a) { return b; }
        if (b != b) { return a; }
        return a == 0.0f ? -(-a + -b) : a;
    }
    return a < b ? a : b;
}
#define maxf tfo_maxf
#define minf tfo_minf

static V3 v3(const float* p) { V3 v = { p[0], p[1], p[2] }; return v; }
static void st3(float* p, V3 v) { p[0] = v.x; p[1] = v.y; p[2] = v.z; }
static Q4 q4(const float* p) { Q4 q = { p[0], p[1], p[2], p[3] }; return q; }
static void stq(float* p, Q4 q) { p[0] = q.x; p[1] = q.y; p[2] = q.z; p[3] = q.w; }
static M3 m3(const float* p) { M3 m; m.r[0] = v3(p); m.r[1] = v3(p + 3); m.r[2] = v3(p + 6); return m; }
static void stm(float* p, M3 m) { st3(p, m.r[0]); st3(p + 3, m.r[1]); st3(p + 6, m.r[2]); }

static V3 vec(float x, float y, float z) { V3 v = { x, y, z }; return v; }
static V3 add(V3 a, V3 b) { return vec(a.x + b.x, a.y + b.y, a.z + b.z); }   /* math.h:40 */
static V3 sub(V3 a, V3 b) { return vec(a.x - b.x, a.y - b.y, a.z - b.z); }   /* math.h:41 */
static V3 mul(V3 a, float s) { return vec(a.x * s, a.y * s, a.z * s); }     /* math.h:43 */
static float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }  /* math.h:53 */
static V3 cross(V3 a, V3 b) {                                              /* math.h:54-56 */
    return vec(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
static float length(V3 v) { return sqrtf(dot(v, v)); }                     /* math.h:58 */
static V3 normalize(V3 v) {                                               /* math.h:59-62 */
    const float l = length(v);
    return l > 0 ? mul(v, 1.0f / l) : vec(0, 0, 0);
}

static M3 mzero(void) { M3 m; m.r[0] = m.r[1] = m.r[2] = vec(0, 0, 0); return m; }  /* math.h:79-83 */
static M3 outer(V3 a, V3 b) { M3 m; m.r[0] = mul(b, a.x); m.r[1] = mul(b, a.y); m.r[2] = mul(b, a.z); return m; } /* math.h:89-93 */
static V3 column(M3 m, int i) {                                            /* math.h:95 */
    return i == 0 ? vec(m.r[0].x, m.r[1].x, m.r[2].x) : i == 1 ? vec(m.r[0].y, m.r[1].y, m.r[2].y)
                                                              : vec(m.r[0].z, m.r[1].z, m.r[2].z);
}
static V3 mv(M3 m, V3 v) { return vec(dot(m.r[0], v), dot(m.r[1], v), dot(m.r[2], v)); }  /* math.h:97 */
static M3 mm(M3 a, M3 b) {                                                 /* math.h:98-102 */
    M3 m;
    for (int i = 0; i < 3; ++i) { m.r[i] = vec(dot(a.r[i], column(b, 0)), dot(a.r[i], column(b, 1)), dot(a.r[i], column(b, 2))); }
    return m;
}
static M3 madd(M3 a, M3 b) { M3 m; for (int i = 0; i < 3; ++i) { m.r[i] = add(a.r[i], b.r[i]); } return m; }  /* 103-107 */
static M3 msub(M3 a, M3 b) { M3 m; for (int i = 0; i < 3; ++i) { m.r[i] = sub(a.r[i], b.r[i]); } return m; }  /* 108-112 */
static M3 mscale(M3 a, float s) { M3 m; for (int i = 0; i < 3; ++i) { m.r[i] = mul(a.r[i], s); } return m; }  /* 113-117 */
static M3 transpose(M3 a) { M3 t; for (int i = 0; i < 3; ++i) { t.r[i] = column(a, i); } return t; }        /* 120-124 */
static M3 inverse(M3 m) {                                                  /* math.h:126-138 */
    const V3 c0 = cross(m.r[1], m.r[2]);
    const V3 c1 = cross(m.r[2], m.r[0]);
    const V3 c2 = cross(m.r[0], m.r[1]);
    float det = dot(m.r[0], c0);
    if (det == 0.0f) { return mzero(); }
    det = 1.0f / det;
    M3 inv;
    inv.r[0] = mul(vec(c0.x, c1.x, c2.x), det);
    inv.r[1] = mul(vec(c0.y, c1.y, c2.y), det);
    inv.r[2] = mul(vec(c0.z, c1.z, c2.z), det);
    return inv;
}

static Q4 quat(float x, float y, float z, float w) { Q4 q = { x, y, z, w }; return q; }
static Q4 qmul(Q4 a, Q4 b) {                                              /* math.h:162-167 */
    return quat(a.w * b.x + a.x * b.w + a.y * b.z - a.z * b.y,
                a.w * b.y - a.x * b.z + a.y * b.w + a.z * b.x,
                a.w * b.z + a.x * b.y - a.y * b.x + a.z * b.w,
                a.w * b.w - a.x * b.x - a.y * b.y - a.z * b.z);
}
static Q4 conjugate(Q4 q) { return quat(-q.x, -q.y, -q.z, q.w); }         /* math.h:169 */
static V3 rotate(Q4 q, V3 v) {                                            /* math.h:171-175 */
    const V3 u = vec(q.x, q.y, q.z);
    const V3 t = mul(cross(u, v), 2.0f);
    return add(add(v, mul(t, q.w)), cross(u, t));
}
static M3 toMat3(Q4 q) {                                                  /* math.h:178-187 */
    const float xx = q.x * q.x, yy = q.y * q.y, zz = q.z * q.z;
    const float xy = q.x * q.y, xz = q.x * q.z, yz = q.y * q.z;
    const float wx = q.w * q.x, wy = q.w * q.y, wz = q.w * q.z;
    M3 m;
    m.r[0] = vec(1 - 2 * (yy + zz), 2 * (xy - wz), 2 * (xz + wy));
    m.r[1] = vec(2 * (xy + wz), 1 - 2 * (xx + zz), 2 * (yz - wx));
    m.r[2] = vec(2 * (xz - wy), 2 * (yz + wx), 1 - 2 * (xx + yy));
    return m;
}
static Q4 qnormalize(Q4 q) {                                              /* math.h:190-194 */
    const float l = sqrtf(q.x * q.x + q.y * q.y + q.z * q.z + q.w * q.w);
    return l > 0 ? quat(q.x / l, q.y / l, q.z / l, q.w / l) : quat(0, 0, 0, 1);
}
static Q4 integrate(Q4 q, V3 w, float h) {                                /* math.h:227-230 */
    const Q4 d = qmul(quat(w.x, w.y, w.z, 0.0f), q);
    return qnormalize(quat(q.x + 0.5f * h * d.x, q.y + 0.5f * h * d.y, q.z + 0.5f * h * d.z, q.w + 0.5f * h * d.w));
}

/* ---- bodies ---------------------------------------------------------------------------------- */

static int dynamic(const tfs_body* b) { return b->type == TFS_DYNAMIC; }
static int moving(const tfs_body* b) { return b->type == TFS_DYNAMIC && b->awake; }   /* lanes.cpp:9 */
static int anchored(const tfs_body* b) { return !moving(b); }                          /* solver.cpp:54 */
static int isPlanar(const tfs_body* b) { return (b->locks & TFS_LOCK_PLANAR) == TFS_LOCK_PLANAR; } /* internal.h:323 */
static V3 planeNormal(const tfs_body* b) { return rotate(q4(b->lock_frame), vec(0, 0, 1)); }       /* internal.h:324 */

/* invertSubspace: src/world.cpp:44-60. */
static M3 invertSubspace(M3 m, const V3* axes, int k) {
    if (k == 0) { return mzero(); }
    if (k == 3) { return inverse(m); }
    if (k == 1) {
        const float d = dot(axes[0], mv(m, axes[0]));
        return d > 0.0f ? mscale(outer(axes[0], axes[0]), 1.0f / d) : mzero();
    }
    const float a = dot(axes[0], mv(m, axes[0]));
    const float b = dot(axes[0], mv(m, axes[1]));
    const float c = dot(axes[1], mv(m, axes[0]));
    const float d = dot(axes[1], mv(m, axes[1]));
    float det = a * d - b * c;
    if (det == 0.0f) { return mzero(); }
    det = 1.0f / det;
    return madd(msub(msub(mscale(outer(axes[0], axes[0]), d * det), mscale(outer(axes[0], axes[1]), b * det)),
                     mscale(outer(axes[1], axes[0]), c * det)),
                mscale(outer(axes[1], axes[1]), a * det));
}

/* updateInverseMass: src/world.cpp:62-92. */
void tfo_update_inverse_mass(tfs_body* b) {
    const M3 lf = toMat3(q4(b->lock_frame));
    V3 ang[3];
    int na = 0;
    M3 lp = mzero(), ap = mzero();
    for (int i = 0; i < 3; ++i) {
        const V3 axis = column(lf, i);
        if (!(b->locks & (TFS_LOCK_TX << i))) { lp = madd(lp, outer(axis, axis)); }
        if (!(b->locks & (TFS_LOCK_RX << i))) {
            ang[na++] = axis;
            ap = madd(ap, outer(axis, axis));
        }
    }
    stm(b->linear_proj, lp);
    stm(b->angular_proj, ap);
    if (b->type != TFS_DYNAMIC || !b->awake) {
        stm(b->inv_mass_world, mzero());
        stm(b->inv_inertia_world, mzero());
        b->plane_inv_mass = 0.0f;
        b->plane_inv_inertia = 0.0f;
        return;
    }
    stm(b->inv_mass_world, mscale(lp, b->inv_mass));
    const M3 r = toMat3(q4(b->q));
    const M3 iw = mm(mm(r, m3(b->inertia)), transpose(r));
    const M3 inv = invertSubspace(iw, ang, na);
    stm(b->inv_inertia_world, inv);
    if (isPlanar(b)) {
        const V3 n = planeNormal(b);
        b->plane_inv_mass = b->inv_mass;
        b->plane_inv_inertia = dot(n, mv(inv, n));
    }
}

/* limitSpeed: src/solver.cpp:30-41. */
static void limitSpeed(tfs_body* b, float maxSpeed, float maxSpin) {
    if (b->type != TFS_DYNAMIC || !b->awake) { return; }
    const float speed = length(v3(b->v)), spin = length(v3(b->w));
    if (speed > maxSpeed) {
        st3(b->v, mul(v3(b->v), maxSpeed / speed));
        b->kept = minf(b->kept, maxSpeed / speed);
    }
    if (spin > maxSpin) {
        st3(b->w, mul(v3(b->w), maxSpin / spin));
        b->kept = minf(b->kept, maxSpin / spin);
    }
}

/* makeSoft: src/solver.cpp:255-262. */
void tfo_make_soft(float hertz, float zeta, float h, float out[3]) {
    if (hertz == 0.0f) { out[0] = 0.0f; out[1] = 1.0f; out[2] = 0.0f; return; }
    const float omega = 2.0f * 3.14159265358979f * hertz;
    const float a1 = 2.0f * zeta + h * omega;
    const float a2 = h * omega * a1;
    const float a3 = 1.0f / (1.0f + a2);
    out[0] = omega / a1;
    out[1] = a2 * a3;
    out[2] = a3;
}

typedef struct {
    float h, invH;
    float contactSoft[3], staticSoft[3];
    float maxPushout, restitutionThreshold;
    float maxSpeed, maxSpin;
    V3 gravity;
} Ctx;

/* ---- axes ------------------------------------------------------------------------------------ */

typedef struct { V3 armA, armB, linearA, angularA, linearB, angularB; float mass; } Axis;

/* makeAxis: src/solver.cpp:5-17; the lane form makeDirection (src/lanes.cpp:38-49) is the same. */
static Axis makeAxis(const tfs_body* a, const tfs_body* b, V3 ra, V3 rb, V3 d) {
    Axis x;
    x.armA = cross(ra, d);
    x.armB = cross(rb, d);
    x.linearA = mv(m3(a->inv_mass_world), d);
    x.angularA = mv(m3(a->inv_inertia_world), x.armA);
    x.linearB = mv(m3(b->inv_mass_world), d);
    x.angularB = mv(m3(b->inv_inertia_world), x.armB);
    const float k = dot(d, x.linearA) + dot(d, x.linearB) + dot(x.armA, x.angularA) + dot(x.armB, x.angularB);
    x.mass = k > 1e-6f * (a->inv_mass + b->inv_mass) ? 1.0f / k : 0.0f;
    return x;
}

/* speedAlong: src/solver.cpp:19-21. */
static float speedAlong(const tfs_body* a, const tfs_body* b, V3 d, const Axis* x) {
    return dot(d, sub(v3(b->v), v3(a->v))) + dot(x->armB, v3(b->w)) - dot(x->armA, v3(a->w));
}

/* push (solver.cpp:23-28) when both moves are set, pushMoving (212-221) and the masked lane push
 * (lanes.cpp:161-166) otherwise. */
static void pushBodies(tfs_body* a, tfs_body* b, const Axis* x, float impulse, int moveA, int moveB) {
    if (moveA) {
        st3(a->v, sub(v3(a->v), mul(x->linearA, impulse)));
        st3(a->w, sub(v3(a->w), mul(x->angularA, impulse)));
    }
    if (moveB) {
        st3(b->v, add(v3(b->v), mul(x->linearB, impulse)));
        st3(b->w, add(v3(b->w), mul(x->angularB, impulse)));
    }
}

/* turned: src/solver.cpp:57-59 (lanes.cpp:168). */
static V3 turned(const tfs_body* b, V3 r) {
    return b->dq[0] == 0.0f && b->dq[1] == 0.0f && b->dq[2] == 0.0f ? r : rotate(q4(b->dq), r);
}

/* ---- planar ---------------------------------------------------------------------------------- */

/* planarContact: src/solver.cpp:66-79. n is written as the reference writes it. */
static int planarContact(const tfs_body* a, const tfs_body* b, V3 normal, float* n) {
    const tfs_body* first = NULL;
    const tfs_body* xs[2] = { a, b };
    V3 nv = v3(n);
    for (int k = 0; k < 2; ++k) {
        const tfs_body* x = xs[k];
        if (x->type == TFS_STATIC) { continue; }
        if (x->type == TFS_KINEMATIC || !isPlanar(x)) { return 0; }
        if (!first) {
            first = x;
            nv = planeNormal(x);
            st3(n, nv);
        } else if (fabsf(dot(nv, planeNormal(x))) < 0.9999f) {
            return 0;
        }
    }
    return first && fabsf(dot(normal, nv)) < 1e-3f;
}

/* planarMass: src/solver.cpp:81-84. */
static float planarMass(const tfs_body* a, const tfs_body* b, float armA, float armB) {
    const float k = a->plane_inv_mass + b->plane_inv_mass + a->plane_inv_inertia * armA * armA + b->plane_inv_inertia * armB * armB;
    return k > 1e-6f * (a->inv_mass + b->inv_mass) ? 1.0f / k : 0.0f;
}

typedef struct { float normalA, normalB, normalMass, tangentA, tangentB, tangentMass; } PlanarAxes;

/* makePlanarAxes: src/solver.cpp:86-98, for point i. */
static PlanarAxes planarAxes(const tfs_body* a, const tfs_body* b, const tfs_contact* c, int i) {
    const V3 n = v3(c->plane);
    const tfs_point* p = &c->points[i];
    PlanarAxes x;
    x.normalA = dot(n, cross(v3(p->anchor_a), v3(c->normal)));
    x.normalB = dot(n, cross(v3(p->anchor_b), v3(c->normal)));
    x.normalMass = planarMass(a, b, x.normalA, x.normalB);
    x.tangentA = dot(n, cross(v3(p->anchor_a), v3(c->tangent1)));
    x.tangentB = dot(n, cross(v3(p->anchor_b), v3(c->tangent1)));
    x.tangentMass = planarMass(a, b, x.tangentA, x.tangentB);
    return x;
}

/* PlanarState: src/solver.cpp:101-129. */
typedef struct { V3 va, vb; float wa, wb; float ma, mb, ia, ib; } PlanarState;

static PlanarState planarState(const tfs_body* a, const tfs_body* b, V3 n) {
    PlanarState s;
    s.va = v3(a->v);
    s.vb = v3(b->v);
    s.wa = dot(v3(a->w), n);
    s.wb = dot(v3(b->w), n);
    s.ma = a->plane_inv_mass;
    s.mb = b->plane_inv_mass;
    s.ia = a->plane_inv_inertia;
    s.ib = b->plane_inv_inertia;
    return s;
}
static float psSpeed(const PlanarState* s, V3 d, float armA, float armB) {
    return dot(d, sub(s->vb, s->va)) + armB * s->wb - armA * s->wa;
}
static void psPush(PlanarState* s, V3 d, float armA, float armB, float impulse) {
    s->va = sub(s->va, mul(d, s->ma * impulse));
    s->wa -= s->ia * armA * impulse;
    s->vb = add(s->vb, mul(d, s->mb * impulse));
    s->wb += s->ib * armB * impulse;
}
static void psStore(const PlanarState* s, tfs_body* a, tfs_body* b, V3 n, int moveA, int moveB) {
    if (moveA) {
        st3(a->v, s->va);
        st3(a->w, mul(n, s->wa));
    }
    if (moveB) {
        st3(b->v, s->vb);
        st3(b->w, mul(n, s->wb));
    }
}

/* planarTurn: src/solver.cpp:132-135. */
static void planarTurn(const tfs_body* b, V3 n, float* c, float* s) {
    const Q4 q = q4(b->dq);
    *c = q.w * q.w - (q.x * q.x + q.y * q.y + q.z * q.z);
    *s = 2.0f * q.w * dot(n, vec(q.x, q.y, q.z));
}

static void softOf(const Ctx* ctx, int moveA, int moveB, float* soft) {
    const float* s = (!moveA || !moveB) ? ctx->staticSoft : ctx->contactSoft;
    soft[0] = s[0];
    soft[1] = s[1];
    soft[2] = s[2];
}

/* solvePlanar: src/solver.cpp:137-209. */
static void solvePlanar(tfs_body* a, tfs_body* b, tfs_contact* c, const Ctx* ctx, int useBias) {
    const int moveA = !anchored(a), moveB = !anchored(b);
    float soft[3];
    softOf(ctx, moveA, moveB, soft);
    const V3 normal = v3(c->normal), tangent1 = v3(c->tangent1), plane = v3(c->plane);
    const float shift = dot(sub(v3(b->dp), v3(a->dp)), normal);
    float tax, tay, tbx, tby;
    planarTurn(a, plane, &tax, &tay);
    planarTurn(b, plane, &tbx, &tby);
    PlanarState st = planarState(a, b, plane);
    int changed = 0;
    PlanarAxes axes[4];
    for (int i = 0; i < c->count; ++i) { axes[i] = planarAxes(a, b, c, i); }

    for (int i = 0; i < c->count; ++i) {
        tfs_point* p = &c->points[i];
        const PlanarAxes* x = &axes[i];
        const float reachA = tax * x->tangentA + tay * x->normalA;
        const float reachB = tbx * x->tangentB + tby * x->normalB;
        const float sep = shift + reachB - reachA + p->adjusted_separation;
        float bias = 0.0f, massScale = 1.0f, impulseScale = 0.0f;
        if (sep > 0.0f) {
            bias = sep * ctx->invH;
        } else if (useBias) {
            bias = maxf(soft[0] * sep, -ctx->maxPushout);
            massScale = soft[1];
            impulseScale = soft[2];
        }
        const float vn = psSpeed(&st, normal, x->normalA, x->normalB);
        float impulse = -x->normalMass * massScale * (vn + bias) - impulseScale * p->normal_impulse;
        const float total = maxf(p->normal_impulse + impulse, 0.0f);
        impulse = total - p->normal_impulse;
        p->normal_impulse = total;
        p->max_normal_impulse = maxf(p->max_normal_impulse, impulse);
        if (impulse != 0.0f) {
            psPush(&st, normal, x->normalA, x->normalB, impulse);
            changed = 1;
        }
    }

    for (int j = 0; j < c->count; ++j) {
        const int i = useBias ? j : c->count - 1 - j;
        tfs_point* p = &c->points[i];
        const PlanarAxes* x = &axes[i];
        const float old = dot(v3(p->tangent_impulse), tangent1);
        if (p->normal_impulse == 0.0f && old == 0.0f) { continue; }
        changed = 1;
        const float limit = c->friction * p->normal_impulse;
        const float t = minf(maxf(old - x->tangentMass * psSpeed(&st, tangent1, x->tangentA, x->tangentB), -limit), limit);
        st3(p->tangent_impulse, mul(tangent1, t));
        psPush(&st, tangent1, x->tangentA, x->tangentB, t - old);
    }

    if (c->rolling_resistance > 0.0f) {
        changed = 1;
        float total = dot(v3(c->rolling_impulse), plane);
        const float spin = st.wb - st.wa;
        const float k = st.ia + st.ib;
        if (spin != 0.0f && k > 0.0f) {
            const float lambda = -spin / k;
            total += lambda;
            st.wa -= st.ia * lambda;
            st.wb += st.ib * lambda;
        }
        float normalSum = 0.0f;
        for (int i = 0; i < c->count; ++i) { normalSum += c->points[i].normal_impulse; }
        const float limit = c->rolling_resistance * c->rolling_radius * normalSum;
        const float held = minf(maxf(total, -limit), limit);
        st.wa -= st.ia * (held - total);
        st.wb += st.ib * (held - total);
        st3(c->rolling_impulse, mul(plane, held));
    }
    if (changed) { psStore(&st, a, b, plane, moveA, moveB); }
}

/* resistTurning: src/solver.cpp:225-251. */
static void resistTurning(tfs_body* a, tfs_body* b, tfs_contact* c) {
    const V3 spin = sub(v3(b->w), v3(a->w));
    const float speed = length(spin);
    V3 total = v3(c->rolling_impulse);
    if (speed > 0.0f) {
        const V3 d = mul(spin, 1.0f / speed);
        const V3 ia = mv(m3(a->inv_inertia_world), d), ib = mv(m3(b->inv_inertia_world), d);
        const float k = dot(d, ia) + dot(d, ib);
        if (k > 0.0f) {
            const float lambda = -speed / k;
            total = add(total, mul(d, lambda));
            st3(a->w, sub(v3(a->w), mul(ia, lambda)));
            st3(b->w, add(v3(b->w), mul(ib, lambda)));
        }
    }
    float normalSum = 0.0f;
    for (int i = 0; i < c->count; ++i) { normalSum += c->points[i].normal_impulse; }
    const float limit = c->rolling_resistance * c->rolling_radius * normalSum;
    const float mag = length(total);
    if (mag > limit) {
        const V3 fix = sub(mul(total, limit / mag), total);
        st3(a->w, sub(v3(a->w), mv(m3(a->inv_inertia_world), fix)));
        st3(b->w, add(v3(b->w), mv(m3(b->inv_inertia_world), fix)));
        total = add(total, fix);
    }
    st3(c->rolling_impulse, total);
}

/* ---- per-contact stages --------------------------------------------------------------------- */

/* prepareContact: src/solver.cpp:264-291. */
static void prepareContact(const tfs_body* a, const tfs_body* b, tfs_contact* c) {
    c->planar = (uint8_t)planarContact(a, b, v3(c->normal), c->plane);
    if (c->planar) {
        const V3 plane = v3(c->plane);
        const V3 n = normalize(sub(v3(c->normal), mul(plane, dot(v3(c->normal), plane))));
        st3(c->normal, n);
        st3(c->tangent1, cross(plane, n));
        st3(c->tangent2, plane);
    }
    const PlanarState st = planarState(a, b, v3(c->plane));
    const V3 normal = v3(c->normal);
    for (int i = 0; i < c->count; ++i) {
        tfs_point* p = &c->points[i];
        const V3 ra = v3(p->anchor_a), rb = v3(p->anchor_b);
        p->adjusted_separation = p->separation - dot(sub(rb, ra), normal);
        if (c->planar) {
            const PlanarAxes x = planarAxes(a, b, c, i);
            p->relative_velocity = psSpeed(&st, normal, x.normalA, x.normalB);
        } else {
            p->relative_velocity = dot(normal, sub(v3(b->v), v3(a->v))) + dot(cross(rb, normal), v3(b->w)) -
                                   dot(cross(ra, normal), v3(a->w));
        }
        p->max_normal_impulse = 0.0f;
    }
    if (c->rolling_resistance > 0.0f && c->count > 0) {
        float ha = 0.0f, hb = 0.0f;
        for (int i = 0; i < c->count; ++i) {
            ha += fabsf(dot(v3(c->points[i].anchor_a), normal));
            hb += fabsf(dot(v3(c->points[i].anchor_b), normal));
        }
        if (a->type != TFS_DYNAMIC) { ha = 0.0f; }
        if (b->type != TFS_DYNAMIC) { hb = 0.0f; }
        c->rolling_radius = maxf(ha, hb) / (float)c->count;
    }
}

/* rebaseContact: src/solver.cpp:293-306 (the axes it rebuilds are rebuilt here where used). */
static void rebaseContact(const tfs_body* a, const tfs_body* b, tfs_contact* c) {
    const float shift = dot(sub(v3(b->dp), v3(a->dp)), v3(c->normal));
    for (int i = 0; i < c->count; ++i) {
        tfs_point* p = &c->points[i];
        st3(p->anchor_a, rotate(q4(a->dq), v3(p->anchor_a)));
        st3(p->anchor_b, rotate(q4(b->dq), v3(p->anchor_b)));
        p->adjusted_separation += shift;
    }
}

/* warmStartLanes (src/lanes.cpp:195-217), one lane. */
static void warmStartLaned(tfs_body* a, tfs_body* b, const tfs_contact* c) {
    const int moveA = moving(a), moveB = moving(b);
    const V3 n = v3(c->normal), t1 = v3(c->tangent1), t2 = v3(c->tangent2);
    for (int i = 0; i < c->count; ++i) {
        const tfs_point* p = &c->points[i];
        const V3 ra = v3(p->anchor_a), rb = v3(p->anchor_b);
        const V3 held = v3(p->tangent_impulse);
        const Axis xn = makeAxis(a, b, ra, rb, n);
        const Axis x1 = makeAxis(a, b, ra, rb, t1);
        const Axis x2 = makeAxis(a, b, ra, rb, t2);
        pushBodies(a, b, &xn, p->normal_impulse, moveA, moveB);
        pushBodies(a, b, &x1, dot(held, t1), moveA, moveB);
        pushBodies(a, b, &x2, dot(held, t2), moveA, moveB);
    }
}

/* warmStart (src/solver.cpp:319-358), one contact that is not laned. */
static void warmStartContact(tfs_body* a, tfs_body* b, const tfs_contact* c) {
    const V3 n = v3(c->normal), t1 = v3(c->tangent1), t2 = v3(c->tangent2);
    if (c->planar) {
        int held = c->rolling_resistance > 0.0f;
        for (int i = 0; i < c->count; ++i) {
            const V3 ti = v3(c->points[i].tangent_impulse);
            held = held || c->points[i].normal_impulse != 0.0f || dot(ti, ti) != 0.0f;
        }
        if (!held) { return; }
        const V3 plane = v3(c->plane);
        PlanarState st = planarState(a, b, plane);
        for (int i = 0; i < c->count; ++i) {
            const tfs_point* p = &c->points[i];
            const PlanarAxes x = planarAxes(a, b, c, i);
            psPush(&st, n, x.normalA, x.normalB, p->normal_impulse);
            psPush(&st, t1, x.tangentA, x.tangentB, dot(v3(p->tangent_impulse), t1));
        }
        if (c->rolling_resistance > 0.0f) {
            const float r = dot(v3(c->rolling_impulse), plane);
            st.wa -= st.ia * r;
            st.wb += st.ib * r;
        }
        psStore(&st, a, b, plane, !anchored(a), !anchored(b));
        return;
    }
    for (int i = 0; i < c->count; ++i) {
        const tfs_point* p = &c->points[i];
        const V3 ra = v3(p->anchor_a), rb = v3(p->anchor_b);
        const Axis xn = makeAxis(a, b, ra, rb, n);
        const Axis x1 = makeAxis(a, b, ra, rb, t1);
        const Axis x2 = makeAxis(a, b, ra, rb, t2);
        pushBodies(a, b, &xn, p->normal_impulse, 1, 1);
        pushBodies(a, b, &x1, dot(v3(p->tangent_impulse), t1), 1, 1);
        pushBodies(a, b, &x2, dot(v3(p->tangent_impulse), t2), 1, 1);
    }
    if (c->rolling_resistance > 0.0f) {
        st3(a->w, sub(v3(a->w), mv(m3(a->inv_inertia_world), v3(c->rolling_impulse))));
        st3(b->w, add(v3(b->w), mv(m3(b->inv_inertia_world), v3(c->rolling_impulse))));
    }
}

/* The 3D contact solve: solveLanes (src/lanes.cpp:220-315) for one lane, and the 3D part of
 * solveContacts (src/solver.cpp:370-419), which takes the same operations. */
static void solve3d(tfs_body* a, tfs_body* b, tfs_contact* c, const Ctx* ctx, int useBias) {
    const int moveA = moving(a), moveB = moving(b);
    float soft[3];
    softOf(ctx, moveA, moveB, soft);
    const V3 n = v3(c->normal), t1 = v3(c->tangent1), t2 = v3(c->tangent2);
    const V3 dp = sub(v3(b->dp), v3(a->dp));

    for (int i = 0; i < c->count; ++i) {
        tfs_point* p = &c->points[i];
        const V3 ra = v3(p->anchor_a), rb = v3(p->anchor_b);
        const V3 d = sub(add(dp, turned(b, rb)), turned(a, ra));
        const float sep = dot(d, n) + p->adjusted_separation;
        float bias = 0.0f, massScale = 1.0f, impulseScale = 0.0f;
        if (sep > 0.0f) {
            bias = sep * ctx->invH;
        } else if (useBias) {
            bias = maxf(soft[0] * sep, -ctx->maxPushout);
            massScale = soft[1];
            impulseScale = soft[2];
        }
        const Axis x = makeAxis(a, b, ra, rb, n);
        const float vn = speedAlong(a, b, n, &x);
        float impulse = -x.mass * massScale * (vn + bias) - impulseScale * p->normal_impulse;
        const float total = maxf(p->normal_impulse + impulse, 0.0f);
        impulse = total - p->normal_impulse;
        p->normal_impulse = total;
        p->max_normal_impulse = maxf(p->max_normal_impulse, impulse);
        pushBodies(a, b, &x, impulse, moveA, moveB);
    }

    for (int j = 0; j < c->count; ++j) {
        const int i = useBias ? j : c->count - 1 - j;
        tfs_point* p = &c->points[i];
        const V3 ra = v3(p->anchor_a), rb = v3(p->anchor_b);
        const Axis x1 = makeAxis(a, b, ra, rb, t1);
        const Axis x2 = makeAxis(a, b, ra, rb, t2);
        const V3 held = v3(p->tangent_impulse);
        const float old1 = dot(held, t1), old2 = dot(held, t2);
        float f1 = old1 - x1.mass * speedAlong(a, b, t1, &x1);
        float f2 = old2 - x2.mass * speedAlong(a, b, t2, &x2);
        const float limit = c->friction * p->normal_impulse;
        const float mag = sqrtf(f1 * f1 + f2 * f2);
        if (mag > limit) {
            const float scale = mag > 0.0f ? limit / mag : 0.0f;
            f1 *= scale;
            f2 *= scale;
        }
        st3(p->tangent_impulse, add(mul(t1, f1), mul(t2, f2)));
        pushBodies(a, b, &x1, f1 - old1, moveA, moveB);
        pushBodies(a, b, &x2, f2 - old2, moveA, moveB);
    }
}

/* solveContacts: src/solver.cpp:360-421, one contact of the rest. */
static void solveRest(tfs_body* a, tfs_body* b, tfs_contact* c, const Ctx* ctx, int useBias) {
    if (c->planar) {
        solvePlanar(a, b, c, ctx, useBias);
        return;
    }
    solve3d(a, b, c, ctx, useBias);
    if (c->rolling_resistance > 0.0f) { resistTurning(a, b, c); }
}

/* restituteLanes: src/lanes.cpp:318-352, one lane. */
static void restituteLaned(tfs_body* a, tfs_body* b, tfs_contact* c, const Ctx* ctx) {
    const int moveA = moving(a), moveB = moving(b);
    const V3 n = v3(c->normal);
    const float e = c->restitution;
    for (int i = 0; i < c->count; ++i) {
        tfs_point* p = &c->points[i];
        const float rv = p->relative_velocity, held = p->normal_impulse, peak = p->max_normal_impulse;
        const int on = !(e == 0.0f) && !(rv > -ctx->restitutionThreshold) && !(peak == 0.0f);
        const Axis x = makeAxis(a, b, v3(p->anchor_a), v3(p->anchor_b), n);
        const float vn = speedAlong(a, b, n, &x);
        float impulse = -x.mass * (vn + e * rv);
        const float total = maxf(held + impulse, 0.0f);
        impulse = total - held;
        if (on) {
            p->normal_impulse = total;
            p->max_normal_impulse = maxf(peak, impulse);
        }
        pushBodies(a, b, &x, impulse, on && moveA, on && moveB);
    }
}

/* applyRestitution: src/solver.cpp:423-460, one contact that is not laned. */
static void restituteContact(tfs_body* a, tfs_body* b, tfs_contact* c, const Ctx* ctx) {
    if (c->restitution == 0.0f || c->laned) { return; }
    const V3 n = v3(c->normal);
    if (c->planar) {
        PlanarState st = planarState(a, b, v3(c->plane));
        for (int i = 0; i < c->count; ++i) {
            tfs_point* p = &c->points[i];
            if (p->relative_velocity > -ctx->restitutionThreshold || p->max_normal_impulse == 0.0f) { continue; }
            const PlanarAxes x = planarAxes(a, b, c, i);
            const float vn = psSpeed(&st, n, x.normalA, x.normalB);
            float impulse = -x.normalMass * (vn + c->restitution * p->relative_velocity);
            const float total = maxf(p->normal_impulse + impulse, 0.0f);
            impulse = total - p->normal_impulse;
            p->normal_impulse = total;
            p->max_normal_impulse = maxf(p->max_normal_impulse, impulse);
            psPush(&st, n, x.normalA, x.normalB, impulse);
        }
        psStore(&st, a, b, v3(c->plane), !anchored(a), !anchored(b));
        return;
    }
    for (int i = 0; i < c->count; ++i) {
        tfs_point* p = &c->points[i];
        if (p->relative_velocity > -ctx->restitutionThreshold || p->max_normal_impulse == 0.0f) { continue; }
        const Axis x = makeAxis(a, b, v3(p->anchor_a), v3(p->anchor_b), n);
        const float vn = speedAlong(a, b, n, &x);
        float impulse = -x.mass * (vn + c->restitution * p->relative_velocity);
        const float total = maxf(p->normal_impulse + impulse, 0.0f);
        impulse = total - p->normal_impulse;
        p->normal_impulse = total;
        p->max_normal_impulse = maxf(p->max_normal_impulse, impulse);
        pushBodies(a, b, &x, impulse, 1, 1);
    }
}

/* ---- world-level stages --------------------------------------------------------------------- */

typedef struct {
    tfs_body* bodies;
    uint32_t nb;
    tfs_contact* contacts;
    uint32_t nc;
    const tfs_settings* s;
    Ctx ctx;
    tfs_contact_ext* ext;   /* the contacts' reuse data, or NULL */
} World;

/* relativeTravel: src/world.cpp:484-486. */
static float relativeTravel(const tfs_body* a, const tfs_body* b, float dt) {
    return (length(sub(v3(b->v), v3(a->v))) + length(v3(a->w)) * a->extent + length(v3(b->w)) * b->extent) * dt;
}

/* integrateVelocities: src/world.cpp:1141-1151. */
static void integrateVelocities(World* w, float h) {
    const V3 g = w->ctx.gravity;
    for (uint32_t i = 0; i < w->nb; ++i) {
        tfs_body* b = &w->bodies[i];
        if (!b->alive || b->type != TFS_DYNAMIC || !b->awake) { continue; }
        V3 v = add(v3(b->v), mul(add(mul(g, b->gravity_scale), mul(v3(b->force), b->inv_mass)), h));
        v = mv(m3(b->linear_proj), mul(v, 1.0f / (1.0f + h * b->linear_damping)));
        V3 om = add(v3(b->w), mul(mv(m3(b->inv_inertia_world), v3(b->torque)), h));
        om = mv(m3(b->angular_proj), mul(om, 1.0f / (1.0f + h * b->angular_damping)));
        st3(b->v, v);
        st3(b->w, om);
        limitSpeed(b, w->ctx.maxSpeed, w->ctx.maxSpin);
    }
}

/* limitSpeeds: src/world.cpp:1153-1159. */
static void limitSpeeds(World* w) {
    for (uint32_t i = 0; i < w->nb; ++i) {
        if (w->bodies[i].alive) { limitSpeed(&w->bodies[i], w->ctx.maxSpeed, w->ctx.maxSpin); }
    }
}

/* integratePositions: src/world.cpp:1161-1172. */
static void integratePositions(World* w, float h) {
    limitSpeeds(w);
    for (uint32_t i = 0; i < w->nb; ++i) {
        tfs_body* b = &w->bodies[i];
        if (!b->alive || b->type == TFS_STATIC) { continue; }
        if (b->type == TFS_DYNAMIC && !b->awake) { continue; }
        const V3 vh = mul(v3(b->v), h);
        st3(b->com, add(v3(b->com), vh));
        st3(b->dp, add(v3(b->dp), vh));
        const Q4 q = integrate(q4(b->q), v3(b->w), h);
        stq(b->q, q);
        stq(b->dq, qmul(q, conjugate(q4(b->q_base))));
        st3(b->p, sub(v3(b->com), rotate(q, v3(b->local_center))));
    }
}

/* impactContacts: src/world.cpp:1029-1049. */
static uint32_t impactContacts(World* w, const uint32_t* active, uint32_t na, float limit, uint32_t* out) {
    uint8_t* struck = (uint8_t*)calloc(w->nb ? w->nb : 1, 1);
    int any = 0;
    for (uint32_t k = 0; k < na; ++k) {
        const tfs_contact* c = &w->contacts[active[k]];
        for (int i = 0; i < c->count; ++i) {
            if (c->points[i].relative_velocity >= -limit) { continue; }
            struck[c->body_a] = dynamic(&w->bodies[c->body_a]);
            struck[c->body_b] = dynamic(&w->bodies[c->body_b]);
            any = 1;
            break;
        }
    }
    uint32_t n = 0;
    if (any) {
        for (uint32_t k = 0; k < na; ++k) {
            const tfs_contact* c = &w->contacts[active[k]];
            if (struck[c->body_a] || struck[c->body_b]) { out[n++] = active[k]; }
        }
    }
    free(struck);
    return n;
}

/* colorContacts: src/lanes.cpp:356-374. */
void tfo_color(const tfs_body* bodies, uint32_t n_bodies, const tfs_contact* contacts, const uint32_t* list,
               uint32_t n, uint8_t* colors) {
    uint32_t* masks = (uint32_t*)calloc(n_bodies ? n_bodies : 1, sizeof(uint32_t));
    for (uint32_t k = 0; k < n; ++k) {
        const uint32_t ci = list[k];
        const tfs_contact* c = &contacts[ci];
        colors[ci] = TFS_COLORS;
        if (c->planar || c->rolling_resistance > 0.0f) { continue; }
        const int ma = moving(&bodies[c->body_a]), mb = moving(&bodies[c->body_b]);
        const uint32_t taken = (ma ? masks[c->body_a] : 0u) | (mb ? masks[c->body_b] : 0u);
        if (taken == 0xffffffffu) { continue; }
        int bit = 0;
        while (taken & (1u << bit)) { ++bit; }
        if (ma) { masks[c->body_a] |= 1u << bit; }
        if (mb) { masks[c->body_b] |= 1u << bit; }
        colors[ci] = (uint8_t)bit;
    }
    free(masks);
}

/* batchContacts' row order: src/lanes.cpp:376-405. Returns the number of coloured rows. */
static uint32_t rowOrder(const tfs_contact* contacts, const uint8_t* colors, const uint32_t* list, uint32_t n,
                         uint32_t* byColor) {
    uint32_t sizes[TFS_COLORS + 1] = { 0 }, start[TFS_COLORS + 1];
    for (uint32_t k = 0; k < n; ++k) { ++sizes[colors[list[k]]]; }
    uint32_t sum = 0;
    for (int k = 0; k <= TFS_COLORS; ++k) {
        start[k] = sum;
        sum += sizes[k];
    }
    for (int points = 4; points >= 0; --points) {
        for (uint32_t k = 0; k < n; ++k) {
            const uint32_t ci = list[k];
            const uint8_t col = colors[ci];
            if (col == TFS_COLORS ? points == 4 : contacts[ci].count == points) { byColor[start[col]++] = ci; }
        }
    }
    return sum - sizes[TFS_COLORS];
}

/* nearestPoint: src/world.cpp:122-133. */
static int nearestPoint(const tfs_contact* old, V3 localA) {
    int best = -1;
    float bestSq = FLT_MAX;
    for (int k = 0; k < old->count; ++k) {
        const V3 d = sub(v3(old->points[k].local_a), localA);
        const float dsq = dot(d, d);
        if (dsq < bestSq) {
            best = k;
            bestSq = dsq;
        }
    }
    return best;
}

/* carryImpulses: src/world.cpp:102-120. */
static void carryImpulses(const tfs_contact* old, tfs_contact* c, float matchSq) {
    int from[4], claims[4] = { 0, 0, 0, 0 };
    for (int i = 0; i < c->count; ++i) {
        from[i] = -1;
        float bestSq = matchSq;
        for (int k = 0; k < old->count; ++k) {
            const V3 d = sub(v3(old->points[k].local_a), v3(c->points[i].local_a));
            const float dsq = dot(d, d);
            if (dsq <= bestSq) {
                from[i] = k;
                bestSq = dsq;
            }
        }
        if (from[i] >= 0) { ++claims[from[i]]; }
    }
    const V3 n = v3(c->normal);
    for (int i = 0; i < c->count; ++i) {
        if (from[i] < 0) { continue; }
        const tfs_point* o = &old->points[from[i]];
        const float share = 1.0f / (float)claims[from[i]];
        c->points[i].normal_impulse = o->normal_impulse * share;
        const V3 t = v3(o->tangent_impulse);
        st3(c->points[i].tangent_impulse, mul(sub(t, mul(n, dot(n, t))), share));
    }
    memcpy(c->rolling_impulse, old->rolling_impulse, sizeof c->rolling_impulse);
}

/* refreshContacts: src/world.cpp:1057-1139. The narrowphase part is the callback's. */
static int refreshContacts(World* w, const uint32_t* active, uint32_t na, const uint32_t* again, uint32_t nagain,
                           int substep, float remaining, tfs_refresh_fn cb, void* user) {
    uint32_t next = 0;
    for (uint32_t k = 0; k < na; ++k) {
        const uint32_t ci = active[k];
        if (next < nagain && again[next] == ci) {
            ++next;
            continue;
        }
        tfs_contact* c = &w->contacts[ci];
        rebaseContact(&w->bodies[c->body_a], &w->bodies[c->body_b], c);
    }
    for (uint32_t i = 0; i < w->nb; ++i) {
        tfs_body* b = &w->bodies[i];
        memset(b->dp, 0, sizeof b->dp);
        stq(b->dq, quat(0, 0, 0, 1));
        memcpy(b->q_base, b->q, sizeof b->q_base);
    }
    /* the callback's compact view: each contact to collide again and its two bodies */
    tfs_contact* fresh = (tfs_contact*)calloc(nagain ? nagain : 1, sizeof(tfs_contact));
    tfs_contact* pairs = (tfs_contact*)malloc((nagain ? nagain : 1) * sizeof(tfs_contact));
    tfs_body* ends = (tfs_body*)malloc((nagain ? 2 * nagain : 1) * sizeof(tfs_body));
    if (!fresh || !pairs || !ends) {
        free(fresh);
        free(pairs);
        free(ends);
        return TFS_E_ARG;
    }
    for (uint32_t k = 0; k < nagain; ++k) {
        pairs[k] = w->contacts[again[k]];
        ends[2 * k] = w->bodies[pairs[k].body_a];
        ends[2 * k + 1] = w->bodies[pairs[k].body_b];
    }
    const int rc = cb(user, substep, remaining, nagain, again, pairs, ends, fresh);
    free(pairs);
    free(ends);
    if (rc != 0) {
        free(fresh);
        return TFS_E_CALLBACK;
    }
    const float matchSq = w->s->speculative * w->s->speculative;
    for (uint32_t k = 0; k < nagain; ++k) {
        tfs_contact* c = &w->contacts[again[k]];
        const tfs_contact* f = &fresh[k];
        if (f->count <= 0) {
            c->count = 0;
            continue;
        }
        const tfs_contact old = *c;
        const tfs_body* a = &w->bodies[c->body_a];
        const tfs_body* b = &w->bodies[c->body_b];
        if (w->ext) {                                             /* world.cpp:1110, 1117 */
            tfs_contact_ext* e = &w->ext[again[k]];
            e->reusable = 0;
            for (int i = 0; i < (f->count > 4 ? 4 : f->count); ++i) {
                memset(e->surface_a[i], 0, sizeof e->surface_a[i]);
                memset(e->surface_b[i], 0, sizeof e->surface_b[i]);
            }
        }
        memcpy(c->normal, f->normal, sizeof c->normal);
        memcpy(c->tangent1, f->tangent1, sizeof c->tangent1);
        memcpy(c->tangent2, f->tangent2, sizeof c->tangent2);
        c->count = f->count > 4 ? 4 : f->count;
        for (int i = 0; i < c->count; ++i) {
            tfs_point* p = &c->points[i];
            memset(p, 0, sizeof *p);
            memcpy(p->local_a, f->points[i].local_a, sizeof p->local_a);
            p->separation = f->points[i].separation;
            memcpy(p->anchor_a, f->points[i].anchor_a, sizeof p->anchor_a);
            memcpy(p->anchor_b, f->points[i].anchor_b, sizeof p->anchor_b);
        }
        prepareContact(a, b, c);
        for (int i = 0; i < c->count; ++i) {
            tfs_point* p = &c->points[i];
            const int near = nearestPoint(&old, v3(p->local_a));
            if (near < 0) { continue; }
            p->relative_velocity = old.points[near].relative_velocity;
            p->max_normal_impulse = old.points[near].max_normal_impulse;
        }
        carryImpulses(&old, c, matchSq);
    }
    free(fresh);
    return 0;
}

/* updateSleep: src/world.cpp:1174-1247 (no joints on this path). */
static uint32_t findRoot(uint32_t* parent, uint32_t x) {
    while (parent[x] != x) {
        parent[x] = parent[parent[x]];
        x = parent[x];
    }
    return x;
}

static int movingAwake(const tfs_body* b) {                     /* src/world.cpp:94-98 */
    if (b->type == TFS_DYNAMIC) { return b->awake; }
    if (b->type == TFS_KINEMATIC) { return dot(v3(b->v), v3(b->v)) > 0.0f || dot(v3(b->w), v3(b->w)) > 0.0f; }
    return 0;
}

static void updateSleep(World* w, float dt) {
    const float threshold = w->s->sleep_threshold * w->s->length_unit;
    for (uint32_t i = 0; i < w->nb; ++i) {
        tfs_body* b = &w->bodies[i];
        if (!b->alive || b->type != TFS_DYNAMIC || !b->awake) { continue; }
        const float speed = maxf(length(v3(b->v)), length(v3(b->w)) * b->extent);
        if (!b->can_sleep || speed > threshold) {
            b->sleep_time = 0.0f;
        } else {
            b->sleep_time += dt;
        }
    }
    uint32_t* parent = (uint32_t*)malloc(sizeof(uint32_t) * (w->nb ? w->nb : 1));
    float* islandTime = (float*)malloc(sizeof(float) * (w->nb ? w->nb : 1));
    for (uint32_t i = 0; i < w->nb; ++i) {
        parent[i] = i;
        islandTime[i] = FLT_MAX;
    }
    for (uint32_t k = 0; k < w->nc; ++k) {
        const tfs_contact* c = &w->contacts[k];
        if (c->count == 0) { continue; }
        tfs_body* a = &w->bodies[c->body_a];
        tfs_body* b = &w->bodies[c->body_b];
        const int da = a->type == TFS_DYNAMIC && a->awake;
        const int db = b->type == TFS_DYNAMIC && b->awake;
        if (da && db) {
            const uint32_t ra = findRoot(parent, c->body_a), rb = findRoot(parent, c->body_b);
            if (ra != rb) { parent[ra > rb ? ra : rb] = ra < rb ? ra : rb; }
        } else if (da && b->type == TFS_KINEMATIC && movingAwake(b)) {
            a->sleep_time = 0.0f;
        } else if (db && a->type == TFS_KINEMATIC && movingAwake(a)) {
            b->sleep_time = 0.0f;
        }
    }
    for (uint32_t i = 0; i < w->nb; ++i) {
        tfs_body* b = &w->bodies[i];
        if (!b->alive || b->type != TFS_DYNAMIC || !b->awake) { continue; }
        const uint32_t r = findRoot(parent, i);
        b->island = r;
        islandTime[r] = minf(islandTime[r], b->sleep_time);
    }
    for (uint32_t i = 0; i < w->nb; ++i) {
        tfs_body* b = &w->bodies[i];
        if (!b->alive || b->type != TFS_DYNAMIC || !b->awake) { continue; }
        if (islandTime[findRoot(parent, i)] < w->s->time_to_sleep) { continue; }
        b->awake = 0;
        memset(b->v, 0, sizeof b->v);
        memset(b->w, 0, sizeof b->w);
        tfo_update_inverse_mass(b);
    }
    free(parent);
    free(islandTime);
}

/* ---- the step --------------------------------------------------------------------------------- */

static void solveRows(World* w, const uint32_t* rows, uint32_t nLaned, uint32_t nRows, int useBias) {
    for (uint32_t k = 0; k < nRows; ++k) {
        tfs_contact* c = &w->contacts[rows[k]];
        tfs_body* a = &w->bodies[c->body_a];
        tfs_body* b = &w->bodies[c->body_b];
        if (k < nLaned) {
            solve3d(a, b, c, &w->ctx, useBias);       /* solveBatches' lanes, lanes.cpp:420-422 */
        } else {
            solveRest(a, b, c, &w->ctx, useBias);     /* then solveContacts(rest), lanes.cpp:423 */
        }
    }
}

/* ---- the contact stream: updateContacts' solver half, src/world.cpp:880-989 ---------------------- */

static V3 unrotate(Q4 q, V3 v) { return rotate(conjugate(q), v); }                       /* math.h:176 */
static float qdot(Q4 a, Q4 b) { return a.x * b.x + a.y * b.y + a.z * b.z + a.w * b.w; } /* math.h:189 */
static V3 perpendicular(V3 n) {                                                           /* math.h:70-73 */
    if (fabsf(n.x) > 0.57735f) { return normalize(vec(n.y, -n.x, 0.0f)); }
    return normalize(vec(0.0f, n.z, -n.y));
}
static V3 apply(const tfs_body* b, V3 v) { return add(v3(b->p), rotate(q4(b->q), v)); }          /* math.h:236 */
static V3 applyInverse(const tfs_body* b, V3 v) { return unrotate(q4(b->q), sub(v, v3(b->p))); } /* math.h:237 */

static V3 lift(const tfs_body* a, const tfs_body* b, V3 n, float sep, float speculative) {       /* world.cpp:537-540 */
    if (isPlanar(a) || isPlanar(b)) { return vec(0, 0, 0); }
    return sep > speculative ? mul(n, 0.5f * sep) : vec(0, 0, 0);
}

static int stillFits(const tfs_contact_ext* c, const tfs_body* a, const tfs_body* b, const tfs_settings* s) {  /* 1002-1008 */
    if (!c->reusable) { return 0; }
    const float reach = s->contact_reuse_distance * s->length_unit;
    const V3 off = sub(unrotate(q4(a->q), sub(v3(b->com), v3(a->com))), v3(c->base_offset));
    if (dot(off, off) > reach * reach) { return 0; }
    const Q4 turn = qmul(conjugate(q4(a->q)), q4(b->q));
    return fabsf(qdot(turn, q4(c->base_turn))) >= cosf(0.5f * s->contact_reuse_angle);
}

static void measureAgain(const tfs_body* a, const tfs_body* b, tfs_contact* c, const tfs_contact_ext* e,
                         float speculative) {                                                  /* 1011-1025 */
    const V3 n = rotate(q4(a->q), v3(e->local_normal));
    const V3 t1 = perpendicular(n);
    st3(c->normal, n);
    st3(c->tangent1, t1);
    st3(c->tangent2, cross(n, t1));
    for (int i = 0; i < c->count; ++i) {
        tfs_point* p = &c->points[i];
        const V3 pa = apply(a, v3(e->surface_a[i])), pb = apply(b, v3(e->surface_b[i]));
        const V3 mid = mul(add(pa, pb), 0.5f);
        p->separation = dot(sub(pb, pa), n);
        const V3 up = lift(a, b, n, p->separation, speculative);
        st3(p->anchor_a, sub(sub(mid, up), v3(a->com)));
        st3(p->anchor_b, sub(add(mid, up), v3(b->com)));
        const V3 t = v3(p->tangent_impulse);
        st3(p->tangent_impulse, sub(t, mul(n, dot(n, t))));
    }
}

/* One pair's contacts when it is collided again: 919-949 (remember: 995-1000). */
static void fromManifold(const tfs_body* a, const tfs_body* b, const tfs_manifold* m, tfs_contact* c, tfs_contact_ext* e,
                         float speculative) {
    const V3 n = v3(m->normal);
    const V3 t1 = perpendicular(n);
    st3(c->normal, n);
    st3(c->tangent1, t1);
    st3(c->tangent2, cross(n, t1));
    c->count = m->count;
    for (int i = 0; i < m->count; ++i) {
        tfs_point* p = &c->points[i];
        const V3 pos = v3(m->position[i]);
        const float sep = m->separation[i];
        const V3 half = mul(n, 0.5f * sep);
        st3(e->surface_a[i], applyInverse(a, sub(pos, half)));
        st3(e->surface_b[i], applyInverse(b, add(pos, half)));
        st3(p->local_a, applyInverse(a, pos));
        p->separation = sep;
        const V3 up = lift(a, b, n, sep, speculative);
        st3(p->anchor_a, sub(sub(pos, up), v3(a->com)));
        st3(p->anchor_b, sub(add(pos, up), v3(b->com)));
    }
    st3(e->local_normal, unrotate(q4(a->q), n));
    st3(e->base_offset, unrotate(q4(a->q), sub(v3(b->com), v3(a->com))));
    stq(e->base_turn, qmul(conjugate(q4(a->q)), q4(b->q)));
    e->reusable = 1;
}

static int movingAwake(const tfs_body* b);

uint32_t tfo_stream_contacts(const tfs_body* bodies, const tfs_contact* old, const tfs_contact_ext* oldExt, uint32_t nOld,
                             const tfs_pair* pairs, uint32_t np, const tfs_manifold* manifolds, const tfs_settings* s,
                             tfs_contact* out, tfs_contact_ext* outExt) {
    const float speculative = s->speculative;
    const float matchSq = speculative * speculative;
    uint32_t n = 0, o = 0;
    for (uint32_t k = 0; k < np; ++k) {                           /* candidates, key order: 883-959 */
        const tfs_pair* x = &pairs[k];
        const tfs_body* a = &bodies[x->body_a];
        const tfs_body* b = &bodies[x->body_b];
        while (o < nOld && old[o].key < x->key) { ++o; }
        uint32_t end = o;
        while (end < nOld && old[end].key == x->key) { ++end; }
        int keep = end > o;
        for (uint32_t i = o; i < end && keep; ++i) { keep = stillFits(&oldExt[i], a, b, s); }
        const uint8_t touched = end > o ? old[o].touched : 0;
        const uint32_t first = n;
        if (keep) {
            for (uint32_t i = o; i < end; ++i) {
                out[n] = old[i];
                outExt[n] = oldExt[i];
                measureAgain(a, b, &out[n], &outExt[n], speculative);
                ++n;
            }
        } else {
            for (uint32_t j = 0; j < x->count; ++j) {
                tfs_contact* c = &out[n];
                tfs_contact_ext* e = &outExt[n];
                memset(c, 0, sizeof *c);
                memset(e, 0, sizeof *e);
                c->key = x->key;
                c->body_a = x->body_a;
                c->body_b = x->body_b;
                c->touched = touched;
                e->part = j;
                fromManifold(a, b, &manifolds[x->first + j], c, e, speculative);
                const tfs_contact* from = NULL;                      /* closestNormal: 543-549 */
                for (uint32_t i = o; i < end; ++i) {
                    if (old[i].count > 0 && (!from || dot(v3(old[i].normal), v3(c->normal)) > dot(v3(from->normal), v3(c->normal)))) {
                        from = &old[i];
                    }
                }
                if (from) { carryImpulses(from, c, matchSq); }
                ++n;
            }
        }
        for (uint32_t i = first; i < n; ++i) {                    /* 951-955 */
            out[i].friction = x->friction;
            out[i].restitution = x->restitution;
            out[i].rolling_resistance = x->rolling_resistance;
        }
    }
    /* resting pairs stay (972-984), merged in key order (985-987); the two key sets are disjoint */
    const uint32_t collided = n;
    tfs_contact* rest = (tfs_contact*)malloc(sizeof(tfs_contact) * (nOld ? nOld : 1));
    tfs_contact_ext* restExt = (tfs_contact_ext*)malloc(sizeof(tfs_contact_ext) * (nOld ? nOld : 1));
    uint32_t nr = 0;
    for (uint32_t i = 0; i < nOld; ++i) {
        const tfs_body* a = &bodies[old[i].body_a];
        const tfs_body* b = &bodies[old[i].body_b];
        if (movingAwake(a) || movingAwake(b)) { continue; }
        if (a->type != TFS_DYNAMIC && b->type != TFS_DYNAMIC) { continue; }
        rest[nr] = old[i];
        restExt[nr] = oldExt[i];
        ++nr;
    }
    tfs_contact* merged = (tfs_contact*)malloc(sizeof(tfs_contact) * (collided + nr ? collided + nr : 1));
    tfs_contact_ext* mergedExt = (tfs_contact_ext*)malloc(sizeof(tfs_contact_ext) * (collided + nr ? collided + nr : 1));
    uint32_t i = 0, j = 0, m = 0;
    while (i < collided || j < nr) {
        if (j >= nr || (i < collided && out[i].key < rest[j].key)) {
            merged[m] = out[i];
            mergedExt[m++] = outExt[i++];
        } else {
            merged[m] = rest[j];
            mergedExt[m++] = restExt[j++];
        }
    }
    memcpy(out, merged, sizeof(tfs_contact) * m);
    memcpy(outExt, mergedExt, sizeof(tfs_contact_ext) * m);
    free(rest);
    free(restExt);
    free(merged);
    free(mergedExt);
    return m;
}

/* reportTouches: src/world.cpp:1424-1468. The pairs are the runs of contacts with one key; each run's
 * state changes when a point was pushed (largest normal impulse above zero) or lies closer than zero
 * (the linear slop, speculative / 4, for a pair that touched before); the fastest approaching pushed
 * point of a pair solved this step is its hit when faster than the hit threshold. */
static void reportTouches(World* w, tfo_lists* L) {
    const tfs_settings* s = w->s;
    const float slop = 0.25f * s->speculative;
    L->n_touch_events = L->n_hits = L->touching = 0;
    uint32_t last = 0;
    for (uint32_t first = 0; first < w->nc; first = last) {
        const uint64_t key = w->contacts[first].key;
        const int was = w->contacts[first].touched != 0;
        const float apart = was ? slop : 0.0f;
        int touching = 0;
        float fastest = s->hit_threshold * s->length_unit;
        int struck = -1, point = 0;
        for (last = first; last < w->nc && w->contacts[last].key == key; ++last) {
            const tfs_contact* c = &w->contacts[last];
            const tfs_body* a = &w->bodies[c->body_a];
            const tfs_body* b = &w->bodies[c->body_b];
            int overlapping = 0;
            for (int i = 0; i < c->count; ++i) {                 /* Contact::touching, internal.h:497-502 */
                if (c->points[i].separation < 0.0f || c->points[i].normal_impulse > 0.0f) { overlapping = 1; }
            }
            L->touching += (uint32_t)overlapping;
            for (int i = 0; i < c->count; ++i) {
                const tfs_point* p = &c->points[i];
                if (p->max_normal_impulse > 0.0f) {
                    touching = 1;
                    if (-p->relative_velocity > fastest && (moving(a) || moving(b))) {
                        fastest = -p->relative_velocity;
                        struck = (int)last;
                        point = i;
                    }
                } else if (p->separation < apart) {
                    touching = 1;
                }
            }
        }
        if (touching != was) {
            L->touch_events[L->n_touch_events++] = (first << 1) | (uint32_t)touching;
            for (uint32_t k = first; k < last; ++k) { w->contacts[k].touched = (uint8_t)touching; }
        }
        if (struck >= 0) {
            const tfs_contact* c = &w->contacts[struck];
            const tfs_body* a = &w->bodies[c->body_a];
            tfs_hit* h = &L->hits[L->n_hits++];
            h->contact = (uint32_t)struck;
            h->point = (uint32_t)point;
            st3(h->position, add(v3(a->p), rotate(q4(a->q), v3(c->points[point].local_a))));   /* math.h:236 */
            for (int d = 0; d < 3; ++d) { h->normal[d] = c->normal[d]; }
            h->speed = fastest;
        }
    }
}

/* src/world.cpp:1564-1691 from the step state on, reportTouches (1693), then updateSleep (1724). */
int tfo_step(tfs_body* bodies, uint32_t nb, tfs_contact* contacts, uint32_t nc, const tfs_settings* s, float dt,
             tfs_refresh_fn cb, void* user, tfo_lists* L) {
    if (!(dt > 0.0f) || !isfinite(dt)) { return 0; }
    World w;
    w.bodies = bodies;
    w.nb = nb;
    w.contacts = contacts;
    w.nc = nc;
    w.s = s;
    w.ext = L->ext;
    const int n = s->substeps > 1 ? s->substeps : 1;
    const float h = dt / (float)n;

    /* step state: src/world.cpp:1591-1601 */
    for (uint32_t i = 0; i < nb; ++i) {
        tfs_body* b = &bodies[i];
        if (!b->alive) { continue; }
        memcpy(b->q0, b->q, sizeof b->q0);
        memcpy(b->q_base, b->q, sizeof b->q_base);
        memcpy(b->com0, b->com, sizeof b->com0);
        b->expected_travel = (length(v3(b->v)) + length(v3(b->w)) * b->extent) * dt;
        memset(b->dp, 0, sizeof b->dp);
        stq(b->dq, quat(0, 0, 0, 1));
        tfo_update_inverse_mass(b);
    }

    /* lists: src/world.cpp:1604-1621 */
    L->n_active = L->n_fast = L->n_fast_solid = 0;
    for (uint32_t i = 0; i < nc; ++i) {
        const tfs_contact* c = &contacts[i];
        const tfs_body* a = &bodies[c->body_a];
        const tfs_body* b = &bodies[c->body_b];
        if (moving(a) || moving(b)) {
            L->active[L->n_active++] = i;
            if (relativeTravel(a, b, dt) > s->speculative) {
                L->fast[L->n_fast++] = i;
                if (!isPlanar(a) && !isPlanar(b)) { L->fast_solid[L->n_fast_solid++] = i; }
            }
        }
    }

    /* SolverContext: src/world.cpp:1634-1642 */
    Ctx* ctx = &w.ctx;
    ctx->h = h;
    ctx->invH = 1.0f / h;
    const float hertz = minf(s->contact_hertz, 0.25f * ctx->invH);
    tfo_make_soft(hertz, s->contact_damping_ratio, h, ctx->contactSoft);
    tfo_make_soft(2.0f * hertz, s->contact_damping_ratio, h, ctx->staticSoft);
    ctx->maxPushout = s->max_pushout_speed * s->length_unit;
    ctx->restitutionThreshold = s->restitution_threshold * s->length_unit;
    ctx->maxSpeed = s->max_linear_speed * s->length_unit;
    ctx->maxSpin = s->max_angular_speed;
    ctx->gravity = v3(s->gravity);

    /* solve block: src/world.cpp:1648-1684 */
    for (uint32_t k = 0; k < L->n_active; ++k) {
        tfs_contact* c = &contacts[L->active[k]];
        prepareContact(&bodies[c->body_a], &bodies[c->body_b], c);
    }
    L->n_impact = impactContacts(&w, L->active, L->n_active, ctx->restitutionThreshold, L->impact);
    for (uint32_t i = 0; i < nc; ++i) { L->colors[i] = 255; }
    tfo_color(bodies, nb, contacts, L->active, L->n_active, L->colors);
    L->n_colors = 0;
    L->points_active = L->points_impact = 0;
    for (uint32_t k = 0; k < L->n_active; ++k) {
        tfs_contact* c = &contacts[L->active[k]];
        const uint8_t col = L->colors[L->active[k]];
        c->laned = col < TFS_COLORS;
        if (col < TFS_COLORS && col + 1u > L->n_colors) { L->n_colors = col + 1u; }
        L->points_active += (uint32_t)c->count;
    }
    for (uint32_t k = 0; k < L->n_impact; ++k) { L->points_impact += (uint32_t)contacts[L->impact[k]].count; }
    const uint32_t laned = rowOrder(contacts, L->colors, L->active, L->n_active, L->by_color);
    const uint32_t impactLaned = rowOrder(contacts, L->colors, L->impact, L->n_impact, L->impact_by_color);
    L->n_rest = L->n_active - laned;
    L->n_impact_rest = L->n_impact - impactLaned;

    for (int i = 0; i < n; ++i) {
        const uint32_t* again = i == n / 2 ? L->fast : L->fast_solid;
        const uint32_t nagain = i == n / 2 ? L->n_fast : L->n_fast_solid;
        if (i > 0 && nagain > 0) {
            if (!cb) { return TFS_E_CALLBACK; }
            for (uint32_t k = 0; k < nb; ++k) {
                tfs_body* b = &bodies[k];
                if (b->alive && b->type == TFS_DYNAMIC && b->awake) { tfo_update_inverse_mass(b); }
            }
            const int rc = refreshContacts(&w, L->active, L->n_active, again, nagain, i, (float)(n - i) * h, cb, user);
            if (rc != 0) { return rc; }
        }
        integrateVelocities(&w, h);
        for (uint32_t k = 0; k < laned; ++k) {                   /* warmStartBatches */
            tfs_contact* c = &contacts[L->by_color[k]];
            warmStartLaned(&bodies[c->body_a], &bodies[c->body_b], c);
        }
        for (uint32_t k = 0; k < L->n_active; ++k) {             /* warmStart(active) */
            tfs_contact* c = &contacts[L->active[k]];
            if (c->laned) { continue; }
            warmStartContact(&bodies[c->body_a], &bodies[c->body_b], c);
        }
        solveRows(&w, L->by_color, laned, L->n_active, 1);
        if (L->n_impact) { solveRows(&w, L->impact_by_color, impactLaned, L->n_impact, 1); }
        integratePositions(&w, h);
        solveRows(&w, L->by_color, laned, L->n_active, 0);
        if (L->n_impact) { solveRows(&w, L->impact_by_color, impactLaned, L->n_impact, 0); }
    }
    for (uint32_t k = 0; k < laned; ++k) {                       /* restituteBatches */
        tfs_contact* c = &contacts[L->by_color[k]];
        if (c->restitution != 0.0f) { restituteLaned(&bodies[c->body_a], &bodies[c->body_b], c, ctx); }
    }
    for (uint32_t k = 0; k < L->n_active; ++k) {                 /* applyRestitution */
        tfs_contact* c = &contacts[L->active[k]];
        restituteContact(&bodies[c->body_a], &bodies[c->body_b], c, ctx);
    }
    limitSpeeds(&w);

    if (L->touch_events) { reportTouches(&w, L); }
    updateSleep(&w, dt);
    return 0;
}
