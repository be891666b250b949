// narrow.h - the host world's own narrowphase for spheres, boxes and half-spaces (scene.cpp), for the
// pairs that the reference's narrowphase restated in collide3d.h does not take: those the reference
// collides in 2D (every dynamic body locked to one plane) and those with a cylinder or a prism.
//
// The manifolds follow the reference's conventions (normal from shape A to shape B, shapes ordered
// by index, points at the midpoint between the two surfaces, a speculative margin grown by the
// pair's travel: src/world.cpp:484-486, 509-511; anchors lifted onto each surface for separated
// points: src/world.cpp:536-540) but not its bits: box-box uses separating axes and face clipping
// of this file's own, where the reference clips polyhedra (src/collide3d.cpp).
#pragma once

#include <stdint.h>
#include <string.h>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define TFN_HD __host__ __device__ __forceinline__
#define TFN_UNROLL _Pragma("unroll")
#else
#include <math.h>
#define TFN_HD inline
#define TFN_UNROLL
#endif

#include "../../include/twofold_solver.h"

namespace tfn {

struct V {
    float x, y, z;
};
TFN_HD V vec3(float x, float y, float z) { return V{ x, y, z }; }
TFN_HD V operator+(V a, V b) { return vec3(a.x + b.x, a.y + b.y, a.z + b.z); }
TFN_HD V operator-(V a, V b) { return vec3(a.x - b.x, a.y - b.y, a.z - b.z); }
TFN_HD V operator*(V a, float s) { return vec3(a.x * s, a.y * s, a.z * s); }
TFN_HD V operator-(V a) { return vec3(-a.x, -a.y, -a.z); }
TFN_HD float dot(V a, V b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
TFN_HD V cross(V a, V b) { return vec3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x); }
TFN_HD float len(V a) { return sqrtf(dot(a, a)); }
TFN_HD V unit(V a) {
    const float l = len(a);
    return l > 0 ? a * (1.0f / l) : vec3(0, 1, 0);
}
TFN_HD V ld(const float* p) { return vec3(p[0], p[1], p[2]); }
TFN_HD void st(float* p, V v) {
    p[0] = v.x;
    p[1] = v.y;
    p[2] = v.z;
}
// std::min / std::max as the standard defines them (the second operand wins only when it compares
// less / greater), so that both builds pick the same operand for signed zeros and NaNs.
TFN_HD float lo(float a, float b) { return b < a ? b : a; }
TFN_HD float hi(float a, float b) { return a < b ? b : a; }

struct Q {
    float x, y, z, w;
};
TFN_HD V rotate(Q q, V v) {
    const V u = vec3(q.x, q.y, q.z);
    const V t = cross(u, v) * 2.0f;
    return v + t * q.w + cross(u, t);
}
TFN_HD V unrotate(Q q, V v) { return rotate(Q{ -q.x, -q.y, -q.z, q.w }, v); }
TFN_HD Q ldq(const float* p) { return Q{ p[0], p[1], p[2], p[3] }; }

TFN_HD V perpendicular(V n) {
    if (fabsf(n.x) > 0.57735f) { return unit(vec3(n.y, -n.x, 0.0f)); }
    return unit(vec3(0.0f, n.z, -n.y));
}

TFN_HD float comp(V v, int i) { return i == 0 ? v.x : i == 1 ? v.y : v.z; }

enum Kind { kSphere = 0, kBox = 1, kCylinder = 2, kPrism = 3, kHalfSpace = 4 };

struct Point {
    V pos;       // midpoint between the surfaces
    float sep;   // along the normal, negative when overlapping
};

struct Manifold {
    V n;
    int count = 0;
    Point p[4];
    bool reference = false;  // made by the reference's narrowphase (collide3d.h), not by this file
};

struct Pose {
    V c;
    Q q;
    V ax[3];  // the shape's axes in the world
};

TFN_HD Pose poseOf(const float* com, const float* q) {
    Pose p;
    p.c = ld(com);
    p.q = ldq(q);
    p.ax[0] = rotate(p.q, vec3(1, 0, 0));
    p.ax[1] = rotate(p.q, vec3(0, 1, 0));
    p.ax[2] = rotate(p.q, vec3(0, 0, 1));
    return p;
}

// Choices between values, made on the values: a conditional between two named objects would pick an
// address, which puts both objects in memory on the device.
TFN_HD float pickf(bool c, float x, float y) { return c ? x : y; }
TFN_HD V pickv(bool c, V x, V y) { return vec3(pickf(c, x.x, y.x), pickf(c, x.y, y.y), pickf(c, x.z, y.z)); }
TFN_HD Point pickp(bool c, Point x, Point y) { return Point{ pickv(c, x.pos, y.pos), pickf(c, x.sep, y.sep) }; }
TFN_HD Pose pickpose(bool c, const Pose& x, const Pose& y) {
    Pose r;
    r.c = pickv(c, x.c, y.c);
    r.q = Q{ pickf(c, x.q.x, y.q.x), pickf(c, x.q.y, y.q.y), pickf(c, x.q.z, y.q.z), pickf(c, x.q.w, y.q.w) };
    r.ax[0] = pickv(c, x.ax[0], y.ax[0]);
    r.ax[1] = pickv(c, x.ax[1], y.ax[1]);
    r.ax[2] = pickv(c, x.ax[2], y.ax[2]);
    return r;
}
// Element i of three.
TFN_HD V of3(const V (&a)[3], int i) { return pickv(i == 0, a[0], pickv(i == 1, a[1], a[2])); }
TFN_HD float of3(const float (&a)[3], int i) { return pickf(i == 0, a[0], pickf(i == 1, a[1], a[2])); }

// Small arrays stay in registers on the device when every index is a constant of an unrolled loop:
// writes at a running count go through put, and loops run to the array's size with a guard.
template <class T, int N>
TFN_HD void put(T (&a)[N], int i, const T& v) {
    TFN_UNROLL
    for (int j = 0; j < N; ++j) {
        if (j == i) { a[j] = v; }
    }
}

// Keeps at most four points of the first n: the deepest, the farthest from it, and the two farthest
// on either side of the line between them.
template <int N>
TFN_HD void reduce(const Point (&pts)[N], int n, V nrm, Manifold& m) {
    if (n <= 4) {
        m.count = n;
        TFN_UNROLL
        for (int i = 0; i < 4; ++i) {
            if (i < n) { m.p[i] = pts[i]; }
        }
        return;
    }
    int a = 0;
    Point pa = pts[0];
    TFN_UNROLL
    for (int i = 1; i < N; ++i) {
        if (i < n && pts[i].sep < pa.sep) {
            a = i;
            pa = pts[i];
        }
    }
    int b = a == 0 ? 1 : 0;
    Point pb = pickp(a == 0, pts[1], pts[0]);
    float best = -1.0f;
    TFN_UNROLL
    for (int i = 0; i < N; ++i) {
        if (i >= n) { continue; }
        const V d = pts[i].pos - pa.pos;
        if (i != a && dot(d, d) > best) {
            best = dot(d, d);
            b = i;
            pb = pts[i];
        }
    }
    const V ab = pb.pos - pa.pos;
    int c = -1, d = -1;
    Point pc = pa, pd = pa;
    float up = 0.0f, down = 0.0f;
    TFN_UNROLL
    for (int i = 0; i < N; ++i) {
        if (i >= n || i == a || i == b) { continue; }
        const float s = dot(cross(ab, pts[i].pos - pa.pos), nrm);
        if (s > up) {
            up = s;
            c = i;
            pc = pts[i];
        }
        if (s < down) {
            down = s;
            d = i;
            pd = pts[i];
        }
    }
    m.p[0] = pa;
    m.p[1] = pb;
    m.count = 2;
    if (c >= 0) {
        m.p[2] = pc;
        m.count = 3;
    }
    if (d >= 0) {
        if (m.count == 3) {
            m.p[3] = pd;
        } else {
            m.p[2] = pd;
        }
        m.count += 1;
    }
}

TFN_HD void sphereSphere(V ca, float ra, V cb, float rb, float margin, Manifold& m) {
    const V d = cb - ca;
    const float dist = len(d);
    const V n = dist > 1e-9f ? d * (1.0f / dist) : vec3(0, 1, 0);
    const float sep = dist - ra - rb;
    m.count = 0;
    if (sep > margin) { return; }
    m.n = n;
    const V pa = ca + n * ra, pb = cb - n * rb;
    m.p[0] = Point{ (pa + pb) * 0.5f, sep };
    m.count = 1;
}

// Half-space A (below its local xy plane) against sphere B.
TFN_HD void halfSpaceSphere(const Pose& a, V cb, float r, float margin, Manifold& m) {
    const V n = a.ax[2];
    const float offset = dot(n, a.c);
    m.count = 0;
    m.n = n;
    const float sep = dot(n, cb) - offset - r;
    if (sep > margin) { return; }
    const V pb = cb - n * r;
    m.p[0] = Point{ pb - n * (0.5f * sep), sep };
    m.count = 1;
}

constexpr int kMaxHalfSpacePoints = 32;

// Half-space A against the first nv vertices of a convex shape: every vertex within the margin,
// reduced.
template <int NV>
TFN_HD void halfSpacePoints(const Pose& a, const V (&verts)[NV], int nv, float margin, Manifold& m) {
    const V n = a.ax[2];
    const float offset = dot(n, a.c);
    m.count = 0;
    m.n = n;
    Point pts[NV];
    int np = 0;
    TFN_UNROLL
    for (int i = 0; i < NV; ++i) {
        if (i >= nv) { continue; }
        const float sep = dot(n, verts[i]) - offset;
        if (sep <= margin) {
            put(pts, np, Point{ verts[i] - n * (0.5f * sep), sep });
            ++np;
        }
    }
    reduce(pts, np, n, m);
}

TFN_HD void boxVertices(const Pose& b, V half, V (&out)[8]) {
    TFN_UNROLL
    for (int k = 0; k < 8; ++k) {
        out[k] = b.c + b.ax[0] * ((k & 1) ? half.x : -half.x) + b.ax[1] * ((k & 2) ? half.y : -half.y) +
                 b.ax[2] * ((k & 4) ? half.z : -half.z);
    }
}

// Box A against sphere B; the normal points from the box to the sphere.
TFN_HD void boxSphere(V h, const Pose& a, V cs, float r, float margin, Manifold& m) {
    const V d = cs - a.c;
    const V c = vec3(dot(d, a.ax[0]), dot(d, a.ax[1]), dot(d, a.ax[2]));
    const V q = vec3(hi(-h.x, lo(h.x, c.x)), hi(-h.y, lo(h.y, c.y)), hi(-h.z, lo(h.z, c.z)));
    V nl, onBox;
    float sep;
    const V off = c - q;
    const float dist = len(off);
    if (dist > 1e-7f) {
        nl = off * (1.0f / dist);
        sep = dist - r;
        onBox = q;
    } else {
        int axis = 0;
        float pen = h.x - fabsf(c.x);
        for (int i = 1; i < 3; ++i) {
            const float p = comp(h, i) - fabsf(comp(c, i));
            if (p < pen) {
                pen = p;
                axis = i;
            }
        }
        const float s = comp(c, axis) >= 0 ? 1.0f : -1.0f;
        nl = vec3(axis == 0 ? s : 0.0f, axis == 1 ? s : 0.0f, axis == 2 ? s : 0.0f);
        sep = -pen - r;
        onBox = c;
        if (axis == 0) { onBox.x = s * h.x; }
        if (axis == 1) { onBox.y = s * h.y; }
        if (axis == 2) { onBox.z = s * h.z; }
    }
    m.count = 0;
    if (sep > margin) { return; }
    const V n = a.ax[0] * nl.x + a.ax[1] * nl.y + a.ax[2] * nl.z;
    const V pa = a.c + a.ax[0] * onBox.x + a.ax[1] * onBox.y + a.ax[2] * onBox.z;
    const V pb = cs - n * r;
    m.n = n;
    m.p[0] = Point{ (pa + pb) * 0.5f, sep };
    m.count = 1;
}

// A convex polygon clipped by planes: a quad by four planes has at most eight vertices; the room
// beyond that only takes vertices that rounding adds near a plane, and both builds drop the same ones.
constexpr int kMaxPolygon = 16;

// Clips the convex polygon of the first k vertices against the plane dot(n, x) <= d.
TFN_HD void clip(V (&poly)[kMaxPolygon], int& k, V n, float d) {
    V out[kMaxPolygon] = {};
    int m = 0;
    TFN_UNROLL
    for (int i = 0; i < kMaxPolygon; ++i) {
        if (i >= k) { continue; }
        const V p = poly[i], q = pickv(i + 1 < k, poly[(i + 1) % kMaxPolygon], poly[0]);
        const float dp = dot(n, p) - d, dq = dot(n, q) - d;
        if (dp <= 0 && m < kMaxPolygon) {
            put(out, m, p);
            ++m;
        }
        if (((dp < 0 && dq > 0) || (dp > 0 && dq < 0)) && m < kMaxPolygon) {
            put(out, m, p + (q - p) * (dp / (dp - dq)));
            ++m;
        }
    }
    TFN_UNROLL
    for (int i = 0; i < kMaxPolygon; ++i) { poly[i] = out[i]; }
    k = m;
}

TFN_HD float extentAlong(const Pose& p, const float* h, V l) {
    return h[0] * fabsf(dot(p.ax[0], l)) + h[1] * fabsf(dot(p.ax[1], l)) + h[2] * fabsf(dot(p.ax[2], l));
}

// Box A against box B by separating axes: a face of one box clipped against the other's incident
// face, or the closest points of two edges.
TFN_HD void boxBox(V halfA, const Pose& a, V halfB, const Pose& b, float margin, Manifold& m) {
    m.count = 0;
    const V t = b.c - a.c;
    const float ha[3] = { halfA.x, halfA.y, halfA.z }, hb[3] = { halfB.x, halfB.y, halfB.z };
    float bestFace = -1e30f, bestEdge = -1e30f;
    int faceAxis = -1, edgeI = -1, edgeJ = -1;
    V faceN = vec3(0, 0, 0), edgeN = vec3(0, 0, 0);
    TFN_UNROLL
    for (int i = 0; i < 6; ++i) {
        const V l = i < 3 ? of3(a.ax, i) : of3(b.ax, i - 3);
        const float sep = fabsf(dot(t, l)) - extentAlong(a, ha, l) - extentAlong(b, hb, l);
        if (sep > margin) { return; }
        if (sep > bestFace) {
            bestFace = sep;
            faceAxis = i;
            faceN = dot(t, l) >= 0 ? l : -l;
        }
    }
    TFN_UNROLL
    for (int i = 0; i < 3; ++i) {
        TFN_UNROLL
        for (int j = 0; j < 3; ++j) {
            V l = cross(of3(a.ax, i), of3(b.ax, j));
            const float l2 = dot(l, l);
            if (l2 < 1e-6f) { continue; }
            l = l * (1.0f / sqrtf(l2));
            const float sep = fabsf(dot(t, l)) - extentAlong(a, ha, l) - extentAlong(b, hb, l);
            if (sep > margin) { return; }
            if (sep > bestEdge) {
                bestEdge = sep;
                edgeI = i;
                edgeJ = j;
                edgeN = dot(t, l) >= 0 ? l : -l;
            }
        }
    }
    if (edgeI >= 0 && bestEdge > 0.95f * bestFace + 0.01f) {
        // Two edges: the edge of A farthest along n and the edge of B farthest along -n.
        V pa = a.c, pb = b.c;
        TFN_UNROLL
        for (int k = 0; k < 3; ++k) {
            if (k != edgeI) { pa = pa + of3(a.ax, k) * (dot(of3(a.ax, k), edgeN) >= 0 ? of3(ha, k) : -of3(ha, k)); }
            if (k != edgeJ) { pb = pb + of3(b.ax, k) * (dot(of3(b.ax, k), edgeN) >= 0 ? -of3(hb, k) : of3(hb, k)); }
        }
        const V da = of3(a.ax, edgeI), db = of3(b.ax, edgeJ);
        const V r = pa - pb;
        const float e = dot(db, db), f = dot(db, r), c = dot(da, r), bb = dot(da, db);
        const float den = 1.0f - bb * bb;
        float s = den > 1e-9f ? (bb * f - c * e) / den : 0.0f;
        const float hae = of3(ha, edgeI), hbe = of3(hb, edgeJ);
        s = hi(-hae, lo(hae, s));
        float u = (bb * s + f) / e;
        u = hi(-hbe, lo(hbe, u));
        const V qa = pa + da * s, qb = pb + db * u;
        const float sep = dot(qb - qa, edgeN);
        if (sep > margin) { return; }
        m.n = edgeN;
        m.p[0] = Point{ (qa + qb) * 0.5f, sep };
        m.count = 1;
        return;
    }
    const bool refA = faceAxis < 3;
    // (copies, not references: a reference chosen at run time would put both boxes in memory)
    const Pose R = pickpose(refA, a, b);
    const Pose I = pickpose(refA, b, a);
    const float hr[3] = { pickf(refA, ha[0], hb[0]), pickf(refA, ha[1], hb[1]), pickf(refA, ha[2], hb[2]) };
    const float hn[3] = { pickf(refA, hb[0], ha[0]), pickf(refA, hb[1], ha[1]), pickf(refA, hb[2], ha[2]) };
    const int ri = refA ? faceAxis : faceAxis - 3;
    const V rn = refA ? faceN : -faceN;  // the reference face's outward normal, toward the other box
    int ii = 0;
    float most = -1.0f;
    TFN_UNROLL
    for (int k = 0; k < 3; ++k) {
        const float d = fabsf(dot(of3(I.ax, k), rn));
        if (d > most) {
            most = d;
            ii = k;
        }
    }
    const V iax = of3(I.ax, ii);
    const V in = dot(iax, rn) > 0 ? -iax : iax;  // incident face normal, facing the reference face
    const int i1 = (ii + 1) % 3, i2 = (ii + 2) % 3;
    const V ax1 = of3(I.ax, i1), ax2 = of3(I.ax, i2);
    const float h1 = of3(hn, i1), h2 = of3(hn, i2);
    const V fc = I.c + in * of3(hn, ii);
    V poly[kMaxPolygon] = { fc + ax1 * h1 + ax2 * h2, fc - ax1 * h1 + ax2 * h2, fc - ax1 * h1 - ax2 * h2,
                            fc + ax1 * h1 - ax2 * h2 };
    // the side planes of the reference face: its two other axes in order, each both ways
    int k = 4;
    const int j0 = ri == 0 ? 1 : 0, j1 = ri == 2 ? 1 : 2;
    const V r0 = of3(R.ax, j0), r1 = of3(R.ax, j1);
    const float h0 = of3(hr, j0), h1r = of3(hr, j1);
    clip(poly, k, r0, dot(r0, R.c) + h0);
    clip(poly, k, -r0, -dot(r0, R.c) + h0);
    clip(poly, k, r1, dot(r1, R.c) + h1r);
    clip(poly, k, -r1, -dot(r1, R.c) + h1r);
    Point pts[kMaxPolygon];
    int np = 0;
    const float plane = dot(rn, R.c) + of3(hr, ri);
    TFN_UNROLL
    for (int i = 0; i < kMaxPolygon; ++i) {
        if (i >= k) { continue; }
        const float sep = dot(rn, poly[i]) - plane;
        if (sep > margin) { continue; }
        put(pts, np, Point{ poly[i] - rn * (0.5f * sep), sep });
        ++np;
    }
    m.n = faceN;
    reduce(pts, np, faceN, m);
}

// The margin of a pair: the speculative distance grown by the pair's travel over `dt`, from the
// bodies' velocities and extents.
TFN_HD float pairMargin(V va, V wa, float extentA, V vb, V wb, float extentB, float dt, float speculative) {
    const float travel = (len(vb - va) + len(wa) * extentA + len(wb) * extentB) * dt;
    return speculative + travel;
}

// Sphere, box and half-space pairs, A before B: 1 with a manifold, 0 without, -1 for a pair this
// file does not collide (another kind of shape, or two boxes that both stay extruded along the world
// z, which the host world collides as a 2D problem).
TFN_HD int collideBasic(int ka, V ha, float ra, const Pose& pa, int kb, V hb, float rb, const Pose& pb, bool bothInPlane,
                        float margin, Manifold& m) {
    const bool basicA = ka == kSphere || ka == kBox || ka == kHalfSpace;
    const bool basicB = kb == kSphere || kb == kBox || kb == kHalfSpace;
    if (!basicA || !basicB) { return -1; }
    m.count = 0;
    if (ka == kHalfSpace && kb != kHalfSpace) {
        if (kb == kSphere) {
            halfSpaceSphere(pa, pb.c, rb, margin, m);
        } else {
            V verts[8];
            boxVertices(pb, hb, verts);
            halfSpacePoints(pa, verts, 8, margin, m);
        }
    } else if (kb == kHalfSpace) {
        return 0;
    } else if (ka == kSphere && kb == kSphere) {
        sphereSphere(pa.c, ra, pb.c, rb, margin, m);
    } else if (kb == kSphere) {
        boxSphere(ha, pa, pb.c, rb, margin, m);
    } else if (ka == kSphere) {
        boxSphere(hb, pb, pa.c, ra, margin, m);
        m.n = -m.n;
    } else {
        if (bothInPlane) { return -1; }
        boxBox(ha, pa, hb, pb, margin, m);
    }
    return m.count > 0 ? 1 : 0;
}

// A body that stays extruded along the world z: static, or locked to a plane whose normal is z
// (a free body that happens to stand upright is not: it may meet others along z).
TFN_HD bool inPlane(const tfs_body& b) {
    if (b.type == TFS_STATIC) { return true; }
    if ((b.locks & TFS_LOCK_PLANAR) != TFS_LOCK_PLANAR) { return false; }
    return fabsf(rotate(ldq(b.lock_frame), vec3(0, 0, 1)).z) > 0.9999f;
}

// The geometric fields of the contact of a pair from its manifold: normal, tangents, count and per
// point local_a, separation and anchors (src/world.cpp:920-950, 536-540 conventions).
TFN_HD void contactGeometry(const tfs_body& a, const tfs_body& b, const Manifold& m, float speculative, tfs_contact& c) {
    const V n = unit(m.n);
    const V t1 = perpendicular(n);
    st(c.normal, n);
    st(c.tangent1, t1);
    st(c.tangent2, cross(n, t1));
    c.count = m.count;
    const Q qa = ldq(a.q);
    TFN_UNROLL
    for (int i = 0; i < 4; ++i) {
        if (i >= m.count) { continue; }
        tfs_point& p = c.points[i];
        const V pos = m.p[i].pos;
        st(p.local_a, unrotate(qa, pos - ld(a.p)));
        p.separation = m.p[i].sep;
        const V up = p.separation > speculative ? n * (0.5f * p.separation) : vec3(0, 0, 0);
        st(p.anchor_a, pos - up - ld(a.com));
        st(p.anchor_b, pos + up - ld(b.com));
    }
}

}  // namespace tfn
