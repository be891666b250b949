// collide3d.h - the reference's narrowphase for spheres, boxes and half-spaces, restated for the host
// and the device: src/collide3d.cpp (collide 1778-1811, collideHalfSpace 1309-1345, collideSpheres
// 708-717, collideSpherePoly 651-706, collideConvex's two-box branch 615-629 with boxQueries 481-596,
// convexContact 599-613, faceContact 296-340, edgeContact 377-414, clipPolygon 275-287,
// closestOnSegments 342-373, supportFace 77-88, reduceManifold 1497-1575, ownFrame / otherFrame 32-75),
// the box polyhedron of src/polyhedron.cpp (makePolyhedron 8-58, makePrism 60-81, makeBox 83-85), the
// gap test `apart` of src/distance.cpp (GJK in double, 44-274) that two boxes fall back on, and
// Motion::predict / swapped (src/internal.h:354-373).
//
// Every function performs the reference's float (and, in the GJK, double) operations in its order,
// so with -ffp-contract=off and IEEE division and square root a pair gives the reference's manifold
// bit for bit on either side. The reference's growable buffers become fixed arrays: a box face clipped
// by four side planes keeps at most 8 points, and a box has 8 vertices.
#pragma once

#include <float.h>

#include "../../include/twofold_solver.h"
#include "fmath.h"

namespace tfc {

using tf::M3;
using tf::Q4;
using tf::V3;
using tf::cross;
using tf::dot;
using tf::length;
using tf::maxf;
using tf::minf;
using tf::normalize;
using tf::quat;
using tf::rotate;
using tf::vec;

// The shape kinds of tfs_shape (include/twofold_solver.h).
constexpr int kSphere = 0;
constexpr int kBox = 1;
constexpr int kHalfSpace = 4;

// Edge pairs closer to parallel than this sine are left to the face queries (src/collide3d.cpp:12).
constexpr float kParallelEdges = 0.05f;
constexpr float kAligned = 0.9999f;

TF_HD V3 neg(V3 a) { return vec(-a.x, -a.y, -a.z); }
TF_HD V3 vdiv(V3 a, float s) { return vec(a.x / s, a.y / s, a.z / s); }
TF_HD float lengthSq(V3 v) { return dot(v, v); }
TF_HD V3 invRotate(Q4 q, V3 v) { return rotate(tf::conjugate(q), v); }
TF_HD float comp(V3 v, int k) { return k == 0 ? v.x : k == 1 ? v.y : v.z; }
TF_HD V3 perpendicularOf(V3 n) {
    if (fabsf(n.x) > 0.57735f) { return normalize(vec(n.y, -n.x, 0.0f)); }
    return normalize(vec(0.0f, n.z, -n.y));
}

// Transform (include/twofold/math.h:232-239).
struct Xf {
    V3 p;
    Q4 q;
};
TF_HD V3 apply(const Xf& x, V3 v) { return x.p + rotate(x.q, v); }
TF_HD V3 applyInverse(const Xf& x, V3 v) { return invRotate(x.q, v - x.p); }
TF_HD Xf compose(const Xf& a, const Xf& b) { return Xf{ apply(a, b.p), tf::operator*(a.q, b.q) }; }

// Motion (src/internal.h:354-373): the two bodies' velocities over the coming time dt.
struct Motion {
    V3 linearA, angularA, centerA, linearB, angularB, centerB;
    float dt;
    bool fast;
};
TF_HD float predict(const Motion& m, V3 p, V3 n, float separation) {
    const V3 va = m.linearA + cross(m.angularA, p - m.centerA);
    const V3 vb = m.linearB + cross(m.angularB, p - m.centerB);
    return separation + m.dt * dot(n, vb - va);
}
TF_HD Motion swapped(const Motion& m) {
    return Motion{ m.linearB, m.angularB, m.centerB, m.linearA, m.angularA, m.centerA, m.dt, m.fast };
}

struct ManifoldPoint {
    V3 position;
    float separation;
    float predicted;
};
struct Manifold {
    V3 normal;
    ManifoldPoint points[4];
    int count;
};

// ---- the box polyhedron ------------------------------------------------------------------------------
//
// makeBox(h) = makePrism of the rectangle (-x,-y), (x,-y), (x,y), (-x,y) between z = -h.z and h.z:
// vertices 0-3 at the bottom, 4-7 at the top; faces 0-3 the sides {i, i+1, 4+(i+1), 4+i} (mod 4 in each
// ring), 4 the top {4,5,6,7}, 5 the bottom {3,2,1,0}; every face has 4 indices, so face f starts at 4f.
// Edges in makePolyhedron's order (directed vertex pairs a < b in lexicographic order):
// (0,1) (0,3) (0,4) (1,2) (1,5) (2,3) (2,6) (3,7) (4,5) (4,7) (5,6) (6,7).
TF_HD int faceVertex(int f, int k) {
    if (f < 4) { return k == 0 ? f : k == 1 ? ((f + 1) & 3) : k == 2 ? 4 + ((f + 1) & 3) : 4 + f; }
    return f == 4 ? 4 + k : 3 - k;
}
TF_HD int edgeV0(int e) { return (int)((0x654432211000ull >> (4 * e)) & 15u); }
TF_HD int edgeV1(int e) { return (int)((0x767576352431ull >> (4 * e)) & 15u); }
// boxTopology (src/collide3d.cpp:450-476) for makeBox's numbering: the face of axis k on side s (1
// positive), and the edge running along axis i between the faces of the next two axes on sides s1, s2.
TF_HD int topoFace(int k, int s) { return (int)((0x452013u >> (4 * (2 * k + s))) & 15u); }
TF_HD int topoEdge(int i, int s1, int s2) { return (int)((0x6472a931b580ull >> (4 * (4 * i + 2 * s1 + s2))) & 15u); }

// A box's vertices follow from its half-extents and their index: makeBox's vertex i is
// ((i + 1) & 2 ? x : -x, i & 2 ? y : -y, i < 4 ? -z : z), and otherFrame carries it into the pair's frame
// as r * v + t. Only the face normals and offsets are kept, and a normal or offset at an index known
// only at run time is picked by selects: no box is ever held in memory indexed at run time, which on
// the device would put it in scratch memory.
struct BoxPoly {
    V3 h;
    M3 r;
    V3 t;
    bool moved;
    V3 faceN[6];
    float faceD[6];
};

TF_HD V3 localVertex(V3 h, int i) { return vec(((i + 1) & 2) ? h.x : -h.x, (i & 2) ? h.y : -h.y, i < 4 ? -h.z : h.z); }
TF_HD V3 vertexOf(const BoxPoly& p, int i) {
    const V3 v = localVertex(p.h, i);
    return p.moved ? tf::operator*(p.r, v) + p.t : v;
}
// Picked component by component by a chain of selects between floats: a loop over the array, an index
// into it or a select between whole vectors is turned back into an access at a run-time index.
TF_HD float pick6(int f, float a0, float a1, float a2, float a3, float a4, float a5) {
    return f == 0 ? a0 : f == 1 ? a1 : f == 2 ? a2 : f == 3 ? a3 : f == 4 ? a4 : a5;
}
TF_HD V3 normalOf(const BoxPoly& p, int f) {
    const V3* n = p.faceN;
    return vec(pick6(f, n[0].x, n[1].x, n[2].x, n[3].x, n[4].x, n[5].x),
               pick6(f, n[0].y, n[1].y, n[2].y, n[3].y, n[4].y, n[5].y),
               pick6(f, n[0].z, n[1].z, n[2].z, n[3].z, n[4].z, n[5].z));
}
TF_HD float offsetOf(const BoxPoly& p, int f) {
    const float* d = p.faceD;
    return pick6(f, d[0], d[1], d[2], d[3], d[4], d[5]);
}

// makePolyhedron's vertices, normals (Newell's method about the centroid) and offsets for makeBox(h).
TF_HD void makeBoxPoly(V3 h, BoxPoly& p) {
    p.h = h;
    p.r = M3{ { vec(1.0f, 0.0f, 0.0f), vec(0.0f, 1.0f, 0.0f), vec(0.0f, 0.0f, 1.0f) } };
    p.t = vec(0.0f, 0.0f, 0.0f);
    p.moved = false;
    V3 c = vec(0.0f, 0.0f, 0.0f);
#pragma unroll
    for (int i = 0; i < 8; ++i) { c = c + localVertex(h, i); }
    c = vdiv(c, 8.0f);
#pragma unroll
    for (int f = 0; f < 6; ++f) {
        V3 n = vec(0.0f, 0.0f, 0.0f);
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const V3 a = localVertex(h, faceVertex(f, i)) - c;
            const V3 b = localVertex(h, faceVertex(f, (i + 1) & 3)) - c;
            n.x += (a.y - b.y) * (a.z + b.z);
            n.y += (a.z - b.z) * (a.x + b.x);
            n.z += (a.x - b.x) * (a.y + b.y);
        }
        p.faceN[f] = normalize(n);
        p.faceD[f] = dot(p.faceN[f], localVertex(h, faceVertex(f, 0)));
    }
}

// otherFrame (src/collide3d.cpp:47-75): the box's vertices, normals and offsets in the frame of the
// pair's first shape (the edge arcs and bounds it also makes are not read for two boxes). p is a box
// in its own frame.
TF_HD void otherFrame(const BoxPoly& p, const Xf& xf, BoxPoly& o) {
    o.h = p.h;
    o.r = tf::toMat3(xf.q);
    o.t = xf.p;
    o.moved = true;
#pragma unroll
    for (int f = 0; f < 6; ++f) {
        o.faceN[f] = tf::operator*(o.r, p.faceN[f]);
        o.faceD[f] = dot(o.faceN[f], vertexOf(o, faceVertex(f, 0)));
    }
}

// supportFace (src/collide3d.cpp:77-88).
TF_HD int supportFace(const BoxPoly& p, V3 n) {
    int best = 0;
    float most = dot(p.faceN[0], n);
#pragma unroll
    for (int i = 1; i < 6; ++i) {
        const float d = dot(p.faceN[i], n);
        if (d > most) {
            most = d;
            best = i;
        }
    }
    return best;
}

// ---- manifold reduction ------------------------------------------------------------------------------

constexpr int kMaxPoints = 8;

// reduceManifold (src/collide3d.cpp:1497-1575).
TF_HD void reduceManifold(const ManifoldPoint* pts, int np, V3 normal, float band, Manifold& m, float tie, bool fast) {
    m.normal = normal;
    m.count = 0;
    if (np <= 4) {
        for (int i = 0; i < np; ++i) { m.points[m.count++] = pts[i]; }
        return;
    }
    int i0 = 0;
    for (int i = 1; i < np; ++i) {
        if (pts[i].predicted < pts[i0].predicted) { i0 = i; }
    }
    for (int i = 0; i < i0; ++i) {
        if (pts[i].predicted <= pts[i0].predicted + tie) {
            i0 = i;
            break;
        }
    }
    const V3 p0 = pts[i0].position;
    const float limit = fast ? maxf(pts[i0].predicted + band, 0.0f) : pts[i0].predicted + band;
    int chosen[4] = { i0, i0, i0, i0 };
#define TFC_USED(i) ((i) == chosen[0] || (i) == chosen[1] || (i) == chosen[2] || (i) == chosen[3])
    m.points[m.count++] = pts[i0];

    int i1 = np;
    float best = 0.0f;
    for (int i = 0; i < np; ++i) {
        if (pts[i].predicted > limit) { continue; }
        const float d = lengthSq(pts[i].position - p0);
        if (d > best) {
            best = d;
            i1 = i;
        }
    }

    int i2 = np;
    float area = 0.0f;
    if (i1 < np) {
        chosen[1] = i1;
        m.points[m.count++] = pts[i1];
        const V3 p1 = pts[i1].position;
        for (int i = 0; i < np; ++i) {
            if (TFC_USED(i) || pts[i].predicted > limit) { continue; }
            const float a = dot(cross(p1 - p0, pts[i].position - p0), normal);
            if (fabsf(a) > fabsf(area)) {
                area = a;
                i2 = i;
            }
        }
    }

    if (i2 < np) {
        chosen[2] = i2;
        m.points[m.count++] = pts[i2];
        const float sign = area < 0.0f ? -1.0f : 1.0f;
        const V3 tri[3] = { p0, m.points[1].position, pts[i2].position };
        int i3 = np;
        float outside = 0.0f;
        for (int i = 0; i < np; ++i) {
            if (TFC_USED(i) || pts[i].predicted > limit) { continue; }
            for (int e = 0; e < 3; ++e) {
                const V3 a = tri[e];
                const V3 b = tri[(e + 1) % 3];
                const float s = -sign * dot(cross(b - a, pts[i].position - a), normal);
                if (s > outside) {
                    outside = s;
                    i3 = i;
                }
            }
        }
        if (i3 < np) {
            chosen[3] = i3;
            m.points[m.count++] = pts[i3];
        }
    }

    while (m.count < 4) {
        int next = np;
        for (int i = 0; i < np; ++i) {
            if (TFC_USED(i)) { continue; }
            if (next == np || pts[i].predicted < pts[next].predicted) { next = i; }
        }
        if (next == np) { break; }
        chosen[m.count] = next;
        m.points[m.count++] = pts[next];
    }
#undef TFC_USED
}

// ---- spheres and half-spaces --------------------------------------------------------------------------

// collideSpheres (src/collide3d.cpp:708-717).
TF_HD void collideSpheres(V3 ca, float ra, V3 cb, float rb, float margin, Manifold& m) {
    const V3 d = cb - ca;
    const float dist = length(d);
    const float sep = dist - ra - rb;
    if (sep > margin) { return; }
    const V3 n = dist > 0.0f ? vdiv(d, dist) : vec(0.0f, 1.0f, 0.0f);
    m.normal = n;
    m.points[0] = ManifoldPoint{ ca + n * (ra + 0.5f * sep), sep, 0.0f };
    m.count = 1;
}

// collideSpherePoly (src/collide3d.cpp:651-706) with the box as the polyhedron: normal from the box to
// the sphere.
TF_HD void collideSphereBox(const BoxPoly& p, const Xf& xp, V3 center, float r, float margin, Manifold& m) {
    const V3 c = applyInverse(xp, center);
    int best = -1;
    float bestSep = -FLT_MAX;
    bool outside = false;
#pragma unroll
    for (int i = 0; i < 6; ++i) {
        const float s = dot(p.faceN[i], c) - p.faceD[i];
        if (s > r + margin) { return; }
        if (s > 0.0f) { outside = true; }
        if (s > bestSep) {
            bestSep = s;
            best = i;
        }
    }
    if (best < 0) { return; }

    if (!outside) {
        const V3 n = normalOf(p, best);
        m.normal = rotate(xp.q, n);
        m.points[0] = ManifoldPoint{ apply(xp, c - n * (0.5f * (bestSep + r))), bestSep - r, 0.0f };
        m.count = 1;
        return;
    }

    V3 closest = vec(0.0f, 0.0f, 0.0f);
    float closestSq = FLT_MAX;
#pragma unroll
    for (int i = 0; i < 6; ++i) {
        const V3 fn = p.faceN[i];
        const float s = dot(fn, c) - p.faceD[i];
        if (s <= 0.0f) { continue; }
        const V3 proj = c - fn * s;
        bool inside = true;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const V3 a = vertexOf(p, faceVertex(i, k));
            const V3 b = vertexOf(p, faceVertex(i, (k + 1) & 3));
            if (dot(cross(b - a, fn), proj - a) > 0.0f) {
                inside = false;
                const V3 ab = b - a;
                const float t = minf(maxf(dot(c - a, ab) / maxf(dot(ab, ab), FLT_MIN), 0.0f), 1.0f);
                const V3 q = a + ab * t;
                const float dsq = lengthSq(c - q);
                if (dsq < closestSq) {
                    closestSq = dsq;
                    closest = q;
                }
            }
        }
        if (inside && s * s < closestSq) {
            closestSq = s * s;
            closest = proj;
        }
    }
    const float dist = sqrtf(closestSq);
    const float sep = dist - r;
    if (sep > margin || dist <= 0.0f) { return; }
    const V3 n = vdiv(c - closest, dist);
    m.normal = rotate(xp.q, n);
    m.points[0] = ManifoldPoint{ apply(xp, (closest + c - n * r) * 0.5f), sep, 0.0f };
    m.count = 1;
}

// collideHalfSpace (src/collide3d.cpp:1309-1345) against a sphere (other = sphere of radius r) or a box.
TF_HD void collideHalfSpace(const Xf& xh, int otherKind, const BoxPoly& box, float radius, const Xf& xo, float margin,
                            float band, const Motion& motion, Manifold& m) {
    const V3 n = rotate(xh.q, vec(0.0f, 0.0f, 1.0f));
    const float offset = dot(n, xh.p);
    if (otherKind == kSphere) {
        const float d = dot(n, xo.p) - offset;
        const float sep = d - radius;
        if (sep > margin) { return; }
        m.normal = n;
        m.points[0] = ManifoldPoint{ xo.p - n * (0.5f * (d + radius)), sep, 0.0f };
        m.count = 1;
        return;
    }
    ManifoldPoint pts[kMaxPoints];
    int np = 0;
    const V3 localN = invRotate(xo.q, n);
    const float localOffset = offset - dot(n, xo.p);
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        const V3 v = vertexOf(box, i);
        const float s = dot(localN, v) - localOffset;
        if (s > margin) { continue; }
        const V3 w = apply(xo, v);
        pts[np++] = ManifoldPoint{ w - n * (0.5f * s), s, predict(motion, w, n, s) };
    }
    reduceManifold(pts, np, n, band, m, 0.0f, motion.fast);
}

// ---- two boxes ---------------------------------------------------------------------------------------

struct FaceQuery {
    int index;
    float separation;
};
struct EdgeQuery {
    int a, b;
    float separation;
    V3 axis;
};

// boxQueries (src/collide3d.cpp:481-596): the face and edge queries of two boxes in closed form, b in
// the frame of a. False once an axis separates them by more than margin; unsure when they can still be
// further apart than margin.
TF_HD bool boxQueries(V3 halfA, V3 halfB, const Xf& x, float margin, FaceQuery& fa, FaceQuery& fb, EdgeQuery& eq,
                      bool& unsure) {
    const M3 rot = tf::toMat3(x.q);
    const float r[3][3] = { { rot.r[0].x, rot.r[0].y, rot.r[0].z },
                            { rot.r[1].x, rot.r[1].y, rot.r[1].z },
                            { rot.r[2].x, rot.r[2].y, rot.r[2].z } };
    float ar[3][3];
    for (int i = 0; i < 3; ++i) {
        for (int j = 0; j < 3; ++j) { ar[i][j] = fabsf(r[i][j]); }
    }
    const float t[3] = { x.p.x, x.p.y, x.p.z };
    const float ha[3] = { halfA.x, halfA.y, halfA.z }, hb[3] = { halfB.x, halfB.y, halfB.z };

    int faceA = 0, faceB = 0;
    for (int i = 0; i < 3; ++i) {
        const float s = fabsf(t[i]) - ha[i] - (hb[0] * ar[i][0] + hb[1] * ar[i][1] + hb[2] * ar[i][2]);
        if (s > margin) { return false; }
        if (s > fa.separation) {
            fa.separation = s;
            faceA = i;
        }
    }
    float tb[3];
    for (int j = 0; j < 3; ++j) {
        tb[j] = t[0] * r[0][j] + t[1] * r[1][j] + t[2] * r[2][j];
        const float s = fabsf(tb[j]) - hb[j] - (ha[0] * ar[0][j] + ha[1] * ar[1][j] + ha[2] * ar[2][j]);
        if (s > margin) { return false; }
        if (s > fb.separation) {
            fb.separation = s;
            faceB = j;
        }
    }

    int edgeI = -1, edgeJ = -1;
    float edgeSign = 1.0f, edgeLength = 1.0f;
    for (int i = 0; i < 3; ++i) {
        const int i1 = (i + 1) % 3, i2 = (i + 2) % 3;
        for (int j = 0; j < 3; ++j) {
            const int j1 = (j + 1) % 3, j2 = (j + 2) % 3;
            const float l1 = -r[i2][j], l2 = r[i1][j];
            if (l1 == 0.0f || l2 == 0.0f || r[i][j1] == 0.0f || r[i][j2] == 0.0f) { continue; }
            const float len = sqrtf(l1 * l1 + l2 * l2);
            if (len < kParallelEdges) { continue; }
            const float d = t[i1] * l1 + t[i2] * l2;
            const float ra = ha[i1] * ar[i2][j] + ha[i2] * ar[i1][j];
            const float rb = hb[j1] * ar[i][j2] + hb[j2] * ar[i][j1];
            const float s = (fabsf(d) - ra - rb) / len;
            if (s > margin) { return false; }
            if (s > eq.separation) {
                eq.separation = s;
                edgeI = i;
                edgeJ = j;
                edgeSign = d < 0.0f ? -1.0f : 1.0f;
                edgeLength = len;
            }
        }
    }

    fa.index = topoFace(faceA, t[faceA] >= 0.0f ? 1 : 0);
    fb.index = topoFace(faceB, tb[faceB] < 0.0f ? 1 : 0);
    if (edgeI >= 0) {
        const int i = edgeI, j = edgeJ;
        const int i1 = (i + 1) % 3, i2 = (i + 2) % 3, j1 = (j + 1) % 3, j2 = (j + 2) % 3;
        const float l1 = -r[i2][j] * edgeSign, l2 = r[i1][j] * edgeSign;
        float n[3];
        n[i] = 0.0f;
        n[i1] = l1 / edgeLength;
        n[i2] = l2 / edgeLength;
        eq.axis = vec(n[0], n[1], n[2]);
        const float m1 = r[i][j2] * edgeSign, m2 = -r[i][j1] * edgeSign;
        eq.a = topoEdge(i, l1 > 0.0f ? 1 : 0, l2 > 0.0f ? 1 : 0);
        eq.b = topoEdge(j, m1 < 0.0f ? 1 : 0, m2 < 0.0f ? 1 : 0);
    }

    unsure = false;
    const float gap = maxf(maxf(fa.separation, fb.separation), eq.separation);
    if (gap <= 0.0f) { return true; }
    if (edgeI >= 0 && gap == eq.separation) {
        const int i = edgeI, j = edgeJ;
        const int i1 = (i + 1) % 3, i2 = (i + 2) % 3, j1 = (j + 1) % 3, j2 = (j + 2) % 3;
        const float sideB1 = r[i][j2] * edgeSign < 0.0f ? hb[j1] : -hb[j1];
        const float sideB2 = -r[i][j1] * edgeSign < 0.0f ? hb[j2] : -hb[j2];
        float w[3];
        for (int k = 0; k < 3; ++k) { w[k] = -t[k] - sideB1 * r[k][j1] - sideB2 * r[k][j2]; }
        w[i1] += -r[i2][j] * edgeSign > 0.0f ? ha[i1] : -ha[i1];
        w[i2] += r[i1][j] * edgeSign > 0.0f ? ha[i2] : -ha[i2];
        const float b = r[i][j], f = w[0] * r[0][j] + w[1] * r[1][j] + w[2] * r[2][j];
        const float s = (b * f - w[i]) / (1.0f - b * b);
        unsure = fabsf(s) > ha[i] || fabsf(f + s * b) > hb[j];
    } else if (gap == fa.separation) {
        const int i = faceA;
        const float side = t[i] < 0.0f ? -1.0f : 1.0f;
        for (int k = 0; k < 3 && !unsure; ++k) {
            if (k == i) { continue; }
            float v = t[k];
            for (int j = 0; j < 3; ++j) { v += side * r[i][j] > 0.0f ? -hb[j] * r[k][j] : hb[j] * r[k][j]; }
            unsure = fabsf(v) > ha[k];
        }
    } else {
        const int j = faceB;
        const float side = tb[j] < 0.0f ? -1.0f : 1.0f;
        for (int k = 0; k < 3 && !unsure; ++k) {
            if (k == j) { continue; }
            float v = -tb[k];
            for (int i = 0; i < 3; ++i) { v += side * r[i][j] > 0.0f ? ha[i] * r[i][k] : -ha[i] * r[i][k]; }
            unsure = fabsf(v) > hb[k];
        }
    }
    return true;
}

// clipPolygon (src/collide3d.cpp:275-287): keeps the part of poly[0..n) on the side of the plane
// through origin that n points away from.
TF_HD void clipPolygon(V3* poly, int& count, V3 n, V3 origin) {
    if (count == 0) { return; }
    V3 out[kMaxPoints];
    int k = 0;
    for (int i = 0; i < count; ++i) {
        const V3 p = poly[i];
        const V3 q = poly[(i + 1) % count];
        const float dp = dot(n, p - origin);
        const float dq = dot(n, q - origin);
        if (dp <= 0.0f && k < kMaxPoints) { out[k++] = p; }
        if ((dp <= 0.0f) != (dq <= 0.0f) && k < kMaxPoints) { out[k++] = p + (q - p) * (dp / (dp - dq)); }
    }
    for (int i = 0; i < k; ++i) { poly[i] = out[i]; }
    count = k;
}

// faceContact (src/collide3d.cpp:296-340): the incident face of inc clipped against the side planes of
// face refFace of ref, in the pair's frame, placed in the world by xf.
TF_HD void faceContact(const BoxPoly& ref, int refFace, const BoxPoly& inc, bool flip, const Xf& xf, float margin,
                       float band, const Motion& motion, Manifold& m) {
    const V3 refN = normalOf(ref, refFace);
    const float refD = offsetOf(ref, refFace);
    const V3 worldN = rotate(xf.q, refN);

    int incFace = 0;
    float minDot = FLT_MAX;
#pragma unroll
    for (int i = 0; i < 6; ++i) {
        const float d = dot(inc.faceN[i], refN);
        if (d < minDot) {
            minDot = d;
            incFace = i;
        }
    }

    V3 poly[kMaxPoints];
    int count = 4;
    for (int k = 0; k < 4; ++k) { poly[k] = vertexOf(inc, faceVertex(incFace, k)); }
    for (int k = 0; k < 4; ++k) {
        const V3 a = vertexOf(ref, faceVertex(refFace, k));
        const V3 b = vertexOf(ref, faceVertex(refFace, (k + 1) & 3));
        clipPolygon(poly, count, cross(b - a, refN), a);
    }

    ManifoldPoint pts[kMaxPoints];
    int np = 0;
    for (int k = 0; k < count; ++k) {
        const V3 p = poly[k];
        const float s = dot(refN, p) - refD;
        if (s > margin) { continue; }
        const V3 w = apply(xf, p);
        pts[np++] = ManifoldPoint{ w - worldN * (0.5f * s), s, predict(motion, w, worldN, s) };
    }
    if (np == 0) {
        V3 deepest = vertexOf(inc, 0);
#pragma unroll
        for (int k = 1; k < 8; ++k) {
            const V3 v = vertexOf(inc, k);
            if (dot(refN, v) < dot(refN, deepest)) { deepest = v; }
        }
        const float s = dot(refN, deepest) - refD;
        if (s <= margin) {
            const V3 w = apply(xf, deepest);
            pts[np++] = ManifoldPoint{ w - worldN * (0.5f * s), s, predict(motion, w, worldN, s) };
        }
    }
    reduceManifold(pts, np, flip ? neg(worldN) : worldN, band, m, 0.0f, motion.fast);
}

// closestOnSegments (src/collide3d.cpp:342-373).
TF_HD void closestOnSegments(V3 p1, V3 q1, V3 p2, V3 q2, V3& c1, V3& c2) {
    const V3 d1 = q1 - p1, d2 = q2 - p2, r = p1 - p2;
    const float a = dot(d1, d1), e = dot(d2, d2), f = dot(d2, r);
    float s = 0.0f, t = 0.0f;
    if (a <= FLT_EPSILON && e <= FLT_EPSILON) {
        c1 = p1;
        c2 = p2;
        return;
    }
    if (a <= FLT_EPSILON) {
        t = minf(maxf(f / e, 0.0f), 1.0f);
    } else {
        const float c = dot(d1, r);
        if (e <= FLT_EPSILON) {
            s = minf(maxf(-c / a, 0.0f), 1.0f);
        } else {
            const float b = dot(d1, d2);
            const float denom = a * e - b * b;
            s = denom != 0.0f ? minf(maxf((b * f - c * e) / denom, 0.0f), 1.0f) : 0.0f;
            t = (b * s + f) / e;
            if (t < 0.0f) {
                t = 0.0f;
                s = minf(maxf(-c / a, 0.0f), 1.0f);
            } else if (t > 1.0f) {
                t = 1.0f;
                s = minf(maxf((b - c) / a, 0.0f), 1.0f);
            }
        }
    }
    c1 = p1 + d1 * s;
    c2 = p2 + d2 * t;
}

// edgeContact (src/collide3d.cpp:377-414): the face of b toward a clipped against the face of a toward b,
// seen along the edge axis; one point between the two edges when fewer than two points are left.
TF_HD void edgeContact(const BoxPoly& a, const BoxPoly& b, const EdgeQuery& eq, const Xf& xf, float margin, float band,
                       const Motion& motion, Manifold& m) {
    const V3 n = eq.axis;
    const V3 worldN = rotate(xf.q, n);
    const int fa = supportFace(a, n);
    const V3 faN = normalOf(a, fa);
    const float faD = offsetOf(a, fa);
    const float cosA = dot(faN, n);
    ManifoldPoint pts[kMaxPoints];
    int np = 0;
    if (cosA > 0.0f) {
        const int fi = supportFace(b, neg(n));
        V3 poly[kMaxPoints];
        int count = 4;
        for (int k = 0; k < 4; ++k) { poly[k] = vertexOf(b, faceVertex(fi, k)); }
        for (int k = 0; k < 4; ++k) {
            const V3 p = vertexOf(a, faceVertex(fa, k));
            const V3 q = vertexOf(a, faceVertex(fa, (k + 1) & 3));
            clipPolygon(poly, count, cross(q - p, n), p);
        }
        for (int k = 0; k < count; ++k) {
            const V3 p = poly[k];
            const float s = maxf((dot(faN, p) - faD) / cosA, eq.separation);
            if (s > margin) { continue; }
            const V3 w = apply(xf, p);
            pts[np++] = ManifoldPoint{ w - worldN * (0.5f * s), s, predict(motion, w, worldN, s) };
        }
    }
    if (np >= 2) {
        reduceManifold(pts, np, worldN, band, m, 0.0f, motion.fast);
        return;
    }
    V3 c1, c2;
    closestOnSegments(vertexOf(a, edgeV0(eq.a)), vertexOf(a, edgeV1(eq.a)), vertexOf(b, edgeV0(eq.b)),
                      vertexOf(b, edgeV1(eq.b)), c1, c2);
    m.normal = worldN;
    m.points[0] = ManifoldPoint{ apply(xf, (c1 + c2) * 0.5f), eq.separation, 0.0f };
    m.count = 1;
}

// convexContact (src/collide3d.cpp:599-613).
TF_HD void convexContact(const BoxPoly& a, const BoxPoly& b, const FaceQuery& fa, const FaceQuery& fb,
                         const EdgeQuery& eq, const Xf& xa, float margin, float linearSlop, const Motion& motion,
                         Manifold& m) {
    const float absTol = 0.5f * linearSlop;
    if (eq.a >= 0 && eq.separation > 0.9f * maxf(fa.separation, fb.separation) + absTol) {
        edgeContact(a, b, eq, xa, margin, linearSlop, motion, m);
    } else if (fb.separation > 0.98f * fa.separation + absTol) {
        faceContact(b, fb.index, a, true, xa, margin, linearSlop, swapped(motion), m);
    } else {
        faceContact(a, fa.index, b, false, xa, margin, linearSlop, motion, m);
    }
}

// ---- apart: GJK in double (src/distance.cpp) ----------------------------------------------------------

struct DP {
    double x, y, z;
};
TF_HD DP dp(V3 v) { return DP{ (double)v.x, (double)v.y, (double)v.z }; }
TF_HD DP operator+(DP a, DP b) { return DP{ a.x + b.x, a.y + b.y, a.z + b.z }; }
TF_HD DP operator-(DP a, DP b) { return DP{ a.x - b.x, a.y - b.y, a.z - b.z }; }
TF_HD DP operator*(DP a, double s) { return DP{ a.x * s, a.y * s, a.z * s }; }
TF_HD double dot(DP a, DP b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
TF_HD DP cross(DP a, DP b) { return DP{ a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x }; }
TF_HD double clampd(double v, double lo, double hi) { return v < lo ? lo : hi < v ? hi : v; }

struct GVertex {
    V3 a, b, w;
};
struct Simplex {
    GVertex v[4];
    double u[4];
    int count;
};

// support (src/distance.cpp:44-57) over a box's vertices.
TF_HD V3 support(const BoxPoly& c, const Xf& xf, V3 d) {
    const V3 local = invRotate(xf.q, d);
    int best = 0;
    float most = dot(vertexOf(c, 0), local);
#pragma unroll
    for (int i = 1; i < 8; ++i) {
        const float s = dot(vertexOf(c, i), local);
        if (s > most) {
            most = s;
            best = i;
        }
    }
    return apply(xf, vertexOf(c, best));
}

// segmentWeights (src/distance.cpp:59-65).
TF_HD void segmentWeights(DP a, DP b, double& ua, double& ub) {
    const DP e = b - a;
    const double ee = dot(e, e);
    const double t = ee > 0.0 ? clampd(-dot(a, e) / ee, 0.0, 1.0) : 0.0;
    ua = 1.0 - t;
    ub = t;
}

// triangleWeights (src/distance.cpp:68-123).
TF_HD void triangleWeights(DP a, DP b, DP c, double* u) {
    u[0] = u[1] = u[2] = 0.0;
    const DP ab = b - a, ac = c - a;
    const double d1 = -dot(ab, a), d2 = -dot(ac, a);
    if (d1 <= 0.0 && d2 <= 0.0) {
        u[0] = 1.0;
        return;
    }
    const double d3 = -dot(ab, b), d4 = -dot(ac, b);
    if (d3 >= 0.0 && d4 <= d3) {
        u[1] = 1.0;
        return;
    }
    const double vc = d1 * d4 - d3 * d2;
    if (vc <= 0.0 && d1 >= 0.0 && d3 <= 0.0) {
        segmentWeights(a, b, u[0], u[1]);
        return;
    }
    const double d5 = -dot(ab, c), d6 = -dot(ac, c);
    if (d6 >= 0.0 && d5 <= d6) {
        u[2] = 1.0;
        return;
    }
    const double vb = d5 * d2 - d1 * d6;
    if (vb <= 0.0 && d2 >= 0.0 && d6 <= 0.0) {
        segmentWeights(a, c, u[0], u[2]);
        return;
    }
    const double va = d3 * d6 - d5 * d4;
    if (va <= 0.0 && d4 - d3 >= 0.0 && d5 - d6 >= 0.0) {
        segmentWeights(b, c, u[1], u[2]);
        return;
    }
    const double sum = va + vb + vc;
    if (sum > DBL_MIN) {
        u[0] = va / sum;
        u[1] = vb / sum;
        u[2] = vc / sum;
        return;
    }
    const DP p[3] = { a, b, c };
    double best = DBL_MAX;
    for (int i = 0; i < 3; ++i) {
        const int j = (i + 1) % 3;
        double ui, uj;
        segmentWeights(p[i], p[j], ui, uj);
        const DP q = p[i] * ui + p[j] * uj;
        if (dot(q, q) < best) {
            best = dot(q, q);
            u[0] = u[1] = u[2] = 0.0;
            u[i] = ui;
            u[j] = uj;
        }
    }
}

// tetrahedronWeights (src/distance.cpp:126-147): false when the origin lies inside.
TF_HD int tetraFace(int f, int k) { return (int)((0x0231132021303210ull >> (4 * (4 * f + k))) & 15u); }
TF_HD bool tetrahedronWeights(const DP* w, double* u) {
    bool outside = false;
    double best = DBL_MAX;
    for (int fi = 0; fi < 4; ++fi) {
        const int f0 = tetraFace(fi, 0), f1 = tetraFace(fi, 1), f2 = tetraFace(fi, 2), f3 = tetraFace(fi, 3);
        const DP n = cross(w[f1] - w[f0], w[f2] - w[f0]);
        const DP apex = w[f3] - w[f0];
        const double height = dot(apex, n);
        if (-dot(w[f0], n) * height > 0.0 && height * height > 1e-8 * dot(n, n) * dot(apex, apex)) { continue; }
        outside = true;
        double t[3];
        triangleWeights(w[f0], w[f1], w[f2], t);
        const DP q = w[f0] * t[0] + w[f1] * t[1] + w[f2] * t[2];
        if (dot(q, q) < best) {
            best = dot(q, q);
            u[0] = u[1] = u[2] = u[3] = 0.0;
            u[f0] = t[0];
            u[f1] = t[1];
            u[f2] = t[2];
        }
    }
    return outside;
}

// solve (src/distance.cpp:149-168).
TF_HD bool solveSimplex(Simplex& s) {
    double u[4] = { 1.0, 0.0, 0.0, 0.0 };
    if (s.count == 2) {
        segmentWeights(dp(s.v[0].w), dp(s.v[1].w), u[0], u[1]);
    } else if (s.count == 3) {
        triangleWeights(dp(s.v[0].w), dp(s.v[1].w), dp(s.v[2].w), u);
    } else if (s.count == 4) {
        const DP w[4] = { dp(s.v[0].w), dp(s.v[1].w), dp(s.v[2].w), dp(s.v[3].w) };
        if (!tetrahedronWeights(w, u)) { return false; }
    }
    int kept = 0;
    for (int i = 0; i < s.count; ++i) {
        if (u[i] <= 0.0) { continue; }
        s.v[kept] = s.v[i];
        s.u[kept] = u[i];
        ++kept;
    }
    s.count = kept;
    return true;
}

// closest (src/distance.cpp:170-174).
TF_HD V3 closestOf(const Simplex& s) {
    DP v = DP{ 0.0, 0.0, 0.0 };
    for (int i = 0; i < s.count; ++i) { v = v + dp(s.v[i].w) * s.u[i]; }
    return vec((float)v.x, (float)v.y, (float)v.z);
}

// search (src/distance.cpp:215-264) from a zero start direction with limit >= 0: 1 when the gap between
// the boxes is known to be beyond the limit, 0 when within it, -1 when the search ran out (Open), with
// v the simplex point nearest the origin.
TF_HD int searchBoxes(const BoxPoly& a, const Xf& xa, const BoxPoly& b, const Xf& xb, float limit, V3& v) {
    const bool bounded = limit >= 0.0f;
    Simplex s;
    s.v[0].a = apply(xa, vertexOf(a, 0));
    s.v[0].b = apply(xb, vertexOf(b, 0));
    s.v[0].w = s.v[0].b - s.v[0].a;
    s.u[0] = 1.0;
    s.u[1] = s.u[2] = s.u[3] = 0.0;
    s.count = 1;
    v = s.v[0].w;
    for (int iter = 0; iter < 32; ++iter) {
        const float vv = dot(v, v);
        if (vv < 1.0e-14f) { return bounded ? 0 : -1; }
        if (bounded && vv <= limit * limit) { return 0; }
        GVertex n;
        n.a = support(a, xa, v);
        n.b = support(b, xb, neg(v));
        n.w = n.b - n.a;
        const float vw = dot(v, n.w);
        if (bounded && vw > 0.0f && vw * vw > limit * limit * vv) { return 1; }
        if (vv - vw <= 1.0e-6f * vv) { break; }
        bool repeated = false;
        for (int i = 0; i < s.count; ++i) { repeated = repeated || lengthSq(s.v[i].w - n.w) < 1.0e-14f; }
        if (repeated) { break; }

        const Simplex previous = s;
        s.v[s.count++] = n;
        if (!solveSimplex(s)) { return bounded ? 0 : -1; }
        const V3 next = closestOf(s);
        if (dot(next, next) >= vv) {
            s = previous;
            break;
        }
        v = next;
    }
    return -1;
}

// apart (src/distance.cpp:266-274) for two boxes (no radius): the gap between them exceeds limit.
TF_HD bool apartBoxes(const BoxPoly& a, const Xf& xa, const BoxPoly& b, const Xf& xb, float limit) {
    V3 v;
    const int verdict = searchBoxes(a, xa, b, xb, limit + 0.0f + 0.0f, v);
    if (verdict >= 0) { return verdict == 1; }
    return length(v) - 0.0f - 0.0f > limit;
}

// collideConvex for two boxes (src/collide3d.cpp:615-629): queries in closed form, the GJK gap test when
// they are unsure, then the contact along the deciding axis.
TF_HD void collideBoxes(V3 halfA, const BoxPoly& a, const Xf& xa, V3 halfB, const BoxPoly& b, const Xf& xb,
                        float margin, float linearSlop, const Motion& motion, Manifold& m) {
    const Xf x{ invRotate(xa.q, xb.p - xa.p), tf::operator*(tf::conjugate(xa.q), xb.q) };
    FaceQuery fa{ -1, -FLT_MAX }, fb{ -1, -FLT_MAX };
    EdgeQuery eq{ -1, -1, -FLT_MAX, vec(0.0f, 0.0f, 0.0f) };
    bool unsure = false;
    if (!boxQueries(halfA, halfB, x, margin, fa, fb, eq, unsure)) { return; }
    if (unsure && apartBoxes(a, xa, b, xb, margin)) { return; }
    BoxPoly bf;
    otherFrame(b, x, bf);
    convexContact(a, bf, fa, fb, eq, xa, margin, linearSlop, motion, m);
}

// A shape as the pair collision reads it: kind, box half-extents (after the reference's solidBox),
// sphere radius (after solidRadius).
struct Solid {
    int kind;
    V3 half;
    float radius;
};

// collide (src/collide3d.cpp:1778-1811) for spheres, boxes and half-spaces; xa and xb are the shapes'
// world transforms. Returns false for any other kind.
TF_HD bool collide(const Solid& a, const Xf& xa, const Solid& b, const Xf& xb, float margin, float linearSlop,
                   const Motion& motion, Manifold& m) {
    m.count = 0;
    m.normal = vec(0.0f, 0.0f, 0.0f);
    const int ka = a.kind, kb = b.kind;
    const bool known = (ka == kSphere || ka == kBox || ka == kHalfSpace) && (kb == kSphere || kb == kBox || kb == kHalfSpace);
    if (!known) { return false; }
    if (ka == kHalfSpace && kb == kHalfSpace) { return true; }
    BoxPoly pa, pb;
    if (ka == kBox) { makeBoxPoly(a.half, pa); }
    if (kb == kBox) { makeBoxPoly(b.half, pb); }
    if (ka == kHalfSpace) {
        collideHalfSpace(xa, kb, pb, b.radius, xb, margin, linearSlop, motion, m);
    } else if (kb == kHalfSpace) {
        collideHalfSpace(xb, ka, pa, a.radius, xa, margin, linearSlop, swapped(motion), m);
        m.normal = neg(m.normal);
    } else if (ka == kSphere && kb == kSphere) {
        collideSpheres(xa.p, a.radius, xb.p, b.radius, margin, m);
    } else if (ka == kSphere) {
        collideSphereBox(pb, xb, xa.p, a.radius, margin, m);
        m.normal = neg(m.normal);
    } else if (kb == kSphere) {
        collideSphereBox(pa, xa, xb.p, b.radius, margin, m);
    } else {
        collideBoxes(a.half, pa, xa, b.half, pb, xb, margin, linearSlop, motion, m);
    }
    return true;
}

// ---- a pair of bodies with one shape each ------------------------------------------------------------

TF_HD bool planarBody(const tfs_body& b) { return (b.locks & TFS_LOCK_PLANAR) == TFS_LOCK_PLANAR; }

// usePlanar (src/world.cpp:737-750): every dynamic body of the pair is locked to a plane, the planes
// aligned, and there is one. The reference collides such a pair in 2D (collidePlanar), which this file
// does not restate.
TF_HD bool usePlanar(const tfs_body& a, const tfs_body& b) {
    const bool da = a.type == TFS_DYNAMIC, db = b.type == TFS_DYNAMIC;
    if ((da && !planarBody(a)) || (db && !planarBody(b))) { return false; }
    if (da && db) {
        const V3 na = rotate(tf::ld4(a.lock_frame), vec(0.0f, 0.0f, 1.0f));
        const V3 nb = rotate(tf::ld4(b.lock_frame), vec(0.0f, 0.0f, 1.0f));
        return !(fabsf(dot(na, nb)) < kAligned);
    }
    return da || db;
}

// The narrowphase of collidePair (src/world.cpp:504-535) for two bodies with one shape each at the body's
// origin (shape transform = body pose x identity), over time dt: the margin grown by the pair's
// relativeTravel (484-486), Motion from the bodies' velocities and centres of mass. 1 with a manifold, 0
// without, -1 for a pair this file does not collide (another kind of shape, or a planar pair).
TF_HD int pairManifold(const tfs_body& a, V3 va, V3 wa, const Solid& sa, const tfs_body& b, V3 vb, V3 wb,
                       const Solid& sb, float dt, float speculative, float linearSlop, Manifold& m) {
    m.count = 0;
    const float travel = (length(vb - va) + length(wa) * a.extent + length(wb) * b.extent) * dt;
    const float margin = speculative + travel;
    if (usePlanar(a, b)) { return -1; }
    const Xf local{ vec(0.0f, 0.0f, 0.0f), quat(0.0f, 0.0f, 0.0f, 1.0f) };
    const Xf xa = compose(Xf{ tf::ld3(a.p), tf::ld4(a.q) }, local);
    const Xf xb = compose(Xf{ tf::ld3(b.p), tf::ld4(b.q) }, local);
    const Motion mo{ va, wa, tf::ld3(a.com), vb, wb, tf::ld3(b.com), dt, travel > speculative };
    if (!collide(sa, xa, sb, xb, margin, linearSlop, mo, m)) { return -1; }
    return m.count > 0 ? 1 : 0;
}

// The geometry of a contact's new points from a manifold, as the refresh makes them (src/world.cpp:
// 1111-1121; the same as a new contact's, 927-943): normal, tangents (perpendicular, 930-931), count, and
// per point local_a, separation and the anchors, lifted onto each surface for points further apart than
// the speculative distance unless a body is planar (lift, 536-540). Other fields are left alone.
TF_HD void fillContact(const tfs_body& a, const tfs_body& b, V3 n, int count, const V3* position,
                       const float* separation, float speculative, tfs_contact& out) {
    const V3 t1 = perpendicularOf(n);
    tf::st3(out.normal, n);
    tf::st3(out.tangent1, t1);
    tf::st3(out.tangent2, cross(n, t1));
    out.count = count;
    const bool flat = planarBody(a) || planarBody(b);
    const Xf pa{ tf::ld3(a.p), tf::ld4(a.q) };
    for (int i = 0; i < 4; ++i) {
        if (i >= count) { continue; }
        tfs_point& p = out.points[i];
        const V3 pos = position[i];
        tf::st3(p.local_a, applyInverse(pa, pos));
        p.separation = separation[i];
        const V3 up = !flat && p.separation > speculative ? n * (0.5f * p.separation) : vec(0.0f, 0.0f, 0.0f);
        tf::st3(p.anchor_a, pos - up - tf::ld3(a.com));
        tf::st3(p.anchor_b, pos + up - tf::ld3(b.com));
    }
}

}  // namespace tfc
