// scene.cpp - a host world that drives the solver end to end: dynamic and static boxes, spheres,
// cylinders and prisms (convex polygons extruded along local z) and static half-spaces, a
// uniform-grid broadphase and a narrowphase that hands the solver contact manifolds in its input form
// (include/twofold_solver.h). Cylinders and prisms collide with half-spaces, spheres, and with other
// shapes extruded along the world z when both bodies stay so (locked to the xy plane, or static): the
// last as a 2D problem in the xy plane over the shared z range.
//
// It stands in for the reference's broadphase and narrowphase (src/world.cpp:752-993,
// src/collide3d.cpp), which sit outside the solve path, so that the solver can be run on the GPU
// box on the piles it is measured on. Pairs of spheres, boxes and half-spaces go through the
// reference's own narrowphase, restated in collide3d.h (the source the device refresh runs too), and
// their contacts take the reference's geometry (src/world.cpp:927-943, 1111-1121): they are the
// reference's manifolds bit for bit. The other pairs (cylinders, prisms, and pairs the reference
// collides in 2D) follow its conventions - normal from shape A to shape B, shapes ordered by index,
// points at the midpoint between the two surfaces, a speculative margin grown by the pair's travel
// (src/world.cpp:484-486, 509-511), anchors lifted onto each surface for separated points
// (src/world.cpp:536-540) - but not its bits: narrow.h, and a 2D version of it for z-extruded shapes.
// Impulses carry to the nearest new point (src/world.cpp:102-120), and friction and restitution mix as
// src/world.cpp:39-40.
#include <omp.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <unordered_map>
#include <vector>

#include "../../include/twofold_solver.h"
#include "collide3d.h"
#include "narrow.h"

namespace {

using namespace tfn;

constexpr int kMaxPrismVertices = 16;

struct Shape {
    int kind = kBox;
    V half{ 0.5f, 0.5f, 0.5f };   // a box's half-extents; the local bounds of a cylinder or prism
    float radius = 0.5f;           // sphere, cylinder (whose axis is local z, from -half.z to half.z)
    float zMin = -0.5f, zMax = 0.5f;
    std::vector<float> poly;       // prism: x0, y0, x1, y1, ... counter-clockwise in the local xy plane
    float density = 1.0f, friction = 0.5f, restitution = 0.0f, rolling = 0.0f;
};

Pose poseOf(const tfs_body& b) { return tfn::poseOf(b.com, b.q); }

// Half-space A (below its local xy plane) against a sphere, box, prism or cylinder B.
void halfSpaceVs(const Pose& a, const Shape& sb, const Pose& b, float margin, Manifold& m) {
    if (sb.kind == kSphere) {
        halfSpaceSphere(a, b.c, sb.radius, margin, m);
        return;
    }
    const V n = a.ax[2];
    V verts[kMaxHalfSpacePoints];
    int nv = 0;
    if (sb.kind == kBox) {
        V box[8];
        boxVertices(b, sb.half, box);
        for (const V& v : box) { verts[nv++] = v; }
    } else if (sb.kind == kPrism) {
        for (size_t i = 0; i + 1 < sb.poly.size(); i += 2) {
            const V xy = b.c + b.ax[0] * sb.poly[i] + b.ax[1] * sb.poly[i + 1];
            verts[nv++] = xy + b.ax[2] * sb.zMin;
            verts[nv++] = xy + b.ax[2] * sb.zMax;
        }
    } else if (sb.kind == kCylinder) {
        // each cap rim's lowest point along n, or four rim points of a cap lying flat on the plane
        for (float side : { -1.0f, 1.0f }) {
            const V cap = b.c + b.ax[2] * (side * sb.half.z);
            const V down = -n - b.ax[2] * dot(-n, b.ax[2]);
            const float dl = len(down);
            if (dl > 1e-4f) {
                verts[nv++] = cap + down * (sb.radius / dl);
            } else {
                for (int k = 0; k < 4; ++k) { verts[nv++] = cap + b.ax[k & 1] * ((k & 2) ? -sb.radius : sb.radius); }
            }
        }
    }
    halfSpacePoints(a, verts, nv, margin, m);
}

// ---- z-extruded shapes as a 2D problem ----------------------------------------------------------

// A shape's cross-section in the world xy plane and its z range, for a shape extruded along the
// world z: a circle or a counter-clockwise convex polygon.
struct Profile {
    bool circle = false;
    float cx = 0, cy = 0, r = 0;
    int n = 0;
    float x[kMaxPrismVertices], y[kMaxPrismVertices];
    float zlo = 0, zhi = 0;
};

bool profileOf(const Shape& s, const Pose& p, Profile& f) {
    const float up = p.ax[2].z;
    if (std::fabs(up) < 0.9999f) { return false; }
    if (s.kind == kCylinder) {
        f.circle = true;
        f.cx = p.c.x;
        f.cy = p.c.y;
        f.r = s.radius;
        f.zlo = p.c.z - s.half.z;
        f.zhi = p.c.z + s.half.z;
        return true;
    }
    float lx[kMaxPrismVertices], ly[kMaxPrismVertices], z0, z1;
    int n;
    if (s.kind == kBox) {
        n = 4;
        const float hx = s.half.x, hy = s.half.y;
        const float cx[4] = { hx, -hx, -hx, hx }, cy[4] = { hy, hy, -hy, -hy };
        for (int i = 0; i < 4; ++i) {
            lx[i] = cx[i];
            ly[i] = cy[i];
        }
        z0 = -s.half.z;
        z1 = s.half.z;
    } else if (s.kind == kPrism) {
        n = (int)s.poly.size() / 2;
        for (int i = 0; i < n; ++i) {
            lx[i] = s.poly[2 * i];
            ly[i] = s.poly[2 * i + 1];
        }
        z0 = s.zMin;
        z1 = s.zMax;
    } else {
        return false;
    }
    f.circle = false;
    f.n = n;
    for (int i = 0; i < n; ++i) {
        const V w = p.c + p.ax[0] * lx[i] + p.ax[1] * ly[i];
        const int k = up > 0 ? i : n - 1 - i;  // a shape turned over mirrors its winding
        f.x[k] = w.x;
        f.y[k] = w.y;
    }
    f.zlo = up > 0 ? p.c.z + z0 : p.c.z - z1;
    f.zhi = up > 0 ? p.c.z + z1 : p.c.z - z0;
    return true;
}

// A 2D manifold: normal from A to B, up to two points (midpoints) with their separations.
struct Manifold2 {
    float nx = 0, ny = 0;
    int count = 0;
    float px[2], py[2], sep[2];
};

void edgeNormal(const Profile& f, int i, float& nx, float& ny) {
    const int j = (i + 1) % f.n;
    const float ex = f.x[j] - f.x[i], ey = f.y[j] - f.y[i];
    const float l = std::sqrt(ex * ex + ey * ey);
    nx = l > 0 ? ey / l : 0.0f;   // outward for counter-clockwise order
    ny = l > 0 ? -ex / l : 0.0f;
}

// The largest separation of Q from P's edges, and that edge.
float edgeSeparation(const Profile& P, const Profile& Q, int& best) {
    float most = -1e30f;
    best = 0;
    for (int i = 0; i < P.n; ++i) {
        float nx, ny;
        edgeNormal(P, i, nx, ny);
        float least = 1e30f;
        for (int j = 0; j < Q.n; ++j) { least = std::min(least, nx * (Q.x[j] - P.x[i]) + ny * (Q.y[j] - P.y[i])); }
        if (least > most) {
            most = least;
            best = i;
        }
    }
    return most;
}

void polygons2(const Profile& A, const Profile& B, float margin, Manifold2& m) {
    m.count = 0;
    int ea, eb;
    const float sa = edgeSeparation(A, B, ea);
    if (sa > margin) { return; }
    const float sb = edgeSeparation(B, A, eb);
    if (sb > margin) { return; }
    const bool refB = sb > 0.98f * sa + 0.0005f;
    const Profile& R = refB ? B : A;
    const Profile& I = refB ? A : B;
    const int k = refB ? eb : ea;
    float nx, ny;
    edgeNormal(R, k, nx, ny);
    int j = 0;
    float most = 1e30f;
    for (int i = 0; i < I.n; ++i) {
        float ix, iy;
        edgeNormal(I, i, ix, iy);
        const float d = ix * nx + iy * ny;
        if (d < most) {
            most = d;
            j = i;
        }
    }
    float sx[2] = { I.x[j], I.x[(j + 1) % I.n] }, sy[2] = { I.y[j], I.y[(j + 1) % I.n] };
    const float r0x = R.x[k], r0y = R.y[k], r1x = R.x[(k + 1) % R.n], r1y = R.y[(k + 1) % R.n];
    const float tx = -ny, ty = nx;  // along the reference edge, r0 -> r1
    // clip the incident edge to the reference edge's extent
    const float lo = tx * r0x + ty * r0y, hi = tx * r1x + ty * r1y;
    for (int side = 0; side < 2; ++side) {
        const float limit = side ? hi : lo;
        const float dir = side ? 1.0f : -1.0f;  // keep dir * (t . p - limit) <= 0
        const float d0 = dir * (tx * sx[0] + ty * sy[0] - limit), d1 = dir * (tx * sx[1] + ty * sy[1] - limit);
        if (d0 > 0 && d1 > 0) { return; }
        if (d0 > 0 || d1 > 0) {
            const float u = d0 / (d0 - d1);
            const float qx = sx[0] + (sx[1] - sx[0]) * u, qy = sy[0] + (sy[1] - sy[0]) * u;
            const int out = d0 > 0 ? 0 : 1;
            sx[out] = qx;
            sy[out] = qy;
        }
    }
    for (int i = 0; i < 2; ++i) {
        const float sep = nx * (sx[i] - r0x) + ny * (sy[i] - r0y);
        if (sep > margin) { continue; }
        m.px[m.count] = sx[i] - nx * (0.5f * sep);
        m.py[m.count] = sy[i] - ny * (0.5f * sep);
        m.sep[m.count] = sep;
        ++m.count;
    }
    m.nx = refB ? -nx : nx;
    m.ny = refB ? -ny : ny;
}

// Circle C against polygon P; the normal points from P to C.
void circlePolygon2(const Profile& C, const Profile& P, float margin, Manifold2& m) {
    m.count = 0;
    float best = -1e30f, bnx = 0, bny = 1;
    bool outside = false;
    for (int i = 0; i < P.n; ++i) {
        float nx, ny;
        edgeNormal(P, i, nx, ny);
        const float s = nx * (C.cx - P.x[i]) + ny * (C.cy - P.y[i]);
        if (s > 0) { outside = true; }
        if (s > best) {
            best = s;
            bnx = nx;
            bny = ny;
        }
    }
    float qx, qy, nx, ny, sep;
    if (!outside) {
        nx = bnx;
        ny = bny;
        sep = best - C.r;
        qx = C.cx - nx * best;
        qy = C.cy - ny * best;
    } else {
        float dsq = 1e30f;
        qx = P.x[0];
        qy = P.y[0];
        for (int i = 0; i < P.n; ++i) {
            const int j = (i + 1) % P.n;
            const float ex = P.x[j] - P.x[i], ey = P.y[j] - P.y[i];
            const float ee = ex * ex + ey * ey;
            float t = ee > 0 ? ((C.cx - P.x[i]) * ex + (C.cy - P.y[i]) * ey) / ee : 0.0f;
            t = std::max(0.0f, std::min(1.0f, t));
            const float x = P.x[i] + ex * t, y = P.y[i] + ey * t;
            const float d = (C.cx - x) * (C.cx - x) + (C.cy - y) * (C.cy - y);
            if (d < dsq) {
                dsq = d;
                qx = x;
                qy = y;
            }
        }
        const float dist = std::sqrt(dsq);
        if (dist <= 0) { return; }
        nx = (C.cx - qx) / dist;
        ny = (C.cy - qy) / dist;
        sep = dist - C.r;
    }
    if (sep > margin) { return; }
    m.nx = nx;
    m.ny = ny;
    m.px[0] = 0.5f * (qx + C.cx - nx * C.r);
    m.py[0] = 0.5f * (qy + C.cy - ny * C.r);
    m.sep[0] = sep;
    m.count = 1;
}

void circles2(const Profile& A, const Profile& B, float margin, Manifold2& m) {
    m.count = 0;
    const float dx = B.cx - A.cx, dy = B.cy - A.cy;
    const float dist = std::sqrt(dx * dx + dy * dy);
    const float sep = dist - A.r - B.r;
    if (sep > margin) { return; }
    m.nx = dist > 1e-9f ? dx / dist : 0.0f;
    m.ny = dist > 1e-9f ? dy / dist : 1.0f;
    m.px[0] = A.cx + m.nx * (A.r + 0.5f * sep);
    m.py[0] = A.cy + m.ny * (A.r + 0.5f * sep);
    m.sep[0] = sep;
    m.count = 1;
}

// Two z-extruded shapes: the 2D contact of their cross-sections at both ends of their shared z range.
void extruded(const Profile& A, const Profile& B, float margin, Manifold& m) {
    m.count = 0;
    const float zlo = std::max(A.zlo, B.zlo), zhi = std::min(A.zhi, B.zhi);
    if (zlo > zhi) { return; }
    Manifold2 q;
    if (A.circle && B.circle) {
        circles2(A, B, margin, q);
    } else if (A.circle) {
        circlePolygon2(A, B, margin, q);
        q.nx = -q.nx;
        q.ny = -q.ny;
    } else if (B.circle) {
        circlePolygon2(B, A, margin, q);
    } else {
        polygons2(A, B, margin, q);
    }
    m.n = V{ q.nx, q.ny, 0.0f };
    const bool flat = zhi - zlo < 1e-6f;
    for (int i = 0; i < q.count; ++i) {
        m.p[m.count++] = { V{ q.px[i], q.py[i], zlo }, q.sep[i] };
        if (!flat) { m.p[m.count++] = { V{ q.px[i], q.py[i], zhi }, q.sep[i] }; }
    }
}

// A prism or cylinder's point nearest a sphere's centre (normal from the shape to the sphere).
void roundOrPrismSphere(const Shape& s, const Pose& a, V cs, float r, float margin, Manifold& m) {
    m.count = 0;
    const V d = cs - a.c;
    const V l{ dot(d, a.ax[0]), dot(d, a.ax[1]), dot(d, a.ax[2]) };
    const float z0 = s.kind == kCylinder ? -s.half.z : s.zMin, z1 = s.kind == kCylinder ? s.half.z : s.zMax;
    V q = l, nl{ 0, 0, 1 };
    float sep;
    bool inside2 = true;
    float deepest = -1e30f, dnx = 0, dny = 0;  // the least deep side, when inside
    if (s.kind == kCylinder) {
        const float rho = std::sqrt(l.x * l.x + l.y * l.y);
        if (rho > s.radius) {
            inside2 = false;
            q.x = l.x * (s.radius / rho);
            q.y = l.y * (s.radius / rho);
        }
        deepest = rho - s.radius;
        dnx = rho > 1e-9f ? l.x / rho : 1.0f;
        dny = rho > 1e-9f ? l.y / rho : 0.0f;
    } else {
        const int n = (int)s.poly.size() / 2;
        for (int i = 0; i < n; ++i) {
            const int j = (i + 1) % n;
            const float ex = s.poly[2 * j] - s.poly[2 * i], ey = s.poly[2 * j + 1] - s.poly[2 * i + 1];
            const float el = std::sqrt(ex * ex + ey * ey);
            const float nx = el > 0 ? ey / el : 0.0f, ny = el > 0 ? -ex / el : 0.0f;
            const float sd = nx * (l.x - s.poly[2 * i]) + ny * (l.y - s.poly[2 * i + 1]);
            if (sd > 0) { inside2 = false; }
            if (sd > deepest) {
                deepest = sd;
                dnx = nx;
                dny = ny;
            }
        }
        if (!inside2) {
            float dsq = 1e30f;
            for (int i = 0; i < n; ++i) {
                const int j = (i + 1) % n;
                const float ax = s.poly[2 * i], ay = s.poly[2 * i + 1];
                const float ex = s.poly[2 * j] - ax, ey = s.poly[2 * j + 1] - ay;
                const float ee = ex * ex + ey * ey;
                float t = ee > 0 ? ((l.x - ax) * ex + (l.y - ay) * ey) / ee : 0.0f;
                t = std::max(0.0f, std::min(1.0f, t));
                const float x = ax + ex * t, y = ay + ey * t;
                const float dd = (l.x - x) * (l.x - x) + (l.y - y) * (l.y - y);
                if (dd < dsq) {
                    dsq = dd;
                    q.x = x;
                    q.y = y;
                }
            }
        }
    }
    q.z = std::max(z0, std::min(z1, l.z));
    const bool insideZ = l.z >= z0 && l.z <= z1;
    if (inside2 && insideZ) {
        // centre inside: out through the nearest side or cap
        float best = deepest;
        nl = V{ dnx, dny, 0.0f };
        if (l.z - z1 > best) {
            best = l.z - z1;
            nl = V{ 0, 0, 1 };
        }
        if (z0 - l.z > best) {
            best = z0 - l.z;
            nl = V{ 0, 0, -1 };
        }
        sep = best - r;
        q = l - nl * best;
    } else {
        const V off = l - q;
        const float dist = len(off);
        if (dist <= 1e-9f) { return; }
        nl = off * (1.0f / dist);
        sep = dist - r;
    }
    if (sep > margin) { return; }
    const V n = a.ax[0] * nl.x + a.ax[1] * nl.y + a.ax[2] * nl.z;
    const V pa = a.c + a.ax[0] * q.x + a.ax[1] * q.y + a.ax[2] * q.z;
    m.n = n;
    m.p[0] = { (pa + (cs - n * r)) * 0.5f, sep };
    m.count = 1;
}

struct World {
    tfs_settings settings{};
    int threads = 1;  // host threads of the narrowphase loops
    // Collision groups: bodies of different groups never collide (independent worlds in one set).
    std::vector<uint32_t> group;
    uint32_t groups = 1;
    std::vector<tfs_body> bodies;
    std::vector<Shape> shapes;  // one per body, at the body's origin
    std::vector<uint32_t> planes, statics, movers;
};

float travelOf(const tfs_body& b, float dt) {
    if (b.type == TFS_STATIC) { return 0.0f; }
    return (len(ld(b.v)) + len(ld(b.w)) * b.extent) * dt;
}

bool basicKind(int kind) { return kind == kSphere || kind == kBox || kind == kHalfSpace; }
tfc::Solid solidOf(const Shape& s) { return tfc::Solid{ s.kind, tf::vec(s.half.x, s.half.y, s.half.z), s.radius }; }

// One manifold for the pair (a < b), or none. Spheres, boxes and half-spaces go through the reference's
// narrowphase (collide3d.h, the source the device refresh runs), except pairs the reference collides in
// 2D, which take the host world's own.
bool collide(const World& w, uint32_t ia, const tfs_body& a, uint32_t ib, const tfs_body& b, float dt, Manifold& m) {
    const Shape& sa = w.shapes[ia];
    const Shape& sb = w.shapes[ib];
    if (basicKind(sa.kind) && basicKind(sb.kind)) {
        tfc::Manifold r;
        const int got = tfc::pairManifold(a, tf::ld3(a.v), tf::ld3(a.w), solidOf(sa), b, tf::ld3(b.v), tf::ld3(b.w),
                                          solidOf(sb), dt, w.settings.speculative, 0.005f * w.settings.length_unit, r);
        if (got >= 0) {
            m.reference = true;
            m.n = vec3(r.normal.x, r.normal.y, r.normal.z);
            m.count = got > 0 ? r.count : 0;
            for (int i = 0; i < m.count; ++i) {
                m.p[i].pos = vec3(r.points[i].position.x, r.points[i].position.y, r.points[i].position.z);
                m.p[i].sep = r.points[i].separation;
            }
            return got > 0;
        }
    }
    const float margin = pairMargin(ld(a.v), ld(a.w), a.extent, ld(b.v), ld(b.w), b.extent, dt, w.settings.speculative);
    const Pose pa = poseOf(a), pb = poseOf(b);
    const bool bothInPlane = inPlane(a) && inPlane(b);
    // spheres, boxes and half-spaces as the device refresh collides them (narrow.h)
    const int basic = collideBasic(sa.kind, sa.half, sa.radius, pa, sb.kind, sb.half, sb.radius, pb, bothInPlane, margin, m);
    if (basic >= 0) { return basic > 0; }
    m.count = 0;
    if (sa.kind == kHalfSpace && sb.kind != kHalfSpace) {
        halfSpaceVs(pa, sb, pb, margin, m);
    } else if (sb.kind == kHalfSpace) {
        return false;
    } else if (sa.kind == kSphere && sb.kind == kSphere) {
        sphereSphere(pa.c, sa.radius, pb.c, sb.radius, margin, m);
    } else if (sb.kind == kSphere) {
        if (sa.kind == kBox) {
            boxSphere(sa.half, pa, pb.c, sb.radius, margin, m);
        } else {
            roundOrPrismSphere(sa, pa, pb.c, sb.radius, margin, m);
        }
    } else if (sa.kind == kSphere) {
        if (sb.kind == kBox) {
            boxSphere(sb.half, pb, pa.c, sa.radius, margin, m);
        } else {
            roundOrPrismSphere(sb, pb, pa.c, sa.radius, margin, m);
        }
        m.n = -m.n;
    } else {
        Profile fa, fb;
        if (bothInPlane && profileOf(sa, pa, fa) && profileOf(sb, pb, fb)) {
            extruded(fa, fb, margin, m);
        } else if (sa.kind == kBox && sb.kind == kBox) {
            boxBox(sa.half, pa, sb.half, pb, margin, m);
        }  // a cylinder or prism against a shape not extruded along the world z: not supported
    }
    return m.count > 0;
}

// The contact of pair (ia, ib) from a manifold (src/world.cpp:920-950 conventions).
void toContact(const World& w, uint32_t ia, const tfs_body& a, uint32_t ib, const tfs_body& b, const Manifold& m,
               tfs_contact& c) {
    std::memset(&c, 0, sizeof c);
    c.key = ((uint64_t)ia << 32) | ib;
    c.body_a = ia;
    c.body_b = ib;
    const Shape& sa = w.shapes[ia];
    const Shape& sb = w.shapes[ib];
    c.friction = std::sqrt(sa.friction * sb.friction);
    c.restitution = std::max(sa.restitution, sb.restitution);
    c.rolling_resistance = std::max(sa.rolling, sb.rolling);
    if (!m.reference) {
        contactGeometry(a, b, m, w.settings.speculative, c);
        return;
    }
    tf::V3 pos[4];
    float sep[4];
    for (int i = 0; i < 4; ++i) {
        pos[i] = tf::vec(m.p[i].pos.x, m.p[i].pos.y, m.p[i].pos.z);
        sep[i] = m.p[i].sep;
    }
    tfc::fillContact(a, b, tf::vec(m.n.x, m.n.y, m.n.z), m.count, pos, sep, w.settings.speculative, c);
}

// carryImpulses: src/world.cpp:102-120.
void carry(const tfs_contact& old, tfs_contact& c, float matchSq) {
    int from[4], claims[4] = {};
    for (int i = 0; i < c.count; ++i) {
        from[i] = -1;
        float bestSq = matchSq;
        for (int k = 0; k < old.count; ++k) {
            const V d = ld(old.points[k].local_a) - ld(c.points[i].local_a);
            if (dot(d, d) <= bestSq) {
                from[i] = k;
                bestSq = dot(d, d);
            }
        }
        if (from[i] >= 0) { ++claims[from[i]]; }
    }
    const V n = ld(c.normal);
    for (int i = 0; i < c.count; ++i) {
        if (from[i] < 0) { continue; }
        const tfs_point& o = old.points[from[i]];
        const float share = 1.0f / (float)claims[from[i]];
        c.points[i].normal_impulse = o.normal_impulse * share;
        const V t = ld(o.tangent_impulse);
        st(c.points[i].tangent_impulse, (t - n * dot(n, t)) * share);
    }
    std::memcpy(c.rolling_impulse, old.rolling_impulse, sizeof c.rolling_impulse);
}

void aabb(const World& w, const tfs_body& b, uint32_t i, float grow, V& lo, V& hi) {
    const Shape& s = w.shapes[i];
    const Pose p = poseOf(b);
    V e;
    if (s.kind == kSphere) {
        e = V{ s.radius, s.radius, s.radius };
    } else {
        e = V{ std::fabs(p.ax[0].x) * s.half.x + std::fabs(p.ax[1].x) * s.half.y + std::fabs(p.ax[2].x) * s.half.z,
               std::fabs(p.ax[0].y) * s.half.x + std::fabs(p.ax[1].y) * s.half.y + std::fabs(p.ax[2].y) * s.half.z,
               std::fabs(p.ax[0].z) * s.half.x + std::fabs(p.ax[1].z) * s.half.y + std::fabs(p.ax[2].z) * s.half.z };
    }
    e = e + V{ grow, grow, grow };
    lo = p.c - e;
    hi = p.c + e;
}

bool overlap(V alo, V ahi, V blo, V bhi) {
    return alo.x <= bhi.x && blo.x <= ahi.x && alo.y <= bhi.y && blo.y <= ahi.y && alo.z <= bhi.z && blo.z <= ahi.z;
}

bool moving(const tfs_body& b) { return b.type == TFS_DYNAMIC && b.awake; }

// Candidate pairs (a < b) whose grown bounds overlap, sorted, with at least one awake dynamic body.
std::vector<uint64_t> candidates(const World& w, const tfs_body* bodies, float dt) {
    const uint32_t nb = (uint32_t)w.bodies.size();
    std::vector<V> lo(nb), hi(nb);
    float cell = 0.0f;
    for (uint32_t i : w.movers) {
        const float g = 0.5f * w.settings.speculative + travelOf(bodies[i], dt);
        aabb(w, bodies[i], i, g, lo[i], hi[i]);
        cell = std::max(cell, std::max(hi[i].x - lo[i].x, std::max(hi[i].y - lo[i].y, hi[i].z - lo[i].z)));
    }
    std::vector<uint64_t> out;
    if (cell <= 0.0f) { return out; }
    const float inv = 1.0f / cell;
    std::unordered_map<uint64_t, std::vector<uint32_t>> grid;
    grid.reserve(w.movers.size() * 2);
    const bool grouped = w.groups > 1;
    // A cell of one group; with several groups the cell key is mixed with the group (a rare clash only
    // costs extra candidates, which the group test below drops).
    auto key = [&](uint32_t body, int x, int y, int z) {
        const uint64_t k = ((uint64_t)(uint32_t)(x + (1 << 20)) << 42) | ((uint64_t)(uint32_t)(y + (1 << 20)) << 21) |
                           (uint64_t)(uint32_t)(z + (1 << 20));
        return grouped ? (k * 0x9E3779B97F4A7C15ull) ^ w.group[body] : k;
    };
    for (uint32_t i : w.movers) {
        const int x = (int)std::floor(lo[i].x * inv), y = (int)std::floor(lo[i].y * inv), z = (int)std::floor(lo[i].z * inv);
        grid[key(i, x, y, z)].push_back(i);
    }
    for (uint32_t i : w.movers) {
        const int x0 = (int)std::floor(lo[i].x * inv), y0 = (int)std::floor(lo[i].y * inv), z0 = (int)std::floor(lo[i].z * inv);
        for (int dx = -1; dx <= 1; ++dx) {
            for (int dy = -1; dy <= 1; ++dy) {
                for (int dz = -1; dz <= 1; ++dz) {
                    auto it = grid.find(key(i, x0 + dx, y0 + dy, z0 + dz));
                    if (it == grid.end()) { continue; }
                    for (uint32_t j : it->second) {
                        if (j <= i || (grouped && w.group[j] != w.group[i])) { continue; }
                        if (!moving(bodies[i]) && !moving(bodies[j])) { continue; }
                        if (overlap(lo[i], hi[i], lo[j], hi[j])) { out.push_back(((uint64_t)i << 32) | j); }
                    }
                }
            }
        }
    }
    // movers by group, for the static bodies and planes of each group
    std::vector<std::vector<uint32_t>> moversOf(w.groups);
    for (uint32_t i : w.movers) { moversOf[grouped ? w.group[i] : 0].push_back(i); }
    for (uint32_t s : w.statics) {
        V slo, shi;
        aabb(w, bodies[s], s, 0.5f * w.settings.speculative, slo, shi);
        for (uint32_t i : moversOf[grouped ? w.group[s] : 0]) {
            if (!moving(bodies[i])) { continue; }
            if (overlap(slo, shi, lo[i], hi[i])) { out.push_back(((uint64_t)std::min(s, i) << 32) | std::max(s, i)); }
        }
    }
    for (uint32_t p : w.planes) {
        const Pose pp = poseOf(bodies[p]);
        const V n = pp.ax[2];
        const float off = dot(n, pp.c);
        for (uint32_t i : moversOf[grouped ? w.group[p] : 0]) {
            if (!moving(bodies[i])) { continue; }
            const V c = (lo[i] + hi[i]) * 0.5f, e = (hi[i] - lo[i]) * 0.5f;
            const float reach = e.x * std::fabs(n.x) + e.y * std::fabs(n.y) + e.z * std::fabs(n.z);
            if (dot(c, n) - off <= reach + 0.5f * w.settings.speculative) {
                out.push_back(((uint64_t)std::min(p, i) << 32) | std::max(p, i));
            }
        }
    }
    std::sort(out.begin(), out.end());
    out.erase(std::unique(out.begin(), out.end()), out.end());
    return out;
}

void massProperties(tfs_body& b, const Shape& s) {
    float m = 0.0f;
    float I[3] = { 0, 0, 0 };
    if (s.kind == kBox) {
        m = s.density * 8.0f * s.half.x * s.half.y * s.half.z;
        const float x2 = s.half.x * s.half.x, y2 = s.half.y * s.half.y, z2 = s.half.z * s.half.z;
        I[0] = m * (y2 + z2) / 3.0f;
        I[1] = m * (x2 + z2) / 3.0f;
        I[2] = m * (x2 + y2) / 3.0f;
        b.extent = len(s.half);
    } else if (s.kind == kSphere) {
        m = s.density * 4.0f / 3.0f * 3.14159265f * s.radius * s.radius * s.radius;
        I[0] = I[1] = I[2] = 0.4f * m * s.radius * s.radius;
        b.extent = s.radius;
    } else if (s.kind == kCylinder) {
        const float r2 = s.radius * s.radius, h = 2.0f * s.half.z;
        m = s.density * 3.14159265f * r2 * h;
        I[0] = I[1] = m * (3.0f * r2 + h * h) / 12.0f;
        I[2] = 0.5f * m * r2;
        b.extent = std::sqrt(r2 + s.half.z * s.half.z);
    } else if (s.kind == kPrism) {
        // polygon area and second moments (shoelace sums), about the local origin (the centroid)
        const int n = (int)s.poly.size() / 2;
        double area = 0, ixx = 0, iyy = 0, far = 0;
        for (int i = 0; i < n; ++i) {
            const int j = (i + 1) % n;
            const double x0 = s.poly[2 * i], y0 = s.poly[2 * i + 1], x1 = s.poly[2 * j], y1 = s.poly[2 * j + 1];
            const double c = x0 * y1 - x1 * y0;
            area += 0.5 * c;
            iyy += c * (x0 * x0 + x0 * x1 + x1 * x1) / 12.0;   // integral of x^2
            ixx += c * (y0 * y0 + y0 * y1 + y1 * y1) / 12.0;   // integral of y^2
            far = std::max(far, x0 * x0 + y0 * y0);
        }
        const double t = s.zMax - s.zMin, zc = 0.5 * (s.zMax + s.zMin);
        const double mass = s.density * area * t;
        const double zz = mass * (t * t / 12.0 + zc * zc);
        m = (float)mass;
        I[0] = (float)(s.density * t * ixx + zz);
        I[1] = (float)(s.density * t * iyy + zz);
        I[2] = (float)(s.density * t * (ixx + iyy));
        b.extent = (float)std::sqrt(far + std::max(s.zMin * s.zMin, s.zMax * s.zMax));
    }
    std::memset(b.inertia, 0, sizeof b.inertia);
    if (b.type == TFS_DYNAMIC && m > 0.0f) {
        b.inv_mass = 1.0f / m;
        b.inertia[0] = I[0];
        b.inertia[4] = I[1];
        b.inertia[8] = I[2];
    } else {
        b.inv_mass = 0.0f;
    }
}

}  // namespace

extern "C" {

// The restated reference narrowphase (collide3d.h, host build) on one pair, in tfr_collide's layout:
// kind 0 sphere (dims[0] = radius), 1 box (half-extents), 4 half-space; pa, pb = p[3], q[4];
// motion = v, w, centre of A then of B. out = normal[3], count, per point position[3], separation,
// predicted. Returns 0, or -1 for another kind.
int tfw_collide3d(int ka, const float* da, const float* pa, int kb, const float* db, const float* pb, float margin,
                  float linearSlop, const float* motion, float dt, int fast, float* out) {
    const tfc::Solid a{ ka, tf::vec(da[0], ka == tfc::kBox ? da[1] : 0.0f, ka == tfc::kBox ? da[2] : 0.0f),
                        ka == tfc::kSphere ? da[0] : 0.0f };
    const tfc::Solid b{ kb, tf::vec(db[0], kb == tfc::kBox ? db[1] : 0.0f, kb == tfc::kBox ? db[2] : 0.0f),
                        kb == tfc::kSphere ? db[0] : 0.0f };
    const tfc::Xf xa{ tf::ld3(pa), tf::ld4(pa + 3) }, xb{ tf::ld3(pb), tf::ld4(pb + 3) };
    const tfc::Motion mo{ tf::ld3(motion), tf::ld3(motion + 3), tf::ld3(motion + 6), tf::ld3(motion + 9),
                          tf::ld3(motion + 12), tf::ld3(motion + 15), dt, fast != 0 };
    tfc::Manifold m;
    if (!tfc::collide(a, xa, b, xb, margin, linearSlop, mo, m)) { return -1; }
    tf::st3(out, m.normal);
    out[3] = (float)m.count;
    for (int i = 0; i < 4; ++i) {
        const tfc::ManifoldPoint p = i < m.count ? m.points[i] : tfc::ManifoldPoint{ tf::vec(0, 0, 0), 0.0f, 0.0f };
        tf::st3(out + 4 + 5 * i, p.position);
        out[4 + 5 * i + 3] = p.separation;
        out[4 + 5 * i + 4] = p.predicted;
    }
    return 0;
}

// Half the OpenMP threads by default (TWOFOLD_HOST_THREADS overrides): a process held to a CPU quota
// that its spinning workers use up is stopped for the rest of the quota period, which stalls a solve
// step waiting on its refresh callback for milliseconds.
void* tfw_create(const tfs_settings* s) {
    World* w = new World();
    w->threads = std::max(1, omp_get_max_threads() / 2);
    if (const char* e = std::getenv("TWOFOLD_HOST_THREADS")) { w->threads = std::max(1, std::atoi(e)); }
    if (s) {
        w->settings = *s;
    } else {
        tfs_settings d{};
        d.gravity[1] = -9.81f;
        d.length_unit = 1.0f;
        d.substeps = 4;
        d.speculative = 0.02f;
        w->settings = d;
    }
    return w;
}

void tfw_destroy(void* h) { delete (World*)h; }

int tfw_add_body(void* h, int type, const float* pos, const float* rot, const float* v, const float* wv, int locks,
                 const float* lockFrame, int awake, float linearDamping, float angularDamping, float gravityScale) {
    World& w = *(World*)h;
    tfs_body b;
    std::memset(&b, 0, sizeof b);
    b.type = (uint8_t)type;
    for (int k = 0; k < 3; ++k) {
        b.p[k] = pos ? pos[k] : 0.0f;
        b.com[k] = b.p[k];
        b.v[k] = v ? v[k] : 0.0f;
        b.w[k] = wv ? wv[k] : 0.0f;
    }
    for (int k = 0; k < 4; ++k) {
        b.q[k] = rot ? rot[k] : (k == 3 ? 1.0f : 0.0f);
        b.lock_frame[k] = lockFrame ? lockFrame[k] : (k == 3 ? 1.0f : 0.0f);
    }
    b.locks = (uint8_t)locks;
    b.awake = awake ? 1 : 0;
    b.can_sleep = 1;
    b.alive = 1;
    b.kept = 1.0f;
    b.island = UINT32_MAX;
    b.linear_damping = linearDamping;
    b.angular_damping = angularDamping;
    b.gravity_scale = gravityScale;
    w.bodies.push_back(b);
    w.shapes.push_back(Shape{ -1 });
    w.group.push_back(0);
    return (int)w.bodies.size() - 1;
}

// Puts a body in a collision group (default 0); bodies of different groups never collide.
int tfw_set_group(void* h, uint32_t body, uint32_t group) {
    World& w = *(World*)h;
    if (body >= w.bodies.size()) { return -1; }
    w.group[body] = group;
    w.groups = std::max(w.groups, group + 1);
    return 0;
}

// One shape per body, at its origin: 0 sphere {r}, 1 box {hx, hy, hz}, 2 cylinder {r, half height}
// (axis local z), 3 prism {zMin, zMax, x0, y0, x1, y1, ...} (3 to 16 vertices of a convex polygon in
// either winding, about the local origin, which should be its centroid), 4 half-space {}.
int tfw_add_shape(void* h, int body, int kind, const float* dims, int ndims, const float* pos, const float* rot,
                  float density, float friction, float restitution, float rolling) {
    World& w = *(World*)h;
    if (body < 0 || (size_t)body >= w.bodies.size() || w.shapes[(size_t)body].kind != -1) { return -1; }
    if ((pos && (pos[0] != 0 || pos[1] != 0 || pos[2] != 0)) || (rot && (rot[0] != 0 || rot[1] != 0 || rot[2] != 0))) {
        return -1;
    }
    Shape s;
    s.kind = kind;
    if (kind == kSphere && ndims >= 1) {
        s.radius = dims[0];
    } else if (kind == kBox && ndims >= 3) {
        s.half = V{ dims[0], dims[1], dims[2] };
    } else if (kind == kCylinder && ndims >= 2) {
        s.radius = dims[0];
        s.half = V{ dims[0], dims[0], dims[1] };
    } else if (kind == kPrism && ndims >= 8 && ndims % 2 == 0 && ndims <= 2 + 2 * kMaxPrismVertices) {
        s.zMin = dims[0];
        s.zMax = dims[1];
        s.poly.assign(dims + 2, dims + ndims);
        const int n = (ndims - 2) / 2;
        double twice = 0;
        for (int i = 0; i < n; ++i) {
            const int j = (i + 1) % n;
            twice += (double)s.poly[2 * i] * s.poly[2 * j + 1] - (double)s.poly[2 * j] * s.poly[2 * i + 1];
        }
        if (twice < 0) {  // clockwise: turn it round
            for (int i = 0, j = n - 1; i < j; ++i, --j) {
                std::swap(s.poly[2 * i], s.poly[2 * j]);
                std::swap(s.poly[2 * i + 1], s.poly[2 * j + 1]);
            }
        }
        float hx = 0, hy = 0;
        for (int i = 0; i < n; ++i) {
            hx = std::max(hx, std::fabs(s.poly[2 * i]));
            hy = std::max(hy, std::fabs(s.poly[2 * i + 1]));
        }
        s.half = V{ hx, hy, std::max(std::fabs(s.zMin), std::fabs(s.zMax)) };
    } else if (kind != kHalfSpace) {
        return -1;
    }
    s.density = density;
    s.friction = friction;
    s.restitution = restitution;
    s.rolling = rolling;
    tfs_body& b = w.bodies[(size_t)body];
    if (kind == kHalfSpace && b.type != TFS_STATIC) { return -1; }
    w.shapes[(size_t)body] = s;
    massProperties(b, s);
    if (kind == kHalfSpace) {
        w.planes.push_back((uint32_t)body);
    } else if (b.type == TFS_STATIC) {
        w.statics.push_back((uint32_t)body);
    } else {
        w.movers.push_back((uint32_t)body);
    }
    return body;
}

uint32_t tfw_bodies(void* h, tfs_body* out, uint32_t cap) {
    World& w = *(World*)h;
    const uint32_t n = (uint32_t)w.bodies.size();
    if (out) { std::memcpy(out, w.bodies.data(), sizeof(tfs_body) * std::min(n, cap)); }
    return n;
}

// The contacts of the current poses, in key order, with impulses carried from `prev` (the last
// step's contacts, in key order). Returns the count; writes at most cap. Sleeping dynamic bodies in
// a touching pair with an awake one are woken in `bodies`; *woke counts them.
uint32_t tfw_collide(void* h, tfs_body* bodies, uint32_t nb, float dt, const tfs_contact* prev, uint32_t nprev,
                     tfs_contact* out, uint32_t cap, uint32_t* woke) {
    World& w = *(World*)h;
    if (nb != w.bodies.size()) { return 0; }
    const std::vector<uint64_t> pairs = candidates(w, bodies, dt);
    std::vector<tfs_contact> found(pairs.size());
    std::vector<uint8_t> hit(pairs.size(), 0);
    const float matchSq = w.settings.speculative * w.settings.speculative;
#pragma omp parallel for schedule(dynamic, 256) num_threads(w.threads)
    for (long k = 0; k < (long)pairs.size(); ++k) {
        const uint32_t ia = (uint32_t)(pairs[(size_t)k] >> 32), ib = (uint32_t)pairs[(size_t)k];
        Manifold m;
        if (!collide(w, ia, bodies[ia], ib, bodies[ib], dt, m)) { continue; }
        tfs_contact& c = found[(size_t)k];
        toContact(w, ia, bodies[ia], ib, bodies[ib], m, c);
        const tfs_contact* it = std::lower_bound(prev, prev + nprev, c.key,
                                                 [](const tfs_contact& x, uint64_t key) { return x.key < key; });
        if (it != prev + nprev && it->key == c.key) {
            carry(*it, c, matchSq);
            c.touched = it->touched;   // as updateContacts keeps it for the pair (src/world.cpp:895, 927)
        }
        hit[(size_t)k] = 1;
    }
    uint32_t n = 0, wakes = 0;
    for (size_t k = 0; k < pairs.size(); ++k) {
        if (!hit[k]) { continue; }
        if (out && n < cap) { out[n] = found[k]; }
        ++n;
        tfs_body& a = bodies[found[k].body_a];
        tfs_body& b = bodies[found[k].body_b];
        for (int s = 0; s < 2; ++s) {
            tfs_body& x = s ? b : a;
            const tfs_body& y = s ? a : b;
            if (x.type == TFS_DYNAMIC && !x.awake && moving(y)) {
                x.awake = 1;
                x.sleep_time = 0.0f;
                ++wakes;
            }
        }
    }
    if (woke) { *woke = wakes; }
    return n;
}

// The contact stream's input (tfs_stream_contacts): the candidate pairs of the current poses (those
// with a dynamic awake body), in key order, with this world's materials mixed for each (friction the
// geometric mean, restitution and rolling resistance the larger) and the narrowphase's manifold for
// each that has one, normal of unit length, points at the surfaces' midpoints. Returns the number of
// pairs and writes at most `cap` of them and of the manifolds (*n_manifolds: their number). Sleeping
// dynamic bodies of a pair with a manifold and an awake body are woken in `bodies`; *woke counts them
// (the caller uploads the bodies again when it is not 0, after streaming: updateContacts wakes them
// after the contacts are made, src/world.cpp:957-958, 987).
uint32_t tfw_pairs(void* h, tfs_body* bodies, uint32_t nb, float dt, tfs_pair* pairs, tfs_manifold* manifolds, uint32_t cap,
                   uint32_t* n_manifolds, uint32_t* woke) {
    World& w = *(World*)h;
    if (nb != w.bodies.size()) { return 0; }
    const std::vector<uint64_t> keys = candidates(w, bodies, dt);
    std::vector<Manifold> found(keys.size());
    std::vector<uint8_t> hit(keys.size(), 0);
#pragma omp parallel for schedule(dynamic, 256) num_threads(w.threads)
    for (long k = 0; k < (long)keys.size(); ++k) {
        const uint32_t ia = (uint32_t)(keys[(size_t)k] >> 32), ib = (uint32_t)keys[(size_t)k];
        if (collide(w, ia, bodies[ia], ib, bodies[ib], dt, found[(size_t)k])) { hit[(size_t)k] = 1; }
    }
    uint32_t nm = 0, wakes = 0;
    for (size_t k = 0; k < keys.size(); ++k) {
        const uint32_t ia = (uint32_t)(keys[k] >> 32), ib = (uint32_t)keys[k];
        if (k < cap && pairs) {
            tfs_pair& p = pairs[k];
            std::memset(&p, 0, sizeof p);
            p.key = keys[k];
            p.body_a = ia;
            p.body_b = ib;
            const Shape& sa = w.shapes[ia];
            const Shape& sb = w.shapes[ib];
            p.friction = std::sqrt(sa.friction * sb.friction);
            p.restitution = std::max(sa.restitution, sb.restitution);
            p.rolling_resistance = std::max(sa.rolling, sb.rolling);
            p.first = nm;
            p.count = hit[k] ? 1u : 0u;
        }
        if (!hit[k]) { continue; }
        if (nm < cap && manifolds) {
            const Manifold& m = found[k];
            tfs_manifold& x = manifolds[nm];
            std::memset(&x, 0, sizeof x);
            st(x.normal, m.reference ? m.n : unit(m.n));
            x.count = m.count;
            for (int i = 0; i < m.count && i < 4; ++i) {
                st(x.position[i], m.p[i].pos);
                x.separation[i] = m.p[i].sep;
            }
        }
        ++nm;
        tfs_body& a = bodies[ia];
        tfs_body& b = bodies[ib];
        for (int s = 0; s < 2; ++s) {
            tfs_body& x = s ? b : a;
            const tfs_body& y = s ? a : b;
            if (x.type == TFS_DYNAMIC && !x.awake && moving(y)) {
                x.awake = 1;
                x.sleep_time = 0.0f;
                ++wakes;
            }
        }
    }
    if (n_manifolds) { *n_manifolds = nm; }
    if (woke) { *woke = wakes; }
    return (uint32_t)keys.size();
}

// The bodies' shapes for the device refresh (tfs_set_body_shapes), which collides spheres, boxes and
// half-spaces with this world's narrowphase (narrow.h): the count, or -1 if a body's shape is a
// cylinder or a prism. Writes at most `cap`.
int tfw_body_shapes(void* h, tfs_shape* out, uint32_t cap) {
    const World& w = *(const World*)h;
    for (const Shape& s : w.shapes) {
        if (s.kind != kSphere && s.kind != kBox && s.kind != kHalfSpace) { return -1; }
    }
    for (uint32_t i = 0; i < cap && i < w.shapes.size(); ++i) {
        const Shape& s = w.shapes[i];
        out[i].kind = (uint32_t)s.kind;
        st(out[i].half, s.half);
        out[i].radius = s.radius;
    }
    return (int)w.shapes.size();
}

// tfs_refresh_fn: collides the listed contacts' pairs again at the bodies' current poses.
int tfw_refresh(void* h, int, float remaining, uint32_t n, const uint32_t*, const tfs_contact* contacts,
                const tfs_body* bodies, tfs_contact* fresh) {
    World& w = *(World*)h;
    // A few pairs are collided faster by this thread than by waking the pool.
#pragma omp parallel for schedule(dynamic, 64) num_threads(w.threads) if (n > 32)
    for (long k = 0; k < (long)n; ++k) {
        const tfs_contact& c = contacts[k];
        const tfs_body& a = bodies[2 * k];
        const tfs_body& b = bodies[2 * k + 1];
        Manifold m;
        tfs_contact& f = fresh[k];
        std::memset(&f, 0, sizeof f);
        if (collide(w, c.body_a, a, c.body_b, b, remaining, m)) { toContact(w, c.body_a, a, c.body_b, b, m, f); }
    }
    return 0;
}

}  // extern "C"
