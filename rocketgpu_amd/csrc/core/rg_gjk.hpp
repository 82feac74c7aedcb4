// Closest points and penetration between a car's hitbox (a Bullet box with a collision margin) and one mesh triangle.
//
// Reference: ComputeCarArenaTriangleContact (include/rocketgpu/detail/step/triangle_and_carcar.hpp:301-520) with
// RunCarArenaTriangleGjkSimplex (:202-299), the Voronoi simplex of raycast_wheels_and_contact_prep.hpp:4043-4478,
// the penetration-depth fallback BulletPortCalcPenDepth (triangle_and_carcar.hpp:98-158) and the gjkepa2 solver of
// bullet_gjk_epa2_port.hpp:560-1730.  These are Bullet's published algorithms (btGjkPairDetector with
// btVoronoiSimplexSolver, then btGjkEpa2 when the cores overlap): a distance query on the margin-less shapes, and
// an expanding-polytope search for the deepest separating direction when it fails.  The float operations follow the
// order Bullet evaluates them in, since contact points and normals must come out bit for bit; the data structures
// (index-linked faces, flat vertex pools, plain structs) are this file's own.
#pragma once

#include "rg_math.hpp"

namespace rg {
namespace core {
namespace gjk {

constexpr float kEps = 1.1920929e-7f;  // FLT_EPSILON

// A rigid transform as Bullet stores it: basis rows and an origin.
struct Xf {
    Vec3 r0, r1, r2;
    Vec3 o;
};

RG_HD Xf IdentityXf() {
    Xf t;
    t.r0 = Vec3(1.0f, 0.0f, 0.0f);
    t.r1 = Vec3(0.0f, 1.0f, 0.0f);
    t.r2 = Vec3(0.0f, 0.0f, 1.0f);
    t.o = Vec3(0.0f, 0.0f, 0.0f);
    return t;
}
// Basis times vector (row dots).
RG_HD Vec3 RowMul(const Xf& t, const Vec3& v) { return Vec3(Dot(t.r0, v), Dot(t.r1, v), Dot(t.r2, v)); }
// Vector times basis (column dots: row0.k * v.x + row1.k * v.y + row2.k * v.z).
RG_HD Vec3 ColMul(const Xf& t, const Vec3& v) {
    return Vec3(t.r0.x * v.x + t.r1.x * v.y + t.r2.x * v.z, t.r0.y * v.x + t.r1.y * v.y + t.r2.y * v.z,
                t.r0.z * v.x + t.r1.z * v.y + t.r2.z * v.z);
}
RG_HD Vec3 Apply(const Xf& t, const Vec3& v) { return RowMul(t, v) + t.o; }
// a^-1 * b: basis a^T b, origin (b.o - a.o) expressed in a's basis.
RG_HD Xf InverseTimes(const Xf& a, const Xf& b) {
    const auto col = [](const Xf& m, int k) {
        return k == 0 ? Vec3(m.r0.x, m.r1.x, m.r2.x) : (k == 1 ? Vec3(m.r0.y, m.r1.y, m.r2.y) : Vec3(m.r0.z, m.r1.z, m.r2.z));
    };
    Xf out;
    Vec3 rows[3];
    for (int i = 0; i < 3; ++i) {
        const Vec3 ai = col(a, i);
        // element (i, j) = a.r0[i] b.r0[j] + a.r1[i] b.r1[j] + a.r2[i] b.r2[j]
        rows[i] = Vec3(ai.x * b.r0.x + ai.y * b.r1.x + ai.z * b.r2.x, ai.x * b.r0.y + ai.y * b.r1.y + ai.z * b.r2.y,
                       ai.x * b.r0.z + ai.y * b.r1.z + ai.z * b.r2.z);
    }
    out.r0 = rows[0];
    out.r1 = rows[1];
    out.r2 = rows[2];
    out.o = ColMul(a, b.o - a.o);
    return out;
}

// Bullet's in-place normalize (divide by the length when its square exceeds 1e-20).
RG_HD Vec3 BtNormalized(const Vec3& v) {
    const float l2 = LengthSq(v);
    if (l2 > 1.0e-20f) return v * (1.0f / sqrtf(l2));
    return v;
}

// ------------------------------------------------------------------------------------------------------------
// The two convex shapes: a box (core half extents plus a margin) and a triangle (no margin).

struct Shape {
    uint8_t is_box;
    Vec3 core;     // box: half extents without the margin
    Vec3 v[3];     // triangle vertices
    float margin;
};

RG_HD Shape BoxShape(const Vec3& half_with_margin, float margin) {
    Shape s;
    s.is_box = 1;
    const float min_dim = Min(half_with_margin.x, Min(half_with_margin.y, half_with_margin.z));
    const float m = Min(margin, 0.1f * min_dim);
    s.core = half_with_margin - Vec3(m, m, m);
    s.margin = m;
    s.v[0] = s.v[1] = s.v[2] = Vec3(0.0f, 0.0f, 0.0f);
    return s;
}

RG_HD Shape TriShape(const Vec3& a, const Vec3& b, const Vec3& c) {
    Shape s;
    s.is_box = 0;
    s.core = Vec3(0.0f, 0.0f, 0.0f);
    s.v[0] = a;
    s.v[1] = b;
    s.v[2] = c;
    s.margin = 0.0f;
    return s;
}

RG_HD Vec3 SupportCore(const Shape& s, const Vec3& d) {
    if (s.is_box) {
        return Vec3(d.x >= 0.0f ? s.core.x : -s.core.x, d.y >= 0.0f ? s.core.y : -s.core.y,
                    d.z >= 0.0f ? s.core.z : -s.core.z);
    }
    const float a = Dot(d, s.v[0]), b = Dot(d, s.v[1]), c = Dot(d, s.v[2]);
    const int k = a < b ? (b < c ? 2 : 1) : (a < c ? 2 : 0);
    return s.v[k];
}

// Support along d of the shape grown by its margin (d normalized, (-1,-1,-1) for a vanishing d).
RG_HD Vec3 SupportWithMargin(const Shape& s, const Vec3& d) {
    Vec3 n = d;
    if (LengthSq(n) < kEps * kEps) n = Vec3(-1.0f, -1.0f, -1.0f);
    n = BtNormalized(n);
    return SupportCore(s, n) + n * s.margin;
}

// World box of a shape under a transform (Bullet's getAabb for the two kinds).
RG_HD void ShapeAabb(const Shape& s, const Xf& t, Vec3& lo, Vec3& hi) {
    if (s.is_box) {
        const Vec3 h = s.core + Vec3(s.margin, s.margin, s.margin);
        const Vec3 a0(Abs(t.r0.x), Abs(t.r0.y), Abs(t.r0.z));
        const Vec3 a1(Abs(t.r1.x), Abs(t.r1.y), Abs(t.r1.z));
        const Vec3 a2(Abs(t.r2.x), Abs(t.r2.y), Abs(t.r2.z));
        const Vec3 ext(Dot(h, a0), Dot(h, a1), Dot(h, a2));
        lo = t.o - ext;
        hi = t.o + ext;
        return;
    }
    float lo_k[3], hi_k[3];
    for (int k = 0; k < 3; ++k) {
        const Vec3 ax(k == 0 ? 1.0f : 0.0f, k == 1 ? 1.0f : 0.0f, k == 2 ? 1.0f : 0.0f);
        const Vec3 up = Apply(t, SupportWithMargin(s, ColMul(t, ax)));
        hi_k[k] = (k == 0 ? up.x : (k == 1 ? up.y : up.z)) + s.margin;
        const Vec3 nax(k == 0 ? -1.0f : 0.0f, k == 1 ? -1.0f : 0.0f, k == 2 ? -1.0f : 0.0f);
        const Vec3 dn = Apply(t, SupportWithMargin(s, ColMul(t, nax)));
        lo_k[k] = (k == 0 ? dn.x : (k == 1 ? dn.y : dn.z)) - s.margin;
    }
    lo = Vec3(lo_k[0], lo_k[1], lo_k[2]);
    hi = Vec3(hi_k[0], hi_k[1], hi_k[2]);
}

// ------------------------------------------------------------------------------------------------------------
// Voronoi-region simplex for the distance query (btVoronoiSimplexSolver semantics): up to four Minkowski vertices
// w = p - q with their witnesses p and q; the point of the simplex nearest the origin, its barycentric weights, and
// the reduction to the vertices that carry it.

struct Nearest {
    Vec3 point;
    uint8_t used;  // bit i: vertex i carries the nearest point
    float bary[4];
    uint8_t degenerate;
};

RG_HD void ResetNearest(Nearest& r) {
    r.point = Vec3(0.0f, 0.0f, 0.0f);
    r.used = 0;
    r.bary[0] = r.bary[1] = r.bary[2] = r.bary[3] = 0.0f;
    r.degenerate = 0;
}
RG_HD void SetBary(Nearest& r, float a, float b, float c, float d = 0.0f) {
    r.bary[0] = a;
    r.bary[1] = b;
    r.bary[2] = c;
    r.bary[3] = d;
}
RG_HD bool BaryNonNegative(const Nearest& r) {
    return r.bary[0] >= 0.0f && r.bary[1] >= 0.0f && r.bary[2] >= 0.0f && r.bary[3] >= 0.0f;
}

// Nearest point of triangle abc to p by Voronoi regions (vertices, edges, face), Ericson's ordering.
RG_HD void NearestOnTriangle(const Vec3& p, const Vec3& a, const Vec3& b, const Vec3& c, Nearest& r) {
    r.used = 0;
    const Vec3 ab = b - a, ac = c - a, ap = p - a;
    const float d1 = Dot(ab, ap), d2 = Dot(ac, ap);
    if (d1 <= 0.0f && d2 <= 0.0f) {
        r.point = a;
        r.used = 1;
        SetBary(r, 1.0f, 0.0f, 0.0f);
        return;
    }
    const Vec3 bp = p - b;
    const float d3 = Dot(ab, bp), d4 = Dot(ac, bp);
    if (d3 >= 0.0f && d4 <= d3) {
        r.point = b;
        r.used = 2;
        SetBary(r, 0.0f, 1.0f, 0.0f);
        return;
    }
    const float vc = d1 * d4 - d3 * d2;
    if (vc <= 0.0f && d1 >= 0.0f && d3 <= 0.0f) {
        const float v = d1 / (d1 - d3);
        r.point = a + ab * v;
        r.used = 3;
        SetBary(r, 1.0f - v, v, 0.0f);
        return;
    }
    const Vec3 cp = p - c;
    const float d5 = Dot(ab, cp), d6 = Dot(ac, cp);
    if (d6 >= 0.0f && d5 <= d6) {
        r.point = c;
        r.used = 4;
        SetBary(r, 0.0f, 0.0f, 1.0f);
        return;
    }
    const float vb = d5 * d2 - d1 * d6;
    if (vb <= 0.0f && d2 >= 0.0f && d6 <= 0.0f) {
        const float w = d2 / (d2 - d6);
        r.point = a + ac * w;
        r.used = 5;
        SetBary(r, 1.0f - w, 0.0f, w);
        return;
    }
    const float va = d3 * d6 - d5 * d4;
    if (va <= 0.0f && (d4 - d3) >= 0.0f && (d5 - d6) >= 0.0f) {
        const float w = (d4 - d3) / ((d4 - d3) + (d5 - d6));
        r.point = b + (c - b) * w;
        r.used = 6;
        SetBary(r, 0.0f, 1.0f - w, w);
        return;
    }
    const float denom = 1.0f / (va + vb + vc);
    const float v = vb * denom, w = vc * denom;
    r.point = a + ab * v + ac * w;
    r.used = 7;
    SetBary(r, 1.0f - v - w, v, w);
}

// Which side of plane abc the point p lies on relative to the fourth vertex d: 1 opposite, 0 same, -1 degenerate.
RG_HD int OutsidePlane(const Vec3& p, const Vec3& a, const Vec3& b, const Vec3& c, const Vec3& d) {
    const Vec3 n = Cross(b - a, c - a);
    const float sp = Dot(p - a, n), sd = Dot(d - a, n);
    if (sd * sd < (1.0e-4f * 1.0e-4f)) return -1;
    return sp * sd < 0.0f ? 1 : 0;
}

// Nearest point of tetrahedron abcd to p: p itself when inside; else the nearest of the faces p is outside of.
// Returns false when p is inside (or the tetrahedron is flat: degenerate set).
RG_HD bool NearestOnTetrahedron(const Vec3& p, const Vec3& a, const Vec3& b, const Vec3& c, const Vec3& d,
                                Nearest& out) {
    out.point = p;
    out.used = 15;
    const int o_abc = OutsidePlane(p, a, b, c, d);
    const int o_acd = OutsidePlane(p, a, c, d, b);
    const int o_adb = OutsidePlane(p, a, d, b, c);
    const int o_bdc = OutsidePlane(p, b, d, c, a);
    if (o_abc < 0 || o_acd < 0 || o_adb < 0 || o_bdc < 0) {
        out.degenerate = 1;
        return false;
    }
    if (!o_abc && !o_acd && !o_adb && !o_bdc) return false;
    float best = 3.402823466e+38f;
    Nearest t;
    ResetNearest(t);
    // Each face's vertices map back to tetrahedron slots: (slot of tri vertex 0, 1, 2).
    const auto take = [&](int s0, int s1, int s2) {
        const float dist = LengthSq(t.point - p);
        if (!(dist < best)) return;
        best = dist;
        out.point = t.point;
        const int slot[3] = {s0, s1, s2};
        float bary[4] = {0.0f, 0.0f, 0.0f, 0.0f};
        uint8_t used = 0;
        for (int k = 0; k < 3; ++k) {
            bary[slot[k]] = t.bary[k];
            if ((t.used >> k) & 1u) used |= static_cast<uint8_t>(1u << slot[k]);
        }
        out.used = used;
        SetBary(out, bary[0], bary[1], bary[2], bary[3]);
    };
    if (o_abc) {
        NearestOnTriangle(p, a, b, c, t);
        take(0, 1, 2);
    }
    if (o_acd) {
        NearestOnTriangle(p, a, c, d, t);
        take(0, 2, 3);
    }
    if (o_adb) {
        NearestOnTriangle(p, a, d, b, t);
        take(0, 3, 1);
    }
    if (o_bdc) {
        NearestOnTriangle(p, b, d, c, t);
        take(1, 3, 2);
    }
    return true;
}

struct Simplex {
    int n;
    Vec3 w[4], p[4], q[4];
    Vec3 near_p, near_q, near_v;
    Vec3 last_w;
    Nearest bc;
    uint8_t valid;
    uint8_t dirty;
};

RG_HD void SimplexReset(Simplex& s) {
    s.valid = 0;
    s.n = 0;
    s.dirty = 1;
    s.last_w = Vec3(1.0e18f, 1.0e18f, 1.0e18f);
    s.near_p = s.near_q = s.near_v = Vec3(0.0f, 0.0f, 0.0f);
    ResetNearest(s.bc);
}

RG_HD void SimplexAdd(Simplex& s, const Vec3& w, const Vec3& p, const Vec3& q) {
    s.last_w = w;
    s.dirty = 1;
    s.w[s.n] = w;
    s.p[s.n] = p;
    s.q[s.n] = q;
    ++s.n;
}

RG_HD bool SimplexHas(const Simplex& s, const Vec3& w) {
    for (int i = 0; i < s.n; ++i) {
        if (LengthSq(s.w[i] - w) <= 0.0001f) return true;
    }
    return w.x == s.last_w.x && w.y == s.last_w.y && w.z == s.last_w.z;
}

RG_HD void SimplexDrop(Simplex& s, int i) {
    --s.n;
    s.w[i] = s.w[s.n];
    s.p[i] = s.p[s.n];
    s.q[i] = s.q[s.n];
}

// Keep only the vertices that carry the nearest point, dropping from the last slot down.
RG_HD void SimplexReduce(Simplex& s, uint8_t used) {
    if (s.n >= 4 && !(used & 8)) SimplexDrop(s, 3);
    if (s.n >= 3 && !(used & 4)) SimplexDrop(s, 2);
    if (s.n >= 2 && !(used & 2)) SimplexDrop(s, 1);
    if (s.n >= 1 && !(used & 1)) SimplexDrop(s, 0);
}

RG_HD bool SimplexUpdate(Simplex& s) {
    if (!s.dirty) return s.valid != 0;
    ResetNearest(s.bc);
    s.dirty = 0;
    switch (s.n) {
        case 1:
            s.near_p = s.p[0];
            s.near_q = s.q[0];
            s.near_v = s.near_p - s.near_q;
            ResetNearest(s.bc);
            SetBary(s.bc, 1.0f, 0.0f, 0.0f, 0.0f);
            s.valid = BaryNonNegative(s.bc);
            break;
        case 2: {
            const Vec3 from = s.w[0], to = s.w[1];
            const Vec3 diff = -from;
            const Vec3 v = to - from;
            float t = Dot(v, diff);
            if (t > 0.0f) {
                const float vv = Dot(v, v);
                if (t < vv) {
                    t /= vv;
                    s.bc.used = 3;
                } else {
                    t = 1.0f;
                    s.bc.used = 2;
                }
            } else {
                t = 0.0f;
                s.bc.used = 1;
            }
            SetBary(s.bc, 1.0f - t, t, 0.0f, 0.0f);
            s.near_p = s.p[0] + (s.p[1] - s.p[0]) * t;
            s.near_q = s.q[0] + (s.q[1] - s.q[0]) * t;
            s.near_v = s.near_p - s.near_q;
            SimplexReduce(s, s.bc.used);
            s.valid = BaryNonNegative(s.bc);
            break;
        }
        case 3: {
            NearestOnTriangle(Vec3(0.0f, 0.0f, 0.0f), s.w[0], s.w[1], s.w[2], s.bc);
            const float* b = s.bc.bary;
            s.near_p = s.p[0] * b[0] + s.p[1] * b[1] + s.p[2] * b[2];
            s.near_q = s.q[0] * b[0] + s.q[1] * b[1] + s.q[2] * b[2];
            s.near_v = s.near_p - s.near_q;
            SimplexReduce(s, s.bc.used);
            s.valid = BaryNonNegative(s.bc);
            break;
        }
        case 4: {
            if (NearestOnTetrahedron(Vec3(0.0f, 0.0f, 0.0f), s.w[0], s.w[1], s.w[2], s.w[3], s.bc)) {
                const float* b = s.bc.bary;
                s.near_p = s.p[0] * b[0] + s.p[1] * b[1] + s.p[2] * b[2] + s.p[3] * b[3];
                s.near_q = s.q[0] * b[0] + s.q[1] * b[1] + s.q[2] * b[2] + s.q[3] * b[3];
                s.near_v = s.near_p - s.near_q;
                SimplexReduce(s, s.bc.used);
                s.valid = BaryNonNegative(s.bc);
            } else if (s.bc.degenerate) {
                s.valid = 0;
            } else {
                s.valid = 1;
                s.near_v = Vec3(0.0f, 0.0f, 0.0f);
            }
            break;
        }
        default:
            s.valid = 0;
            break;
    }
    return s.valid != 0;
}

// ------------------------------------------------------------------------------------------------------------
// gjkepa2: the GJK of the penetration solver (support vertices with or without margins, Johnson-style projection of
// the origin), then EPA over the simplex that encloses the origin.

struct Pair {
    Shape s0, s1;
    Vec3 to1_r0, to1_r1, to1_r2;  // rotation from shape 0's frame to shape 1's
    Xf to0;                       // shape 1's frame in shape 0's
    uint8_t margins;
};

RG_HD Vec3 PairSupport0(const Pair& m, const Vec3& d) {
    return m.margins ? SupportWithMargin(m.s0, d) : SupportCore(m.s0, d);
}
RG_HD Vec3 PairSupport1(const Pair& m, const Vec3& d) {
    const Vec3 d1(Dot(m.to1_r0, d), Dot(m.to1_r1, d), Dot(m.to1_r2, d));
    return Apply(m.to0, m.margins ? SupportWithMargin(m.s1, d1) : SupportCore(m.s1, d1));
}
RG_HD Vec3 PairSupport(const Pair& m, const Vec3& d) { return PairSupport0(m, d) - PairSupport1(m, -d); }

RG_HD Pair MakePair(const Shape& s0, const Xf& t0, const Shape& s1, const Xf& t1, bool margins) {
    Pair m;
    m.s0 = s0;
    m.s1 = s1;
    // t1.basis^T * t0.basis
    const Xf rel = InverseTimes(Xf{t1.r0, t1.r1, t1.r2, Vec3(0.0f, 0.0f, 0.0f)}, Xf{t0.r0, t0.r1, t0.r2, Vec3(0.0f, 0.0f, 0.0f)});
    m.to1_r0 = rel.r0;
    m.to1_r1 = rel.r1;
    m.to1_r2 = rel.r2;
    m.to0 = InverseTimes(t0, t1);
    m.margins = margins ? 1 : 0;
    return m;
}

constexpr int kEpaMaxVerts = 64;
constexpr int kEpaMaxFaces = 128;
constexpr int kPoolVerts = 4 + kEpaMaxVerts;  // slots 0..3: the GJK's own vertices
constexpr uint8_t kNone = 0xFF;

struct Vert {
    Vec3 d, w;  // search direction (unit) and Minkowski support point
};

struct GjkState {
    Pair m;
    Vert pool[kPoolVerts];
    uint8_t free_stack[4];
    int free_count;
    // Two simplices used alternately; each a list of pool indices with weights.
    uint8_t c[2][4];
    float p[2][4];
    int rank[2];
    int cur;
    Vec3 ray;
    int status;  // 0 valid, 1 inside, 2 failed
};

RG_HD void GjkVertex(GjkState& g, const Vec3& d, int slot) {
    g.pool[slot].d = d * (1.0f / Length(d));
    g.pool[slot].w = PairSupport(g.m, g.pool[slot].d);
}
RG_HD void GjkPush(GjkState& g, int sx, const Vec3& d) {
    const int r = g.rank[sx];
    g.p[sx][r] = 0.0f;
    const int slot = g.free_stack[--g.free_count];
    g.c[sx][r] = static_cast<uint8_t>(slot);
    g.rank[sx] = r + 1;
    GjkVertex(g, d, slot);
}
RG_HD void GjkPop(GjkState& g, int sx) { g.free_stack[g.free_count++] = g.c[sx][--g.rank[sx]]; }

RG_HD float Det3(const Vec3& a, const Vec3& b, const Vec3& c) {
    return (a.y * b.z * c.x + a.z * b.x * c.y - a.x * b.z * c.y - a.y * b.x * c.z + a.x * b.y * c.z - a.z * b.y * c.x);
}

// Origin projected on segment ab: squared distance, weights, mask of the vertices used (-1: degenerate).
RG_HD float ProjectOrigin2(const Vec3& a, const Vec3& b, float* w, uint32_t& m) {
    const Vec3 d = b - a;
    const float l = LengthSq(d);
    if (!(l > 0.0f)) return -1.0f;
    const float t = l > 0.0f ? -Dot(a, d) / l : 0.0f;
    if (t >= 1.0f) {
        w[0] = 0.0f;
        w[1] = 1.0f;
        m = 2;
        return LengthSq(b);
    }
    if (t <= 0.0f) {
        w[0] = 1.0f;
        w[1] = 0.0f;
        m = 1;
        return LengthSq(a);
    }
    w[1] = t;
    w[0] = 1.0f - t;
    m = 3;
    return LengthSq(a + d * t);
}

RG_HD float ProjectOrigin3(const Vec3& a, const Vec3& b, const Vec3& c, float* w, uint32_t& m) {
    const Vec3* vt[3] = {&a, &b, &c};
    const Vec3 dl[3] = {a - b, b - c, c - a};
    const Vec3 n = Cross(dl[0], dl[1]);
    const float l = LengthSq(n);
    if (!(l > 0.0f)) return -1.0f;
    float mindist = -1.0f;
    float subw[2] = {0.0f, 0.0f};
    uint32_t subm = 0;
    for (int i = 0; i < 3; ++i) {
        if (Dot(*vt[i], Cross(dl[i], n)) > 0.0f) {
            const int j = i == 2 ? 0 : i + 1;
            const float subd = ProjectOrigin2(*vt[i], *vt[j], subw, subm);
            if (mindist < 0.0f || subd < mindist) {
                mindist = subd;
                m = ((subm & 1) ? 1u << i : 0u) + ((subm & 2) ? 1u << j : 0u);
                w[i] = subw[0];
                w[j] = subw[1];
                w[j == 2 ? 0 : j + 1] = 0.0f;
            }
        }
    }
    if (mindist < 0.0f) {
        const float d = Dot(a, n);
        const float s = sqrtf(l);
        const Vec3 p = n * (d / l);
        mindist = LengthSq(p);
        m = 7;
        w[0] = Length(Cross(dl[1], b - p)) / s;
        w[1] = Length(Cross(dl[2], c - p)) / s;
        w[2] = 1.0f - (w[0] + w[1]);
    }
    return mindist;
}

RG_HD float ProjectOrigin4(const Vec3& a, const Vec3& b, const Vec3& c, const Vec3& d, float* w, uint32_t& m) {
    const Vec3* vt[4] = {&a, &b, &c, &d};
    const Vec3 dl[3] = {a - d, b - d, c - d};
    const float vl = Det3(dl[0], dl[1], dl[2]);
    const bool ng = (vl * Dot(a, Cross(b - c, a - b))) <= 0.0f;
    if (!(ng && Abs(vl) > 0.0f)) return -1.0f;
    float mindist = -1.0f;
    float subw[3] = {0.0f, 0.0f, 0.0f};
    uint32_t subm = 0;
    for (int i = 0; i < 3; ++i) {
        const int j = i == 2 ? 0 : i + 1;
        const float s = vl * Dot(d, Cross(dl[i], dl[j]));
        if (s > 0.0f) {
            const float subd = ProjectOrigin3(*vt[i], *vt[j], d, subw, subm);
            if (mindist < 0.0f || subd < mindist) {
                mindist = subd;
                m = ((subm & 1) ? 1u << i : 0u) + ((subm & 2) ? 1u << j : 0u) + ((subm & 4) ? 8u : 0u);
                w[i] = subw[0];
                w[j] = subw[1];
                w[j == 2 ? 0 : j + 1] = 0.0f;
                w[3] = subw[2];
            }
        }
    }
    if (mindist < 0.0f) {
        mindist = 0.0f;
        m = 15;
        w[0] = Det3(c, b, d) / vl;
        w[1] = Det3(a, c, d) / vl;
        w[2] = Det3(b, a, d) / vl;
        w[3] = 1.0f - (w[0] + w[1] + w[2]);
    }
    return mindist;
}

RG_HD int GjkEvaluate(GjkState& g, const Vec3& guess) {
    constexpr float kAccuracy = 0.0001f, kMinDistance = 0.0001f, kDuplicated = 0.0001f;
    constexpr int kMaxIterations = 128;
    int iterations = 0;
    float sqdist = 0.0f, alpha = 0.0f;
    Vec3 lastw[4];
    int clastw = 0;
    for (int k = 0; k < 4; ++k) g.free_stack[k] = static_cast<uint8_t>(k);
    g.free_count = 4;
    g.cur = 0;
    g.status = 0;
    g.rank[0] = 0;
    g.ray = guess;
    const float sqrl = LengthSq(g.ray);
    GjkPush(g, 0, sqrl > 0.0f ? -g.ray : Vec3(1.0f, 0.0f, 0.0f));
    g.p[0][0] = 1.0f;
    g.ray = g.pool[g.c[0][0]].w;
    sqdist = sqrl;
    lastw[0] = lastw[1] = lastw[2] = lastw[3] = g.ray;
    do {
        const int next = 1 - g.cur;
        const int cs = g.cur;
        const float rl = Length(g.ray);
        if (rl < kMinDistance) {
            g.status = 1;
            break;
        }
        GjkPush(g, cs, -g.ray);
        const Vec3 w = g.pool[g.c[cs][g.rank[cs] - 1]].w;
        bool found = false;
        for (int i = 0; i < 4; ++i) {
            if (LengthSq(w - lastw[i]) < kDuplicated) {
                found = true;
                break;
            }
        }
        if (found) {
            GjkPop(g, g.cur);
            break;
        }
        clastw = (clastw + 1) & 3;
        lastw[clastw] = w;
        const float omega = Dot(g.ray, w) / rl;
        alpha = omega > alpha ? omega : alpha;
        if (((rl - alpha) - (kAccuracy * rl)) <= 0.0f) {
            GjkPop(g, g.cur);
            break;
        }
        float weights[4];
        uint32_t mask = 0;
        const uint8_t* c = g.c[cs];
        switch (g.rank[cs]) {
            case 2:
                sqdist = ProjectOrigin2(g.pool[c[0]].w, g.pool[c[1]].w, weights, mask);
                break;
            case 3:
                sqdist = ProjectOrigin3(g.pool[c[0]].w, g.pool[c[1]].w, g.pool[c[2]].w, weights, mask);
                break;
            case 4:
                sqdist = ProjectOrigin4(g.pool[c[0]].w, g.pool[c[1]].w, g.pool[c[2]].w, g.pool[c[3]].w, weights, mask);
                break;
        }
        if (sqdist >= 0.0f) {
            g.rank[next] = 0;
            g.ray = Vec3(0.0f, 0.0f, 0.0f);
            g.cur = next;
            for (int i = 0, ni = g.rank[cs]; i < ni; ++i) {
                if (mask & (1u << i)) {
                    g.c[next][g.rank[next]] = g.c[cs][i];
                    g.p[next][g.rank[next]++] = weights[i];
                    g.ray += g.pool[g.c[cs][i]].w * weights[i];
                } else {
                    g.free_stack[g.free_count++] = g.c[cs][i];
                }
            }
            if (mask == 15) g.status = 1;
        } else {
            GjkPop(g, g.cur);
            break;
        }
        if (++iterations >= kMaxIterations) g.status = 2;
    } while (g.status == 0);
    return g.status;
}

// Grow the final simplex to a tetrahedron that encloses the origin (rank by rank, trying both signs of each axis).
RG_HD bool EncloseRank4(const GjkState& g) {
    const uint8_t* c = g.c[g.cur];
    const Vec3 w3 = g.pool[c[3]].w;
    return Abs(Det3(g.pool[c[0]].w - w3, g.pool[c[1]].w - w3, g.pool[c[2]].w - w3)) > 0.0f;
}
RG_HD bool EncloseRank3(GjkState& g) {
    const uint8_t* c = g.c[g.cur];
    const Vec3 w0 = g.pool[c[0]].w;
    const Vec3 n = Cross(g.pool[c[1]].w - w0, g.pool[c[2]].w - w0);
    if (LengthSq(n) > 0.0f) {
        GjkPush(g, g.cur, n);
        if (EncloseRank4(g)) return true;
        GjkPop(g, g.cur);
        GjkPush(g, g.cur, -n);
        if (EncloseRank4(g)) return true;
        GjkPop(g, g.cur);
    }
    return false;
}
RG_HD bool EncloseRank2(GjkState& g) {
    const uint8_t* c = g.c[g.cur];
    const Vec3 d = g.pool[c[1]].w - g.pool[c[0]].w;
    for (int i = 0; i < 3; ++i) {
        const Vec3 axis(i == 0 ? 1.0f : 0.0f, i == 1 ? 1.0f : 0.0f, i == 2 ? 1.0f : 0.0f);
        const Vec3 p = Cross(d, axis);
        if (LengthSq(p) > 0.0f) {
            GjkPush(g, g.cur, p);
            if (EncloseRank3(g)) return true;
            GjkPop(g, g.cur);
            GjkPush(g, g.cur, -p);
            if (EncloseRank3(g)) return true;
            GjkPop(g, g.cur);
        }
    }
    return false;
}
RG_HD bool EncloseRank1(GjkState& g) {
    for (int i = 0; i < 3; ++i) {
        const Vec3 axis(i == 0 ? 1.0f : 0.0f, i == 1 ? 1.0f : 0.0f, i == 2 ? 1.0f : 0.0f);
        GjkPush(g, g.cur, axis);
        if (EncloseRank2(g)) return true;
        GjkPop(g, g.cur);
        GjkPush(g, g.cur, -axis);
        if (EncloseRank2(g)) return true;
        GjkPop(g, g.cur);
    }
    return false;
}
RG_HD bool EncloseOrigin(GjkState& g) {
    switch (g.rank[g.cur]) {
        case 1: return EncloseRank1(g);
        case 2: return EncloseRank2(g);
        case 3: return EncloseRank3(g);
        case 4: return EncloseRank4(g);
    }
    return false;
}

// EPA polytope: faces in a fixed pool, linked into the hull list or the free list (most recent first).
struct Face {
    Vec3 n;
    float d;
    uint8_t v[3];     // pool vertices
    uint8_t adj[3];   // neighbouring face across edge k
    uint8_t adj_e[3]; // that neighbour's edge index
    uint8_t prev, next;
    uint8_t pass;
};

struct Epa {
    Face f[kEpaMaxFaces];
    uint8_t hull, stock;
    int hull_count, stock_count;
    int next_vert;  // next free pool slot (from 4)
    int status;     // 0 valid, 1 accuracy reached, 2 other terminal states, 3 fallback
    Vec3 normal;
    float depth;
    uint8_t res_v[3];
    float res_p[3];
    int res_rank;
};

RG_HD void ListAppend(Epa& e, uint8_t& root, int& count, uint8_t face) {
    e.f[face].prev = kNone;
    e.f[face].next = root;
    if (root != kNone) e.f[root].prev = face;
    root = face;
    ++count;
}
RG_HD void ListRemove(Epa& e, uint8_t& root, int& count, uint8_t face) {
    if (e.f[face].next != kNone) e.f[e.f[face].next].prev = e.f[face].prev;
    if (e.f[face].prev != kNone) e.f[e.f[face].prev].next = e.f[face].next;
    if (face == root) root = e.f[face].next;
    --count;
}
RG_HD void Bind(Epa& e, uint8_t fa, int ea, uint8_t fb, int eb) {
    e.f[fa].adj_e[ea] = static_cast<uint8_t>(eb);
    e.f[fa].adj[ea] = fb;
    e.f[fb].adj_e[eb] = static_cast<uint8_t>(ea);
    e.f[fb].adj[eb] = fa;
}

// Distance of the origin from the face plane when the origin projects outside edge ab (nearest edge point).
RG_HD bool EdgeDistance(const Vec3& n, const Vec3& a, const Vec3& b, float& dist) {
    const Vec3 ba = b - a;
    const Vec3 n_ab = Cross(ba, n);
    if (!(Dot(a, n_ab) < 0.0f)) return false;
    const float ba_l2 = LengthSq(ba);
    const float a_dot_ba = Dot(a, ba);
    const float b_dot_ba = Dot(b, ba);
    if (a_dot_ba > 0.0f) {
        dist = Length(a);
    } else if (b_dot_ba < 0.0f) {
        dist = Length(b);
    } else {
        const float a_dot_b = Dot(a, b);
        const float v = (LengthSq(a) * LengthSq(b) - a_dot_b * a_dot_b) / ba_l2;
        dist = sqrtf(v > 0.0f ? v : 0.0f);
    }
    return true;
}

RG_HD uint8_t NewFace(Epa& e, const GjkState& g, uint8_t a, uint8_t b, uint8_t c, bool forced) {
    constexpr float kAccuracy = 0.0001f, kPlaneEps = 0.00001f;
    if (e.stock == kNone) {
        e.status = 2;
        return kNone;
    }
    const uint8_t face = e.stock;
    ListRemove(e, e.stock, e.stock_count, face);
    ListAppend(e, e.hull, e.hull_count, face);
    Face& f = e.f[face];
    f.pass = 0;
    f.v[0] = a;
    f.v[1] = b;
    f.v[2] = c;
    const Vec3 wa = g.pool[a].w, wb = g.pool[b].w, wc = g.pool[c].w;
    f.n = Cross(wb - wa, wc - wa);
    const float l = Length(f.n);
    if (l > kAccuracy) {
        if (!EdgeDistance(f.n, wa, wb, f.d) && !EdgeDistance(f.n, wb, wc, f.d) && !EdgeDistance(f.n, wc, wa, f.d)) {
            f.d = Dot(wa, f.n) / l;
        }
        f.n = f.n * (1.0f / l);
        if (forced || f.d >= -kPlaneEps) return face;
        e.status = 2;
    } else {
        e.status = 2;
    }
    ListRemove(e, e.hull, e.hull_count, face);
    ListAppend(e, e.stock, e.stock_count, face);
    return kNone;
}

RG_HD uint8_t BestFace(const Epa& e) {
    uint8_t best = e.hull;
    float mind = e.f[best].d * e.f[best].d;
    for (uint8_t f = e.f[best].next; f != kNone; f = e.f[f].next) {
        const float sqd = e.f[f].d * e.f[f].d;
        if (sqd < mind) {
            best = f;
            mind = sqd;
        }
    }
    return best;
}

struct Horizon {
    uint8_t cf, ff;
    int nf;
};

// Silhouette walk from face `face` across edge `edge` as seen from the new vertex w: faces w sees are removed, and a
// new face is stitched along every horizon edge.  Depth-first as Bullet's recursion (both neighbours of a visible
// face must succeed before it is removed), on an explicit stack.
RG_HD bool Expand(Epa& e, const GjkState& g, int pass, uint8_t w, uint8_t face, int edge, Horizon& h) {
    constexpr float kPlaneEps = 0.00001f;
    struct Frame {
        uint8_t face, edge, stage;
    };
    Frame st[kEpaMaxFaces];
    int sp = 0;
    st[sp++] = Frame{face, static_cast<uint8_t>(edge), 0};
    bool ret = false;
    while (sp > 0) {
        Frame& fr = st[sp - 1];
        Face& cf = e.f[fr.face];
        const int ce = fr.edge;
        if (fr.stage == 0) {
            if (cf.pass == pass) {
                ret = false;
                --sp;
                continue;
            }
            const int e1 = ce == 2 ? 0 : ce + 1;
            if ((Dot(cf.n, g.pool[w].w) - cf.d) < -kPlaneEps) {
                const uint8_t nf = NewFace(e, g, cf.v[e1], cf.v[ce], w, false);
                if (nf != kNone) {
                    Bind(e, nf, 0, fr.face, ce);
                    if (h.cf != kNone) {
                        Bind(e, h.cf, 1, nf, 2);
                    } else {
                        h.ff = nf;
                    }
                    h.cf = nf;
                    ++h.nf;
                    ret = true;
                } else {
                    ret = false;
                }
                --sp;
                continue;
            }
            cf.pass = static_cast<uint8_t>(pass);
            fr.stage = 1;
            if (sp < kEpaMaxFaces) {
                st[sp++] = Frame{cf.adj[e1], cf.adj_e[e1], 0};
                continue;
            }
            ret = false;
            --sp;
            continue;
        }
        if (fr.stage == 1) {
            if (!ret) {
                --sp;
                continue;
            }
            const int e2 = ce == 0 ? 2 : ce - 1;
            fr.stage = 2;
            if (sp < kEpaMaxFaces) {
                st[sp++] = Frame{cf.adj[e2], cf.adj_e[e2], 0};
                continue;
            }
            ret = false;
            --sp;
            continue;
        }
        if (ret) {
            ListRemove(e, e.hull, e.hull_count, fr.face);
            ListAppend(e, e.stock, e.stock_count, fr.face);
        }
        --sp;
    }
    return ret;
}

RG_HD void EpaInit(Epa& e) {
    e.status = 2;
    e.normal = Vec3(0.0f, 0.0f, 0.0f);
    e.depth = 0.0f;
    e.next_vert = 4;
    e.hull = kNone;
    e.stock = kNone;
    e.hull_count = e.stock_count = 0;
    for (int i = 0; i < kEpaMaxFaces; ++i) {
        ListAppend(e, e.stock, e.stock_count, static_cast<uint8_t>(kEpaMaxFaces - i - 1));
    }
}

RG_HD void EpaEvaluate(Epa& e, GjkState& g, const Vec3& guess) {
    constexpr float kAccuracy = 0.0001f;
    constexpr int kMaxIterations = 255;
    uint8_t* sc = g.c[g.cur];
    float* spw = g.p[g.cur];
    if (g.rank[g.cur] > 1 && EncloseOrigin(g)) {
        e.status = 0;
        const Vec3 w0 = g.pool[sc[0]].w, w1 = g.pool[sc[1]].w, w2 = g.pool[sc[2]].w, w3 = g.pool[sc[3]].w;
        if (Det3(w0 - w3, w1 - w3, w2 - w3) < 0.0f) {
            const uint8_t tc = sc[0];
            sc[0] = sc[1];
            sc[1] = tc;
            const float tp = spw[0];
            spw[0] = spw[1];
            spw[1] = tp;
        }
        const uint8_t tetra[4] = {NewFace(e, g, sc[0], sc[1], sc[2], true), NewFace(e, g, sc[1], sc[0], sc[3], true),
                                  NewFace(e, g, sc[2], sc[1], sc[3], true), NewFace(e, g, sc[0], sc[2], sc[3], true)};
        if (e.hull_count == 4) {
            uint8_t best = BestFace(e);
            Face outer = e.f[best];
            int pass = 0;
            Bind(e, tetra[0], 0, tetra[1], 0);
            Bind(e, tetra[0], 1, tetra[2], 0);
            Bind(e, tetra[0], 2, tetra[3], 0);
            Bind(e, tetra[1], 1, tetra[3], 2);
            Bind(e, tetra[1], 2, tetra[2], 1);
            Bind(e, tetra[2], 2, tetra[3], 1);
            e.status = 0;
            for (int it = 0; it < kMaxIterations; ++it) {
                if (e.next_vert >= kPoolVerts) {
                    e.status = 2;
                    break;
                }
                Horizon h{kNone, kNone, 0};
                const uint8_t w = static_cast<uint8_t>(e.next_vert++);
                bool valid = true;
                e.f[best].pass = static_cast<uint8_t>(++pass);
                GjkVertex(g, e.f[best].n, w);
                const float wdist = Dot(e.f[best].n, g.pool[w].w) - e.f[best].d;
                if (!(wdist > kAccuracy)) {
                    e.status = 1;
                    break;
                }
                for (int j = 0; j < 3 && valid; ++j) {
                    valid = Expand(e, g, pass, w, e.f[best].adj[j], e.f[best].adj_e[j], h) && valid;
                }
                if (!(valid && h.nf >= 3)) {
                    e.status = 2;
                    break;
                }
                Bind(e, h.cf, 1, h.ff, 2);
                ListRemove(e, e.hull, e.hull_count, best);
                ListAppend(e, e.stock, e.stock_count, best);
                best = BestFace(e);
                outer = e.f[best];
            }
            const Vec3 projection = outer.n * outer.d;
            e.normal = outer.n;
            e.depth = outer.d;
            e.res_rank = 3;
            for (int k = 0; k < 3; ++k) e.res_v[k] = outer.v[k];
            const Vec3 a = g.pool[outer.v[0]].w, b = g.pool[outer.v[1]].w, c = g.pool[outer.v[2]].w;
            e.res_p[0] = Length(Cross(b - projection, c - projection));
            e.res_p[1] = Length(Cross(c - projection, a - projection));
            e.res_p[2] = Length(Cross(a - projection, b - projection));
            const float sum = e.res_p[0] + e.res_p[1] + e.res_p[2];
            e.res_p[0] /= sum;
            e.res_p[1] /= sum;
            e.res_p[2] /= sum;
            return;
        }
    }
    e.status = 3;
    e.normal = -guess;
    const float nl = Length(e.normal);
    e.normal = nl > 0.0f ? e.normal * (1.0f / nl) : Vec3(1.0f, 0.0f, 0.0f);
    e.depth = 0.0f;
    e.res_rank = 1;
    e.res_v[0] = sc[0];
    e.res_p[0] = 1.0f;
}

struct DepthResult {
    Vec3 witness0, witness1, normal;
};

// gjkepa2 Penetration: GJK with margins from -guess; inside -> EPA; witnesses on shape 0 from the EPA face.
RG_HD bool Penetration(GjkState& g, Epa& e, const Shape& s0, const Xf& t0, const Shape& s1, const Xf& t1,
                       const Vec3& guess, DepthResult& out) {
    g.m = MakePair(s0, t0, s1, t1, true);
    if (GjkEvaluate(g, -guess) != 1) return false;
    EpaInit(e);
    EpaEvaluate(e, g, -guess);
    Vec3 w0(0.0f, 0.0f, 0.0f);
    for (int i = 0; i < e.res_rank; ++i) {
        Vec3 dir = g.pool[e.res_v[i]].d;
        if (LengthSq(dir) < kEps * kEps) dir = Vec3(-1.0f, -1.0f, -1.0f);
        dir = BtNormalized(dir);
        const Vec3 sup = SupportCore(s0, dir) + dir * s0.margin;
        w0 += sup * e.res_p[i];
    }
    out.witness0 = Apply(t0, w0);
    out.witness1 = Apply(t0, w0 - e.normal * e.depth);
    out.normal = -e.normal;
    return true;
}

// gjkepa2 Distance: GJK without margins from +guess; witnesses from the final simplex.
RG_HD bool Distance(GjkState& g, const Shape& s0, const Xf& t0, const Shape& s1, const Xf& t1, const Vec3& guess,
                    DepthResult& out) {
    g.m = MakePair(s0, t0, s1, t1, false);
    if (GjkEvaluate(g, guess) != 0) return false;
    Vec3 w0(0.0f, 0.0f, 0.0f), w1(0.0f, 0.0f, 0.0f);
    const int sx = g.cur;
    for (int i = 0; i < g.rank[sx]; ++i) {
        const float p = g.p[sx][i];
        const Vec3 d = g.pool[g.c[sx][i]].d;
        w0 += PairSupport0(g.m, d) * p;
        w1 += PairSupport1(g.m, -d) * p;
    }
    out.witness0 = Apply(t0, w0);
    out.witness1 = Apply(t0, w1);
    Vec3 n = w0 - w1;
    const float dist = Length(n);
    n = n * (1.0f / (dist > 0.0001f ? dist : 1.0f));
    out.normal = n;
    return true;
}

// The penetration-depth fallback: EPA from the direction between the frames, else the margin-less distance.
// Returns 1 when penetrating, 0 when separated with witnesses, -1 when both fail (zero outputs).
RG_HD int PenetrationDepth(const Shape& s0, const Xf& t0, const Shape& s1, const Xf& t1, Vec3& v, Vec3& wa, Vec3& wb) {
    GjkState g;
    Epa e;
    Vec3 guess = t1.o - t0.o;
    {
        const float l2 = LengthSq(guess);
        guess = l2 > 1.0e-20f ? guess * (1.0f / sqrtf(l2)) : Vec3(0.0f, 0.0f, 0.0f);
    }
    DepthResult r;
    if (Penetration(g, e, s0, t0, s1, t1, guess, r)) {
        wa = r.witness0;
        wb = r.witness1;
        v = r.normal;
        return 1;
    }
    if (Distance(g, s0, t0, s1, t1, guess, r)) {
        wa = r.witness0;
        wb = r.witness1;
        v = r.normal;
        return 0;
    }
    wa = wb = v = Vec3(0.0f, 0.0f, 0.0f);
    return -1;
}

// ------------------------------------------------------------------------------------------------------------
// Box vs triangle: btGjkPairDetector's closest points (Voronoi simplex on the margin-less shapes, margins applied
// after), EPA when the cores overlap or the result is degenerate and shallow, then the distance filter and the
// normal turned to point from the triangle toward the box.

struct BoxTriHit {
    Vec3 normal;      // unit, from the triangle toward the box
    float distance;   // signed (negative: penetration), bt
    Vec3 point_a;     // on the box
    Vec3 point_b;     // on the triangle
};

// box_half: the box's half extents (bt) before the margin is taken off; box: its world transform; tri in world.
RG_HD bool BoxTriangleContact(const Vec3& box_half, const Xf& box, const Vec3& ta, const Vec3& tb, const Vec3& tc,
                              BoxTriHit& out) {
    constexpr float kRelError2 = 1.0e-6f;
    constexpr float kMaxDistSq = 100.0f;
    const Shape s_box = BoxShape(box_half, 0.04f);
    const Shape s_tri = TriShape(ta, tb, tc);
    const float margin_a = s_box.margin, margin_b = s_tri.margin;
    const float margin = margin_a + margin_b;
    Xf xa = box, xb = IdentityXf();
    const Vec3 offset = (xa.o + xb.o) * 0.5f;
    xa.o = xa.o - offset;
    xb.o = xb.o - offset;

    // Distance query on the cores.
    Simplex sim;
    SimplexReset(sim);
    Vec3 axis(0.0f, 1.0f, 0.0f);
    float sqd = 1.0e30f;
    int degenerate = 0;
    bool check = false;
    int iter = 0;
    for (;;) {
        const Vec3 pin = SupportCore(s_box, ColMul(xa, -axis));
        const Vec3 qin = SupportCore(s_tri, ColMul(xb, axis));
        const Vec3 pw = Apply(xa, pin), qw = Apply(xb, qin);
        const Vec3 w = pw - qw;
        const float delta = Dot(axis, w);
        const bool seen = SimplexHas(sim, w);
        const float f0 = sqd - delta, f1 = sqd * kRelError2;
        if (delta > 0.0f && delta * delta > sqd * kMaxDistSq) {
            degenerate = 10;
            check = true;
            break;
        }
        if (seen) {
            degenerate = 1;
            check = true;
            break;
        }
        if (f0 <= f1) {
            degenerate = f0 <= 0.0f ? 2 : 11;
            check = true;
            break;
        }
        SimplexAdd(sim, w, pw, qw);
        const bool ok = SimplexUpdate(sim);
        const Vec3 next = sim.near_v;
        if (!ok) {
            degenerate = 3;
            check = true;
            break;
        }
        if (LengthSq(next) < kRelError2) {
            axis = next;
            degenerate = 6;
            check = true;
            break;
        }
        const float prev = sqd;
        sqd = LengthSq(next);
        if (prev - sqd <= kEps * prev) {
            check = true;
            degenerate = 12;
            break;
        }
        axis = next;
        if (iter++ > 1000) break;
        if (sim.n == 4) {
            degenerate = 13;
            break;
        }
    }
    Vec3 pa(0.0f, 0.0f, 0.0f), pb(0.0f, 0.0f, 0.0f), normal(0.0f, 0.0f, 0.0f);
    float distance = 0.0f;
    bool valid = false;
    if (check) {
        SimplexUpdate(sim);
        pa = sim.near_p;
        pb = sim.near_q;
        const float len_sqr = LengthSq(axis);
        if (len_sqr <= 0.0001f) degenerate = 5;
        if (len_sqr > kEps * kEps) {
            const float rlen = 1.0f / sqrtf(len_sqr);
            normal = axis * rlen;
            const float s = sqrtf(sqd);
            if (s > 0.0f) {
                pa -= axis * (margin_a / s);
                pb += axis * (margin_b / s);
                distance = (1.0f / rlen) - margin;
                valid = true;
            }
        }
    }
    const bool shallow_degenerate = degenerate != 0 && (distance + margin) <= 0.01f;
    if (!valid || shallow_degenerate) {
        Vec3 v, wa, wb;
        const int pen = PenetrationDepth(s_box, xa, s_tri, xb, v, wa, wb);
        if (pen == 1) {
            Vec3 n = wb - wa;
            float l2 = LengthSq(n);
            if (l2 <= kEps * kEps) {
                n = v;
                l2 = LengthSq(v);
            }
            if (l2 > kEps * kEps) {
                n = n * (1.0f / sqrtf(l2));
                const float d2 = -Length(wa - wb);
                if (!valid || d2 < distance) {
                    distance = d2;
                    pa = wa;
                    pb = wb;
                    normal = n;
                    valid = true;
                }
            }
        } else if (LengthSq(v) > 0.0f) {
            const float d2 = Length(wa - wb) - margin;
            if (!valid || d2 < distance) {
                pa = wa - v * margin_a;
                pb = wb + v * margin_b;
                normal = Normalize(v, v);
                distance = d2;
                valid = true;
            }
        }
    }
    if (!valid) return false;
    if (!(distance < 0.0f || distance * distance < kMaxDistSq)) return false;
    pa += offset;
    pb += offset;
    // Turn the normal to point from the triangle's box center toward the box's.
    Vec3 lo, hi;
    ShapeAabb(s_box, xa, lo, hi);
    const Vec3 ca = (hi + lo) * 0.5f;
    ShapeAabb(s_tri, xb, lo, hi);
    const Vec3 cb = (hi + lo) * 0.5f;
    if (Dot(ca - cb, normal) < 0.0f) normal = normal * -1.0f;
    out.normal = normal;
    out.distance = distance;
    out.point_a = pa;
    out.point_b = pb;
    return true;
}

}  // namespace gjk
}  // namespace core
}  // namespace rg
