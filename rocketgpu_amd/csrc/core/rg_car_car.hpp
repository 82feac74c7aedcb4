// Car against car: each pair's persistent hitbox manifold, the bumps the contact gives, the pair's solver rows, and
// the joint solve of the cars that touch each other.
//
// Reference (include/rocketgpu/detail/step/):
//   the gate and pair order WorldStepSoccar env_step.hpp:4914-5031, CarCarCcmGateMasks :139-205;
//   ResolveCarCarCollisionImpl<true> :734-860 -> PrepareSharedCarCarCollisionContacts :207-380 (manifold refresh
//   :42-71, point merge AddCarCarRetainedPoint :78-132) and EmitSharedCarCarContactRows :382-732;
//   the box query ComputeObbContact triangle_and_carcar.hpp:2221-2365 in game order :2367-2392; the face clip
//   ComputeFaceClippedCarCarContactAnchorPairs :1934-2097 with IntersectRectQuad2 :1817-1868 and CullFacePoints2
//   :1870-1932; the edge pair ComputeEdgeEdgeCarCarContactAnchorsBulletStyle :2121-2202; the anchors
//   ComputeApproxCarCarContactAnchorPairsWithBtObb :3093-3226;
//   the bump ResolveCarCarDirectionalContact :2901-2975, BumpCone :2830-2898, InForwardConePose :2782-2814,
//   RecordCarHit :2821-2828; the demolition test ShouldDemolish :2757-2780;
//   the joint solve SolveUnifiedVelocityContactRowsGeneral arena_geometry.hpp:5527-5866 with
//   SolveSharedCarVelocityNormalRow :3103-3196 and SolveSharedCarVelocityFrictionRow :3198-3272, and
//   SolveUnifiedSplitContactRows :5035-5276 with MarkBtSplitAccumulators :3318-3432 and SolveSharedCarSplitRow
//   :3435-3502.
//
// What it computes, in Bullet's terms: the two hitbox boxes go through the box-box detector (separating axes over
// the 15 candidate axes; a face axis clips the incident face against the reference face, keeping up to four points
// spread around the deepest; an edge pair gives one point), the points enter the pair's persistent manifold
// (nearest-point replacement by the point on b, the deepest kept, largest area when full), and every manifold point
// gives one non-penetration row (restitution 0.1), one friction row (0.09) and, when the cars overlap, one
// split-impulse row.  The rows couple the two cars, so every car that has such rows is solved jointly with its mesh
// rows, in the order of the arena's contact-pair list.
#pragma once

#include "rg_car_world.hpp"

namespace rg {
namespace core {

// Bounding-sphere radius of a car's hitbox about its origin (uu): half the hitbox diagonal plus the offset, plus 5.
RG_HD float CarBoundingRadius(const CarConfig& cfg) {
    return Length(cfg.hitbox_size * 0.5f) + Length(cfg.hitbox_offset) + 5.0f;
}

// ------------------------------------------------------------------------------------------------------------
// The hitbox box and the box-box query.

struct Obb {
    Vec3 c;       // center (bt)
    Vec3 ax[3];   // axes (the body's orientation columns)
    Vec3 h;       // half extents with the margin (bt)
};

RG_HD float Elem(const Vec3& v, int i) { return i == 0 ? v.x : (i == 1 ? v.y : v.z); }

RG_HD Obb CarObb(const CarState& car, const CarConfig& cfg) {
    Obb o;
    CarBasis(car, o.ax[0], o.ax[1], o.ax[2]);
    const Vec3 off = cfg.hitbox_offset * (1.0f / 50.0f);
    o.c = car.pos_bt + ((o.ax[0] * off.x + o.ax[1] * off.y) + o.ax[2] * off.z);
    o.h = HitboxHalfBt(cfg);
    return o;
}

enum : uint8_t { kObbFaceA = 0, kObbFaceB = 1, kObbEdges = 2 };

struct ObbFeature {
    uint8_t kind;
    uint8_t axis_a;  // face or edge axis on the first box
    uint8_t axis_b;  // face or edge axis on the second box
};

// Separating-axis test of boxes p and q.  False when an axis separates them (a face axis by more than `threshold`,
// an edge axis by more than FLT_EPSILON).  Otherwise the axis of least overlap, turned to point from p toward q, the
// overlap along it (positive when the boxes interpenetrate) and the feature that gave it.  Face axes of p, then of
// q, then the nine edge cross products; an edge axis wins only when its depth beats the best by the 1.05 fudge.
RG_HD bool ObbLeastOverlap(const Obb& p, const Obb& q, float threshold, Vec3& normal, float& overlap, ObbFeature& f) {
    constexpr float kEdgeEps = 1.1920929e-7f;
    float R[3][3], AR[3][3], ARf[3][3];
    for (int i = 0; i < 3; ++i) {
        for (int j = 0; j < 3; ++j) {
            R[i][j] = Dot(p.ax[i], q.ax[j]);
            ARf[i][j] = Abs(R[i][j]);
            AR[i][j] = ARf[i][j] + 1e-5f;
        }
    }
    const Vec3 d = q.c - p.c;
    const float t[3] = {Dot(d, p.ax[0]), Dot(d, p.ax[1]), Dot(d, p.ax[2])};
    float best = -1.0e30f;
    Vec3 axis(1.0f, 0.0f, 0.0f);
    bool flip = false;
    ObbFeature bf{kObbEdges, 0, 0};
    for (int i = 0; i < 3; ++i) {
        const float e2 = ((Elem(p.h, i) + q.h.x * ARf[i][0]) + q.h.y * ARf[i][1]) + q.h.z * ARf[i][2];
        const float s = Abs(t[i]) - e2;
        if (s > threshold) return false;
        if (i == 0 || s > best) {
            best = s;
            axis = p.ax[i];
            flip = t[i] < 0.0f;
            bf = ObbFeature{kObbFaceA, static_cast<uint8_t>(i), 0};
        }
    }
    for (int j = 0; j < 3; ++j) {
        const float e1 = Dot(d, q.ax[j]);
        const float e2 = ((p.h.x * ARf[0][j] + p.h.y * ARf[1][j]) + p.h.z * ARf[2][j]) + Elem(q.h, j);
        const float s = Abs(e1) - e2;
        if (s > threshold) return false;
        if (s > best) {
            best = s;
            axis = q.ax[j];
            flip = e1 < 0.0f;
            bf = ObbFeature{kObbFaceB, 0, static_cast<uint8_t>(j)};
        }
    }
    for (int i = 0; i < 3; ++i) {
        const int i1 = (i + 1) % 3, i2 = (i + 2) % 3;
        const int plo = i == 0 ? 1 : 0, phi = i == 2 ? 1 : 2;
        for (int j = 0; j < 3; ++j) {
            const int qlo = j == 0 ? 1 : 0, qhi = j == 2 ? 1 : 2;
            const Vec3 n = i == 0 ? Vec3(0.0f, -R[2][j], R[1][j])
                                  : (i == 1 ? Vec3(R[2][j], 0.0f, -R[0][j]) : Vec3(-R[1][j], R[0][j], 0.0f));
            const float e1 = t[i2] * R[i1][j] - t[i1] * R[i2][j];
            const float e2 = ((Elem(p.h, plo) * AR[phi][j] + Elem(p.h, phi) * AR[plo][j]) + Elem(q.h, qlo) * AR[i][qhi]) +
                             Elem(q.h, qhi) * AR[i][qlo];
            const float s_raw = Abs(e1) - e2;
            if (s_raw > kEdgeEps) return false;
            const float len = sqrtf((n.x * n.x + n.y * n.y) + n.z * n.z);
            if (len <= kEdgeEps) continue;
            const float s = s_raw / len;
            if (s * 1.05f > best) {
                best = s;
                const Vec3 nc(n.x / len, n.y / len, n.z / len);
                axis = Vec3((p.ax[0].x * nc.x + p.ax[1].x * nc.y) + p.ax[2].x * nc.z,
                            (p.ax[0].y * nc.x + p.ax[1].y * nc.y) + p.ax[2].y * nc.z,
                            (p.ax[0].z * nc.x + p.ax[1].z * nc.y) + p.ax[2].z * nc.z);
                flip = e1 < 0.0f;
                bf = ObbFeature{kObbEdges, static_cast<uint8_t>(i), static_cast<uint8_t>(j)};
            }
        }
    }
    normal = flip ? -axis : axis;
    overlap = -best;
    f = bf;
    return true;
}

// The game's query order: the test runs with b as the first box; the normal is turned to point from a toward b and
// the feature's boxes are swapped back.
RG_HD bool CarBoxContact(const Obb& a, const Obb& b, Vec3& normal, float& overlap, ObbFeature& f) {
    Vec3 n;
    ObbFeature g;
    if (!ObbLeastOverlap(b, a, 0.0f, n, overlap, g)) return false;
    normal = -n;
    f.kind = g.kind == kObbFaceA ? kObbFaceB : (g.kind == kObbFaceB ? kObbFaceA : g.kind);
    f.axis_a = g.axis_b;
    f.axis_b = g.axis_a;
    return true;
}

// Clips a quadrilateral (four u, v points) to the rectangle |u| < h[0], |v| < h[1], one half-plane at a time
// (u < h0, u > -h0 ... in the order -u, +u, -v, +v sides); stops as soon as a pass has produced eight points.
RG_HD int ClipQuadToRect(const float h[2], const float quad[8], float out[16]) {
    float buf[16];
    const float* src = quad;
    float* dst = out;
    int n_in = 4, n_out = 0;
    for (int axis = 0; axis < 2; ++axis) {
        for (int side = 0; side < 2; ++side) {
            const float s = side == 0 ? -1.0f : 1.0f;
            const float bound = s * h[axis];
            n_out = 0;
            for (int k = 0; k < n_in; ++k) {
                const float* cur = src + 2 * k;
                const float* nxt = k + 1 < n_in ? cur + 2 : src;
                const bool in_cur = s * cur[axis] < h[axis];
                if (in_cur) {
                    dst[2 * n_out] = cur[0];
                    dst[2 * n_out + 1] = cur[1];
                    if (++n_out == 8) goto full;
                }
                if (in_cur != (s * nxt[axis] < h[axis])) {
                    const int o = 1 - axis;
                    dst[2 * n_out + o] = cur[o] + (nxt[o] - cur[o]) / (nxt[axis] - cur[axis]) * (bound - cur[axis]);
                    dst[2 * n_out + axis] = bound;
                    if (++n_out == 8) goto full;
                }
            }
            src = dst;
            dst = src == out ? buf : out;
            n_in = n_out;
        }
    }
    if (src != out) {
        for (int k = 0; k < 2 * n_out; ++k) out[k] = src[k];
    }
    return n_out;
full:
    if (dst != out) {
        for (int k = 0; k < 16; ++k) out[k] = dst[k];
    }
    return 8;
}

// Picks `keep` of `count` (u, v) points: the deepest first, then for each further slot the remaining point whose
// angle about the polygon's centroid is nearest the deepest's angle plus slot * 2 pi / keep.
RG_HD void SpreadPoints(int count, const float* pts, int keep, int deepest, int* chosen) {
    float cx, cy;
    if (count == 1) {
        cx = pts[0];
        cy = pts[1];
    } else if (count == 2) {
        cx = 0.5f * (pts[0] + pts[2]);
        cy = 0.5f * (pts[1] + pts[3]);
    } else {
        float area = 0.0f, sx = 0.0f, sy = 0.0f;
        for (int i = 0; i + 1 < count; ++i) {
            const float q = pts[2 * i] * pts[2 * i + 3] - pts[2 * i + 2] * pts[2 * i + 1];
            area += q;
            sx += q * (pts[2 * i] + pts[2 * i + 2]);
            sy += q * (pts[2 * i + 1] + pts[2 * i + 3]);
        }
        const float q = pts[2 * count - 2] * pts[1] - pts[0] * pts[2 * count - 1];
        const float den = area + q;
        const float scale = Abs(den) > 1.1920929e-7f ? 1.0f / (3.0f * den) : 1.0e18f;
        cx = scale * (sx + q * (pts[2 * count - 2] + pts[0]));
        cy = scale * (sy + q * (pts[2 * count - 1] + pts[1]));
    }
    constexpr float kPi = 3.14159265f;
    float ang[8];
    bool free_pt[8];
    for (int i = 0; i < count; ++i) {
        ang[i] = SafeAtan2(pts[2 * i + 1] - cy, pts[2 * i] - cx);
        free_pt[i] = true;
    }
    free_pt[deepest] = false;
    chosen[0] = deepest;
    const float step = (2.0f * kPi) / static_cast<float>(keep);
    for (int slot = 1; slot < keep; ++slot) {
        float target = static_cast<float>(slot) * step + ang[deepest];
        if (target > kPi) target -= 2.0f * kPi;
        float best = 1.0e9f;
        int pick = deepest;
        for (int i = 0; i < count; ++i) {
            if (!free_pt[i]) continue;
            float diff = Abs(ang[i] - target);
            if (diff > kPi) diff = 2.0f * kPi - diff;
            if (diff < best) {
                best = diff;
                pick = i;
            }
        }
        free_pt[pick] = false;
        chosen[slot] = pick;
    }
}

// Face contact: the face of `ref` on axis `ref_axis` that faces `dir`, the face of `inc` most anti-parallel to it,
// the incident face clipped to the reference face's rectangle, the clipped points below the reference face kept
// (up to four, spread about the deepest).  Outputs, per point, the point on car a (the reference face's point when
// ref is a's box, else the incident point) and minus its depth.
RG_HD int FaceContactPoints(const Obb& ref, int ref_axis, const Obb& inc, const Vec3& dir, bool ref_is_a,
                            Vec3* pts_on_a, float* depths) {
    const float sgn = Dot(ref.ax[ref_axis], dir) >= 0.0f ? 1.0f : -1.0f;
    const Vec3 rn = ref.ax[ref_axis] * sgn;
    const float nr[3] = {Dot(inc.ax[0], rn), Dot(inc.ax[1], rn), Dot(inc.ax[2], rn)};
    const float anr[3] = {Abs(nr[0]), Abs(nr[1]), Abs(nr[2])};
    int lanr, a1, a2;
    if (anr[1] > anr[0]) {
        if (anr[1] > anr[2]) {
            lanr = 1, a1 = 0, a2 = 2;
        } else {
            lanr = 2, a1 = 0, a2 = 1;
        }
    } else if (anr[0] > anr[2]) {
        lanr = 0, a1 = 1, a2 = 2;
    } else {
        lanr = 2, a1 = 0, a2 = 1;
    }
    const Vec3 rel = inc.c - ref.c;
    const Vec3 lift = inc.ax[lanr] * Elem(inc.h, lanr);
    const Vec3 center = nr[lanr] < 0.0f ? rel + lift : rel - lift;
    const int code1 = ref_axis == 0 ? 1 : 0;
    const int code2 = ref_axis == 2 ? 1 : 2;
    const float c1 = Dot(center, ref.ax[code1]);
    const float c2 = Dot(center, ref.ax[code2]);
    const float m11 = Dot(ref.ax[code1], inc.ax[a1]);
    const float m12 = Dot(ref.ax[code1], inc.ax[a2]);
    const float m21 = Dot(ref.ax[code2], inc.ax[a1]);
    const float m22 = Dot(ref.ax[code2], inc.ax[a2]);
    const float k1 = m11 * Elem(inc.h, a1), k2 = m21 * Elem(inc.h, a1);
    const float k3 = m12 * Elem(inc.h, a2), k4 = m22 * Elem(inc.h, a2);
    const float quad[8] = {c1 - k1 - k3, c2 - k2 - k4, c1 - k1 + k3, c2 - k2 + k4,
                           c1 + k1 + k3, c2 + k2 + k4, c1 + k1 - k3, c2 + k2 - k4};
    const float rect[2] = {Elem(ref.h, code1), Elem(ref.h, code2)};
    float uv[16];
    const int n = ClipQuadToRect(rect, quad, uv);
    if (n < 1) return 0;
    const float det = m11 * m22 - m12 * m21;
    if (Abs(det) <= 1.0e-8f) return 0;
    const float inv = 1.0f / det;
    const float im11 = m11 * inv, im12 = m12 * inv, im21 = m21 * inv, im22 = m22 * inv;
    Vec3 local[8];
    float depth[8];
    int kept = 0;
    for (int i = 0; i < n; ++i) {
        const float u = uv[2 * i], v = uv[2 * i + 1];
        const float s1 = im22 * (u - c1) - im12 * (v - c2);
        const float s2 = -im21 * (u - c1) + im11 * (v - c2);
        const Vec3 pt = (center + inc.ax[a1] * s1) + inc.ax[a2] * s2;
        const float dp = Elem(ref.h, ref_axis) - Dot(rn, pt);
        if (dp >= 0.0f) {
            local[kept] = pt;
            depth[kept] = dp;
            uv[2 * kept] = u;
            uv[2 * kept + 1] = v;
            ++kept;
        }
    }
    if (kept < 1) return 0;
    int chosen[8];
    int out = kept;
    if (kept <= 4) {
        for (int i = 0; i < kept; ++i) chosen[i] = i;
    } else {
        int deepest = 0;
        for (int i = 1; i < kept; ++i) {
            if (depth[i] > depth[deepest]) deepest = i;
        }
        out = 4;
        SpreadPoints(kept, uv, 4, deepest, chosen);
    }
    for (int i = 0; i < out; ++i) {
        const int k = chosen[i];
        const Vec3 incident = ref.c + local[k];
        pts_on_a[i] = ref_is_a ? incident + rn * depth[k] : incident;
        depths[i] = -depth[k];
    }
    return out;
}

// Edge-edge contact of edge axis `axis_b` of b's box and `axis_a` of a's box: the normal of the two edges (from b
// toward a), the closest point of a's edge to b's edge (on the supporting edges of each box along that normal) and
// the separation along the normal less both boxes' projected extents (with the 1e-5 fudge).  False when the edges
// are parallel.
RG_HD bool EdgeContactPoint(const Obb& a, int axis_a, const Obb& b, int axis_b, Vec3& normal, Vec3& point_on_a,
                            float& depth) {
    const int i = axis_b, j = axis_a;
    const Vec3 p = a.c - b.c;
    const float pp[3] = {Dot(b.ax[0], p), Dot(b.ax[1], p), Dot(b.ax[2], p)};
    const float r[3] = {Dot(b.ax[0], a.ax[j]), Dot(b.ax[1], a.ax[j]), Dot(b.ax[2], a.ax[j])};
    float e1;
    Vec3 n;
    if (i == 0) {
        e1 = pp[2] * r[1] - pp[1] * r[2];
        n = Vec3(0.0f, -r[2], r[1]);
    } else if (i == 1) {
        e1 = pp[0] * r[2] - pp[2] * r[0];
        n = Vec3(r[2], 0.0f, -r[0]);
    } else {
        e1 = pp[1] * r[0] - pp[0] * r[1];
        n = Vec3(-r[1], r[0], 0.0f);
    }
    const float len = sqrtf((n.x * n.x + n.y * n.y) + n.z * n.z);
    if (!(len > 1.1920929e-07f)) return false;
    {
        const int lo = i == 0 ? 1 : 0, hi = i == 2 ? 1 : 2;
        const int k1 = j == 0 ? 1 : 0, k2 = j == 2 ? 1 : 2;
        const float q_hi = Abs(r[hi]) + 1.0e-5f;
        const float q_lo = Abs(r[lo]) + 1.0e-5f;
        const float q_k2 = Abs(Dot(b.ax[i], a.ax[k2])) + 1.0e-5f;
        const float q_k1 = Abs(Dot(b.ax[i], a.ax[k1])) + 1.0e-5f;
        const float e2 = ((Elem(b.h, lo) * q_hi + Elem(b.h, hi) * q_lo) + Elem(a.h, k1) * q_k2) + Elem(a.h, k2) * q_k1;
        depth = (Abs(e1) - e2) / len;
    }
    const Vec3 nc(n.x / len, n.y / len, n.z / len);
    Vec3 nw((b.ax[0].x * nc.x + b.ax[1].x * nc.y) + b.ax[2].x * nc.z,
            (b.ax[0].y * nc.x + b.ax[1].y * nc.y) + b.ax[2].y * nc.z,
            (b.ax[0].z * nc.x + b.ax[1].z * nc.y) + b.ax[2].z * nc.z);
    if (e1 < 0.0f) nw = -nw;
    Vec3 pb = b.c;
    for (int k = 0; k < 3; ++k) pb += b.ax[k] * ((Dot(nw, b.ax[k]) > 0.0f ? 1.0f : -1.0f) * Elem(b.h, k));
    Vec3 pa = a.c;
    for (int k = 0; k < 3; ++k) pa += a.ax[k] * ((Dot(nw, a.ax[k]) > 0.0f ? -1.0f : 1.0f) * Elem(a.h, k));
    // Closest approach of the lines pb + s ub and pa + t ua: only a's parameter is needed.
    const Vec3 ub = b.ax[i], ua = a.ax[j];
    const Vec3 w = pa - pb;
    const float uu = Dot(ub, ua);
    const float q1 = Dot(ub, w);
    const float q2 = -Dot(ua, w);
    float den = 1.0f - uu * uu;
    float t = 0.0f;
    if (!(den <= 1.0e-4f)) {
        den = 1.0f / den;
        t = (uu * q1 + q2) * den;
    }
    normal = nw;
    point_on_a = pa + ua * t;
    return true;
}

// ------------------------------------------------------------------------------------------------------------
// The pair's manifold.

struct PairPoint {
    Vec3 local_a, local_b;  // body frames
    Vec3 pos_a, pos_b;      // world (bt), from the last refresh
    Vec3 normal;            // from a toward b
    float distance;         // from a's point to b's along the normal; negative when the cars overlap
};

struct PairWork {
    int count;
    PairPoint p[4];
};

// Moves the points with the bodies and drops those that separated beyond the threshold or slid apart sideways
// (Bullet's refreshContactPoints; a dropped point takes the last point's place).
RG_HD void RefreshPair(PairWork& w, const gjk::Xf& xa, const gjk::Xf& xb, float threshold) {
    if (w.count <= 0) return;
    for (int i = 0; i < w.count; ++i) {
        PairPoint& q = w.p[i];
        q.pos_a = BodyToWorld(xa, q.local_a);
        q.pos_b = BodyToWorld(xb, q.local_b);
        q.distance = Dot(q.pos_b - q.pos_a, q.normal);
    }
    const float thr2 = threshold * threshold;
    int kept = w.count;
    for (int i = w.count - 1; i >= 0; --i) {
        const PairPoint q = w.p[i];
        bool drop = q.distance > threshold;
        if (!drop) {
            const Vec3 diff = q.pos_a - (q.pos_b - q.normal * q.distance);
            drop = Dot(diff, diff) > thr2;
        }
        if (drop) {
            w.p[i] = w.p[kept - 1];
            --kept;
        }
    }
    w.count = kept;
}

RG_HD float QuadArea2(const Vec3& l, const Vec3& p, const Vec3& q, const Vec3& r) {
    const Vec3 c = Cross(l - p, q - r);
    return Dot(c, c);
}

// Adds a point: it replaces the point nearest by its point on b (within the threshold); else it is appended; else
// (full) it replaces the point whose removal leaves the largest quadrilateral, the deepest point never replaced.
RG_HD void AddPairPoint(PairWork& w, const PairPoint& np, float threshold) {
    const int count = w.count;
    float shortest = threshold * threshold;
    int nearest = -1;
    for (int i = 0; i < count; ++i) {
        const Vec3 d = w.p[i].local_b - np.local_b;
        const float d2 = Dot(d, d);
        if (d2 < shortest) {
            shortest = d2;
            nearest = i;
        }
    }
    if (nearest >= 0) {
        w.p[nearest] = np;
        return;
    }
    if (count < 4) {
        w.p[count] = np;
        w.count = count + 1;
        return;
    }
    int deepest = 4;
    float deepest_d = np.distance;
    for (int i = 0; i < 4; ++i) {
        if (w.p[i].distance < deepest_d) {
            deepest = i;
            deepest_d = w.p[i].distance;
        }
    }
    const Vec3& l = np.local_b;
    const Vec3 &p0 = w.p[0].local_b, &p1 = w.p[1].local_b, &p2 = w.p[2].local_b, &p3 = w.p[3].local_b;
    const float r0 = deepest == 0 ? 0.0f : QuadArea2(l, p1, p3, p2);
    const float r1 = deepest == 1 ? 0.0f : QuadArea2(l, p0, p3, p2);
    const float r2 = deepest == 2 ? 0.0f : QuadArea2(l, p0, p3, p1);
    const float r3 = deepest == 3 ? 0.0f : QuadArea2(l, p0, p2, p1);
    int pick = 0;
    float most = r0;
    if (r1 > most) pick = 1, most = r1;
    if (r2 > most) pick = 2, most = r2;
    if (r3 > most) pick = 3;
    w.p[pick] = np;
}

RG_HD void LoadPair(const CarPairManifold& m, PairWork& w) {
    w.count = m.count <= 4 ? m.count : 4;
    for (int i = 0; i < w.count; ++i) {
        w.p[i].local_a = m.pts[i].local_a;
        w.p[i].local_b = m.pts[i].local_b;
        w.p[i].normal = m.pts[i].normal;
    }
}

RG_HD void StorePair(const PairWork& w, CarPairManifold& m) {
    m.count = static_cast<uint8_t>(w.count);
    for (int i = 0; i < w.count; ++i) {
        m.pts[i].local_a = w.p[i].local_a;
        m.pts[i].local_b = w.p[i].local_b;
        m.pts[i].normal = w.p[i].normal;
    }
}

// ------------------------------------------------------------------------------------------------------------
// Bumps.

// The game's rotator axes (forward, right, up) of an orientation.
RG_HD void Ue3RotatorAxes(const Ue3Rotator& r, Vec3& x, Vec3& y, Vec3& z) {
    const float sp = Ue3RotatorSin(r.pitch), cp = Ue3RotatorCos(r.pitch);
    const float sy = Ue3RotatorSin(r.yaw), cy = Ue3RotatorCos(r.yaw);
    const float sr = Ue3RotatorSin(r.roll), cr = Ue3RotatorCos(r.roll);
    const float sp_sr = sp * sr, cr_sp = cr * sp;
    x = Vec3(cy * cp, cp * sy, sp);
    y = Vec3(sp_sr * cy - cr * sy, cy * cr + sp_sr * sy, -sr * cp);
    z = Vec3(-(cr_sp * cy + sy * sr), cy * sr - cr_sp * sy, cp * cr);
}

// Is the victim's hitbox center in front of the hitter's: the direction between the centers, projected off the
// hitter's right axis and then off its up axis (each projection shortened by 1 %), within 36.87 degrees of forward in
// the first and `max_yaw_deg` in the second.
RG_HD bool InForwardCone(const CarState& h, const CarConfig& hc, const CarState& v, const CarConfig& vc,
                         float max_yaw_deg) {
    Vec3 fwd, right, up, vx, vy, vz;
    Ue3RotatorAxes(Ue3QuatToRotator(h.rot), fwd, right, up);
    Ue3RotatorAxes(Ue3QuatToRotator(v.rot), vx, vy, vz);
    const Vec3 ho = hc.hitbox_offset, vo = vc.hitbox_offset;
    const Vec3 ph = h.pos + Vec3(Dot(fwd, ho), Dot(right, ho), Dot(up, ho));
    const Vec3 pv = v.pos + Vec3(Dot(vx, vo), Dot(vy, vo), Dot(vz, vo));
    const auto unit = [](Vec3 d) {
        const float d2 = LengthSq(d);
        if (d2 != 1.0f) d = d2 < 1e-8f ? Vec3(0.0f, 0.0f, 0.0f) : d * (1.0f / sqrtf(d2));
        return d;
    };
    const Vec3 dir = unit(pv - ph);
    const auto angle_deg = [&](const Vec3& axis) {
        float k = Dot(dir, axis);
        k = k >= 0.0f ? k / 1.01f : k * 1.01f;
        const Vec3 p = unit(dir - axis * k);
        float c = Dot(fwd, p);
        c = c < -1.0f ? -1.0f : (c > 1.0f ? 1.0f : c);
        return UcrtAcosf(c) * 57.29578f;
    };
    if (angle_deg(right) > 36.8698959f) return false;
    return angle_deg(up) <= max_yaw_deg;
}

RG_HD Vec3 SafeUnit(const Vec3& v) {
    const float l2 = LengthSq(v);
    const float len = l2 > 0.0f ? sqrtf(l2) : 0.0f;
    constexpr float kEps2 = 1.1920928955078125e-7f * 1.1920928955078125e-7f;
    if (len > kEps2) return Vec3(v.x / len, v.y / len, v.z / len);
    return Vec3(0.0f, 0.0f, 0.0f);
}

// One direction of a car-car touch: the hitter, moving toward the victim faster than the victim moves away (both at
// their tick-start velocities), bumps it when the victim is in its forward cone and the pair is not cooling down.  A
// supersonic hitter fast enough along its nose may demolish instead; that test sweeps both hitboxes over the tick
// (time of impact), which this build does not step, so such a touch marks the arena.
RG_HD void CarBumpAttempt(CarState& hitter, uint32_t hitter_id, uint8_t hitter_team, const CarConfig& hitter_cfg,
                          const Vec3& hitter_vel, CarState& victim, uint32_t victim_id, uint8_t victim_team,
                          const CarConfig& victim_cfg, const Vec3& victim_vel, const CarParams& m, float dt,
                          uint32_t& unsupported) {
    if (victim.is_demoed || hitter.just_respawned || victim.just_respawned) return;
    const bool cooling = hitter.car_contact.other_car_id == victim_id && hitter.car_contact.cooldown_timer > 0.0f;
    const float speed = sqrtf(LengthSq(hitter_vel));
    if (!(speed > 0.0f)) return;
    const float toward = Dot(SafeUnit(victim.pos - hitter.pos), hitter_vel);
    if (toward <= 0.0f) return;
    const Vec3 vel_dir = hitter_vel * (1.0f / speed);
    if (Dot(victim_vel, vel_dir) >= toward) return;
    if (m.enable_demos && hitter.is_supersonic && (m.enable_team_demos || hitter_team != victim_team)) {
        Vec3 fx, fy, fz;
        Ue3RotatorAxes(Ue3QuatToRotator(hitter.rot), fx, fy, fz);
        if (fabsf(Dot(hitter_vel, fx)) >= m.supersonic_maintain_min_speed) unsupported |= kUnsupportedCarCar;
    }
    if (cooling) return;
    const bool bumper = InForwardCone(hitter, hitter_cfg, victim, victim_cfg, 70.0f);
    if (bumper) {
        const auto curve = [speed](float y0, float y1, float y2) {
            if (speed <= 0.0f) return y0;
            if (speed < 1400.0f) return y0 + (y1 - y0) * (speed / 1400.0f);
            if (speed < 2200.0f) return y1 + (y2 - y1) * ((speed - 1400.0f) / 800.0f);
            return y2;
        };
        Vec3 impulse;
        if (victim.is_on_ground != 0) {
            const Vec3 up = Ue3RotatorUpAxis(Ue3QuatToRotator(victim.rot));
            impulse = vel_dir * curve(150.0f, 200000.0f, 275000.0f) + up * curve(50.0f, 50000.0f, 75000.0f);
        } else {
            impulse = vel_dir * curve(150.0f, 250000.0f, 350000.0f);
        }
        const Vec3 ibt = ((impulse * m.bump_force_scale) * (1.0f / 50.0f)) * (1.0f / m.car_mass);
        victim.vel_impulse_cache_bt += ibt;
        victim.vel_impulse_cache += ibt * 50.0f;
        if (victim.pending_bump_count < CarState::kMaxPendingBumps) {
            int slot = victim.pending_bump_count++;
            for (; slot > 0 && victim.pending_bump_hitter[slot - 1] > hitter_id; --slot) {
                victim.pending_bump_bt[slot] = victim.pending_bump_bt[slot - 1];
                victim.pending_bump_hitter[slot] = victim.pending_bump_hitter[slot - 1];
            }
            victim.pending_bump_bt[slot] = ibt;
            victim.pending_bump_hitter[slot] = hitter_id;
        }
    }
    hitter.car_contact.other_car_id = victim_id;
    hitter.car_contact.cooldown_timer = m.bump_cooldown_time + dt * 0.5f;
    if (bumper) {
        const uint32_t k = ((victim_id - hitter_id) & 7u) - 1u;
        hitter.bump_event_bumps[k] = static_cast<uint8_t>(hitter.bump_event_bumps[k] + 1u);
    }
}

// ------------------------------------------------------------------------------------------------------------
// Pair rows.

// A row between cars a and b (bt units): car a is pushed along `dir`, car b along -dir.
struct CarPairRow {
    Vec3 dir;
    Vec3 rxa, rxb;    // contact offset x direction, for a (with dir) and for b (with -dir)
    Vec3 anga, angb;  // world inverse inertia times those
    float jac, rhs, applied;
    int8_t a, b;        // car slots
    int8_t normal_row;  // friction rows: their non-penetration row
    int8_t pad;
};

// A pair's split row: the direction terms and denominator of its point's non-penetration row.
struct CarPairSplitRow {
    float rhs;
    float applied;
    int normal_row;
};

constexpr int kMaxPairRows = 8;
constexpr float kCarCarFriction = 0.09f;

struct CarPairRows {
    CarPairRow n[kMaxPairRows];       // non-penetration
    CarPairRow f[kMaxPairRows];       // friction
    CarPairSplitRow s[kMaxPairRows];  // split impulse
    int nn, nf, ns;
    uint32_t cars;  // slots with pair rows
};

RG_HD void CarInvInertiaBt(const CarState& car, const CarConfig& cfg, float mass, Vec3& r0, Vec3& r1, Vec3& r2) {
    if (car.restored_inv_inertia_valid) {
        r0 = car.restored_inv_inertia_bt[0];
        r1 = car.restored_inv_inertia_bt[1];
        r2 = car.restored_inv_inertia_bt[2];
        return;
    }
    InvInertiaWorldRows(car.tick_start_basis_forward, car.tick_start_basis_right, car.tick_start_basis_up,
                        InvInertiaLocalBt(cfg, mass), r0, r1, r2);
}

RG_HD Vec3 PlaneTangent(const Vec3& n) {
    if (Abs(n.z) > 0.7071067811865475244008443621048490f) {
        const float k = 1.0f / sqrtf(n.y * n.y + n.z * n.z);
        return Vec3(0.0f, -n.z * k, n.y * k);
    }
    const float k = 1.0f / sqrtf(n.x * n.x + n.y * n.y);
    return Vec3(-n.y * k, n.x * k, 0.0f);
}

// Rows for the pair's manifold points.  The non-penetration row's target is the restitution of the velocity
// without the step's external impulses; its and the friction row's current velocities include them (angular ones
// only for the non-penetration row).  The split row exists when the point penetrates in game units; this build
// solves it in Bullet units, which needs a penetration there too (else the arena is marked).
RG_HD void EmitPairRows(CarState& a, const CarConfig& ca, const CarPreContact& pa, int slot_a, CarState& b,
                        const CarConfig& cb, const CarPreContact& pb, int slot_b, const PairWork& w,
                        const CarParams& m, float dt, CarPairRows& out, uint32_t& unsupported) {
    const int count = w.count < 4 ? w.count : 4;
    if (count <= 0) return;
    float deepest_uu = 0.0f;
    for (int i = 0; i < count; ++i) deepest_uu = Max(deepest_uu, -w.p[i].distance * 50.0f);
    const float penetration = Max(0.0f, deepest_uu);
    Vec3 ia0, ia1, ia2, ib0, ib1, ib2;
    CarInvInertiaBt(a, ca, m.car_mass, ia0, ia1, ia2);
    CarInvInertiaBt(b, cb, m.car_mass, ib0, ib1, ib2);
    const float inv_m = 1.0f / m.car_mass;
    const Vec3 lin_a = pa.vel_bt, lin_b = pb.vel_bt;
    const Vec3 ext_ang_a = a.ext_torque_impulse, ext_ang_b = b.ext_torque_impulse;
    const Vec3 base_lin_a = lin_a + a.ext_force_impulse_bt;
    const Vec3 base_lin_b = lin_b + b.ext_force_impulse_bt;
    for (int idx = 0; idx < count; ++idx) {
        const PairPoint& q = w.p[idx];
        const Vec3 npt = q.normal;
        const Vec3 ra = q.pos_a - a.pos_bt, rb = q.pos_b - b.pos_bt;
        const Vec3 na = -npt;
        const Vec3 rxa = Cross(ra, na);
        const Vec3 anga = MulRows(ia0, ia1, ia2, rxa);
        const float den_a = inv_m + Dot(na, Cross(anga, ra));
        const Vec3 rxb = Cross(rb, -na);
        const Vec3 angb = MulRows(ib0, ib1, ib2, rxb);
        const float den_b = inv_m + Dot(na, Cross(-angb, rb));
        const float den = den_a + den_b;
        int normal_idx = -1;
        if (!(den <= 1.0e-6f)) {
            if (out.nn >= kMaxPairRows) {
                unsupported |= kUnsupportedCarCar;
            } else {
                const float v1 = Dot(na, base_lin_a) + Dot(rxa, pa.ang_vel + ext_ang_a);
                const float v2 = Dot(npt, base_lin_b) + Dot(rxb, pb.ang_vel + ext_ang_b);
                const float jac = 1.0f / den;
                const Vec3 va = lin_a + Cross(pa.ang_vel, ra);
                const Vec3 vb = lin_b + Cross(pb.ang_vel, rb);
                const float rest = Max(0.1f * -Dot(npt, vb - va), 0.0f);
                CarPairRow& r = out.n[out.nn];
                r.dir = na;
                r.rxa = rxa;
                r.rxb = rxb;
                r.anga = anga;
                r.angb = angb;
                r.jac = jac;
                r.rhs = (rest - (v1 + v2)) * jac;
                r.applied = 0.0f;
                r.a = static_cast<int8_t>(slot_a);
                r.b = static_cast<int8_t>(slot_b);
                r.normal_row = -1;
                normal_idx = out.nn++;
                // Friction along the sliding direction at the contact (with the external impulses).
                const Vec3 bva = base_lin_a + Cross(pa.ang_vel + ext_ang_a, ra);
                const Vec3 bvb = base_lin_b + Cross(pb.ang_vel + ext_ang_b, rb);
                const Vec3 rel = bvb - bva;
                Vec3 tan = rel - npt * Dot(rel, npt);
                const float lat = LengthSq(tan);
                const float tl = sqrtf(lat);
                tan = lat > 1.1920928955078125e-07f ? tan * (1.0f / tl) : PlaneTangent(npt);
                const Vec3 ta = tan, tb = -tan;
                const Vec3 rxta = Cross(ra, ta), rxtb = Cross(rb, tb);
                const Vec3 angta = MulRows(ia0, ia1, ia2, rxta);
                const Vec3 angtb = MulRows(ib0, ib1, ib2, rxtb);
                const float dta = Dot(ta, Cross(angta, ra));
                const float dtb = Dot(ta, Cross(-angtb, rb));
                const float den_t = (inv_m + dta) + (inv_m + dtb);
                if (!(den_t <= 1.0e-6f)) {
                    if (out.nf >= kMaxPairRows) {
                        unsupported |= kUnsupportedCarCar;
                    } else {
                        CarPairRow& fr = out.f[out.nf++];
                        fr.dir = ta;
                        fr.rxa = rxta;
                        fr.rxb = rxtb;
                        fr.anga = angta;
                        fr.angb = angtb;
                        fr.jac = 1.0f / den_t;
                        const float t1 = Dot(ta, base_lin_a) + Dot(rxta, pa.ang_vel);
                        const float t2 = Dot(tb, base_lin_b) + Dot(rxtb, pb.ang_vel);
                        fr.rhs = -(t1 + t2) * fr.jac;
                        fr.applied = 0.0f;
                        fr.a = static_cast<int8_t>(slot_a);
                        fr.b = static_cast<int8_t>(slot_b);
                        fr.normal_row = static_cast<int8_t>(normal_idx);
                    }
                }
            }
        }
        if (penetration > 0.0f && dt > 0.0f) {
            const Vec3 pa_uu = q.pos_a * 50.0f, pb_uu = q.pos_b * 50.0f;
            const float pen_uu = Max(0.0f, -Dot(pb_uu - pa_uu, npt));
            if (!(pen_uu <= 0.0f)) {
                // The row as the game units see it (its existence only).
                Vec3 ua0, ua1, ua2, ub0, ub1, ub2;
                InvInertiaWorldRowsUu(a, ca, m.car_mass, ua0, ua1, ua2);
                InvInertiaWorldRowsUu(b, cb, m.car_mass, ub0, ub1, ub2);
                const Vec3 sra = (pa_uu * (1.0f / 50.0f) - a.pos_bt) * 50.0f;
                const Vec3 srb = (pb_uu * (1.0f / 50.0f) - b.pos_bt) * 50.0f;
                const Vec3 uanga = MulRows(ua0, ua1, ua2, Cross(sra, na));
                const Vec3 uangb = MulRows(ub0, ub1, ub2, Cross(srb, npt));
                const float uden = (inv_m + Dot(na, Cross(uanga, sra))) + (inv_m + Dot(npt, Cross(uangb, srb)));
                if (!(uden <= 1.0e-8f)) {
                    const float ujac = 1.0f / uden;
                    const float urhs = ((pen_uu * 0.8f) / dt) * ujac;
                    if (urhs > 0.0f && ujac > 0.0f) {
                        // In Bullet units the row's denominator equals the non-penetration row's (the same terms,
                        // b's with both signs flipped), so it shares that row's direction terms and inverse.
                        const float pen_bt = Max(0.0f, -Dot(q.pos_b - q.pos_a, npt));
                        const float den_bt = den_a + (inv_m + Dot(npt, Cross(angb, rb)));
                        const float jac_bt = den_bt > 1.0e-8f ? 1.0f / den_bt : 0.0f;
                        if (!(jac_bt > 0.0f && pen_bt > 0.0f) || normal_idx < 0 || out.ns >= kMaxPairRows ||
                            FloatBits(jac_bt) != FloatBits(out.n[normal_idx].jac)) {
                            unsupported |= kUnsupportedCarCar;
                        } else {
                            CarPairSplitRow& s = out.s[out.ns++];
                            s.rhs = ((pen_bt * 0.8f) / dt) * jac_bt;
                            s.applied = 0.0f;
                            s.normal_row = normal_idx;
                        }
                    }
                }
            }
        }
    }
    out.cars |= (1u << slot_a) | (1u << slot_b);
    // The pair solve writes the cars' velocities and origins back unchanged (a signed zero becomes +0).
    a.ang_vel += a.ang_vel - a.ang_vel;
    SetCarPosBt(a, a.pos_bt + (a.pos_bt - a.pos_bt));
    b.ang_vel += b.ang_vel - b.ang_vel;
    SetCarPosBt(b, b.pos_bt + (b.pos_bt - b.pos_bt));
}

// ------------------------------------------------------------------------------------------------------------
// One pair of cars: the manifold, the bumps, the rows.

struct CarPairSide {
    CarState* car;
    const CarConfig* cfg;
    const CarPreContact* pre;
    Vec3 start_vel;   // the car's velocity at the start of the tick (uu)
    uint32_t id;      // batch car id (slot + 1 past the arena's first)
    uint8_t team;
    int slot;
};

RG_HD void CarPairContact(const CarPairSide& sa, const CarPairSide& sb, CarPairManifold& stored, const CarParams& m,
                          float dt, CarPairRows& rows, uint32_t& unsupported) {
    CarState& a = *sa.car;
    CarState& b = *sb.car;
    if (a.is_demoed || b.is_demoed || a.just_respawned || b.just_respawned) {
        stored.count = 0;
        return;
    }
    if (!a.basis_authoritative || !b.basis_authoritative) {
        unsupported |= kUnsupportedCarCar;
        return;
    }
    const float reach = CarBoundingRadius(*sa.cfg) + CarBoundingRadius(*sb.cfg);
    const float threshold = Min(CarBreakingThresholdBt(*sa.cfg), CarBreakingThresholdBt(*sb.cfg));
    const gjk::Xf xa = CarBodyXf(a), xb = CarBodyXf(b);
    PairWork w;
    LoadPair(stored, w);
    RefreshPair(w, xa, xb, threshold);
    const bool near = LengthSq(b.pos - a.pos) < reach * reach;
    if (!near && w.count == 0) {
        StorePair(w, stored);
        return;
    }
    const Obb oa = CarObb(a, *sa.cfg), ob = CarObb(b, *sb.cfg);
    bool boxes_meet = true;
    for (int k = 0; k < 3; ++k) {
        float ea = 0.0f, eb = 0.0f;
        for (int i = 0; i < 3; ++i) {
            ea += Abs(Elem(oa.ax[i], k)) * Elem(oa.h, i);
            eb += Abs(Elem(ob.ax[i], k)) * Elem(ob.h, i);
        }
        const float ca = Elem(oa.c, k), cb = Elem(ob.c, k);
        boxes_meet = boxes_meet && !(ca - ea > cb + eb || ca + ea < cb - eb);
    }
    if (!boxes_meet) {
        stored.count = 0;
        return;
    }
    Vec3 normal;
    float overlap = 0.0f;
    ObbFeature feat{kObbEdges, 0, 0};
    const bool fresh = near && CarBoxContact(oa, ob, normal, overlap, feat);
    if (!fresh && w.count == 0) {
        StorePair(w, stored);
        return;
    }
    Vec3 pts_on_a[4];
    float depths[4];
    Vec3 pt_normal = normal;
    int n_pts = w.count;
    if (fresh) {
        if (feat.kind == kObbEdges) {
            Vec3 nba, p;
            float s;
            if (EdgeContactPoint(oa, feat.axis_a, ob, feat.axis_b, nba, p, s)) {
                pts_on_a[0] = p;
                depths[0] = s;
                pt_normal = -nba;
                n_pts = 1;
            } else {
                // Unreachable: the query only selects edge pairs that are not parallel.
                unsupported |= kUnsupportedCarCar;
                n_pts = 0;
            }
        } else {
            const bool face_a = feat.kind == kObbFaceA;
            n_pts = FaceContactPoints(face_a ? oa : ob, face_a ? feat.axis_a : feat.axis_b, face_a ? ob : oa,
                                      face_a ? normal : -normal, face_a, pts_on_a, depths);
        }
    }
    // Bumps, once per contact point and direction, the car with the larger id first.
    if (n_pts > 0) {
        const bool a_first = sa.id > sb.id;
        const CarPairSide& h1 = a_first ? sa : sb;
        const CarPairSide& v1 = a_first ? sb : sa;
        for (int k = 0; k < n_pts; ++k) {
            if (a.is_demoed && b.is_demoed) break;
            if (!v1.car->is_demoed && !h1.car->is_demoed) {
                CarBumpAttempt(*h1.car, h1.id, h1.team, *h1.cfg, h1.start_vel, *v1.car, v1.id, v1.team, *v1.cfg,
                               v1.start_vel, m, dt, unsupported);
            }
            if (a.is_demoed && b.is_demoed) break;
            if (!h1.car->is_demoed && !v1.car->is_demoed) {
                CarBumpAttempt(*v1.car, v1.id, v1.team, *v1.cfg, v1.start_vel, *h1.car, h1.id, h1.team, *h1.cfg,
                               h1.start_vel, m, dt, unsupported);
            }
        }
    }
    if (fresh && n_pts > 0) {
        for (int k = 0; k < n_pts; ++k) {
            if (depths[k] > threshold) continue;
            PairPoint np;
            np.normal = pt_normal;
            np.local_a = WorldToBody(xa, pts_on_a[k]);
            np.local_b = WorldToBody(xb, pts_on_a[k] + pt_normal * depths[k]);
            np.pos_a = np.pos_b = Vec3(0.0f, 0.0f, 0.0f);
            np.distance = depths[k];
            AddPairPoint(w, np, threshold);
        }
        RefreshPair(w, xa, xb, threshold);
    }
    StorePair(w, stored);
    EmitPairRows(a, *sa.cfg, *sa.pre, sa.slot, b, *sb.cfg, *sb.pre, sb.slot, w, m, dt, rows, unsupported);
}

// ------------------------------------------------------------------------------------------------------------
// The joint solve of the cars that touch each other.

// Visit order of the joint rows: the arena's contact-pair list (a car-mesh entry brings that car's rows for that
// mesh, a car-car entry that pair's rows), then the rows no entry claimed, mesh rows before pair rows.
struct GroupOrder {
    uint8_t kind[kMaxCarWorldRows + kMaxPairRows];  // 0 mesh row, 1 pair row
    uint8_t row[kMaxCarWorldRows + kMaxPairRows];
    int count;
};

// Rows are named through the non-penetration row they belong to: w_of(r) / p_of(r) give the mesh row r's and the
// pair row r's non-penetration row.
template <typename WorldOf, typename PairOf>
RG_HD void BuildGroupOrder(const ManifoldEntry* entries, int entry_count, int n_world, const CarWorldRows& w,
                           WorldOf w_of, int n_pair, const CarPairRows& p, PairOf p_of, GroupOrder& o) {
    uint32_t wused = 0u, pused = 0u;
    o.count = 0;
    for (int e = 0; e < entry_count; ++e) {
        const uint32_t key = entries[e].key;
        const int kind = KeyKind(key);
        if (kind == 2) {
            const int car = KeyBodyA(key);
            const int sid = KeyStatic(key);
            if (sid < 0 || sid >= 100) continue;
            for (int r = 0; r < n_world; ++r) {
                if (((wused >> r) & 1u) || w.n_car[w_of(r)] != car || w.n_mesh[w_of(r)] != sid) continue;
                wused |= 1u << r;
                o.kind[o.count] = 0;
                o.row[o.count++] = static_cast<uint8_t>(r);
            }
        } else if (kind == 3) {
            const int lo = KeyBodyA(key), hi = KeyBodyB(key);
            for (int r = 0; r < n_pair; ++r) {
                if ((pused >> r) & 1u) continue;
                const CarPairRow& pn = p.n[p_of(r)];
                const int plo = MinI(pn.a, pn.b), phi = MaxI(pn.a, pn.b);
                if (plo != lo || phi != hi) continue;
                pused |= 1u << r;
                o.kind[o.count] = 1;
                o.row[o.count++] = static_cast<uint8_t>(r);
            }
        }
    }
    for (int r = 0; r < n_world; ++r) {
        if (!((wused >> r) & 1u)) {
            o.kind[o.count] = 0;
            o.row[o.count++] = static_cast<uint8_t>(r);
        }
    }
    for (int r = 0; r < n_pair; ++r) {
        if (!((pused >> r) & 1u)) {
            o.kind[o.count] = 1;
            o.row[o.count++] = static_cast<uint8_t>(r);
        }
    }
}

RG_HD bool SolvePairVelocityRow(CarPairRow& r, float inv_mass, Vec3* dl, Vec3* da, float max_impulse, bool friction) {
    const int a = r.a, b = r.b;
    const float dv_b = Dot(-r.dir, dl[b]) + Dot(r.rxb, da[b]);
    const float dv_a = Dot(r.dir, dl[a]) + Dot(r.rxa, da[a]);
    float delta = r.rhs;
    delta = delta - dv_b * r.jac;
    delta = delta - dv_a * r.jac;
    const float before = r.applied;
    const float sum = before + delta;
    if (friction) {
        if (sum < -max_impulse) {
            delta = -max_impulse - before;
            r.applied = -max_impulse;
        } else if (sum > max_impulse) {
            delta = max_impulse - before;
            r.applied = max_impulse;
        } else {
            r.applied = sum;
        }
    } else if (sum < 0.0f) {
        delta = -before;
        r.applied = 0.0f;
    } else {
        r.applied = sum;
    }
    if (delta == 0.0f) return FloatBits(before) != FloatBits(r.applied);
    dl[a] += (r.dir * inv_mass) * delta;
    da[a] += r.anga * delta;
    dl[b] += (-r.dir * inv_mass) * delta;
    da[b] += r.angb * delta;
    return true;
}

RG_HD bool SolveWorldVelocityRow(CarWorldRow& row, float inv_mass, Vec3& dl, Vec3& da) {
    const float dv = Dot(row.normal, dl) + Dot(row.rxn, da);
    float delta = row.rhs;
    delta = delta - 0.0f * row.jac;
    delta = delta - dv * row.jac;
    const float before = row.applied;
    const float sum = before + delta;
    if (sum < 0.0f) {
        delta = -before;
        row.applied = 0.0f;
    } else {
        row.applied = sum;
    }
    if (delta == 0.0f) return FloatBits(before) != FloatBits(row.applied);
    dl += (row.normal * inv_mass) * delta;
    da += row.ang * delta;
    return true;
}

RG_HD bool SolveWorldFrictionRow(CarWorldFrictionRow& row, float max_impulse, float total, float inv_mass, Vec3& dl,
                                  Vec3& da) {
    if (total <= 0.0f) return false;
    float delta = row.rhs;
    const float dv = Dot(row.tangent, dl) + Dot(row.rxt, da);
    delta = delta - dv * row.jac;
    const float before = row.applied;
    const float sum = before + delta;
    if (sum < -max_impulse) {
        delta = -max_impulse - before;
        row.applied = -max_impulse;
    } else if (sum > max_impulse) {
        delta = max_impulse - before;
        row.applied = max_impulse;
    } else {
        row.applied = sum;
    }
    if (delta == 0.0f) return FloatBits(before) != FloatBits(row.applied);
    dl += (row.tangent * inv_mass) * delta;
    da += row.ang * delta;
    return true;
}

// Velocity solve of the joint rows: non-penetration rows in visit order, then friction rows in the order of their
// non-penetration rows (rows without one last, mesh before pair), ten iterations or until one changes nothing.
// g holds only the touching cars' mesh rows.  dl / da: the velocity changes per car slot.
RG_HD void SolveGroupVelocity(const ManifoldEntry* entries, int entry_count, CarWorldRows& g, CarPairRows& p,
                              float inv_mass, float mesh_friction, Vec3* dl, Vec3* da) {
    GroupOrder on;
    BuildGroupOrder(entries, entry_count, g.nn, g, [](int r) { return r; }, p.nn, p, [](int r) { return r; }, on);
    uint8_t fk[kMaxCarWorldRows + kMaxPairRows], fr[kMaxCarWorldRows + kMaxPairRows];
    int fcount = 0;
    uint32_t wfu = 0u, pfu = 0u;
    for (int k = 0; k < on.count; ++k) {
        const int nrow = on.row[k];
        if (on.kind[k] == 0) {
            for (int r = 0; r < g.nf; ++r) {
                if (!((wfu >> r) & 1u) && g.f[r].normal_row == nrow) {
                    wfu |= 1u << r;
                    fk[fcount] = 0;
                    fr[fcount++] = static_cast<uint8_t>(r);
                    break;
                }
            }
        } else {
            for (int r = 0; r < p.nf; ++r) {
                if (!((pfu >> r) & 1u) && p.f[r].normal_row == nrow) {
                    pfu |= 1u << r;
                    fk[fcount] = 1;
                    fr[fcount++] = static_cast<uint8_t>(r);
                }
            }
        }
    }
    for (int r = 0; r < g.nf; ++r) {
        if (!((wfu >> r) & 1u)) fk[fcount] = 0, fr[fcount++] = static_cast<uint8_t>(r);
    }
    for (int r = 0; r < p.nf; ++r) {
        if (!((pfu >> r) & 1u)) fk[fcount] = 1, fr[fcount++] = static_cast<uint8_t>(r);
    }
    for (int iter = 0; iter < 10; ++iter) {
        bool changed = false;
        for (int k = 0; k < on.count; ++k) {
            const int r = on.row[k];
            if (on.kind[k] == 0) {
                const int c = g.n_car[r];
                changed = SolveWorldVelocityRow(g.n[r], inv_mass, dl[c], da[c]) || changed;
            } else {
                changed = SolvePairVelocityRow(p.n[r], inv_mass, dl, da, 0.0f, false) || changed;
            }
        }
        for (int k = 0; k < fcount; ++k) {
            const int r = fr[k];
            if (fk[k] == 0) {
                CarWorldFrictionRow& row = g.f[r];
                const int c = g.n_car[row.normal_row];
                changed = SolveWorldFrictionRow(row, mesh_friction * g.n[row.normal_row].applied,
                                                g.n[row.normal_row].applied, inv_mass, dl[c], da[c]) || changed;
            } else {
                CarPairRow& row = p.f[r];
                const float total = p.n[row.normal_row].applied;
                if (total <= 0.0f) continue;
                changed = SolvePairVelocityRow(row, inv_mass, dl, da, kCarCarFriction * total, true) || changed;
            }
        }
        if (!changed) break;
    }
}

// Split-impulse solve of the joint rows in visit order; push / turn: pseudo-velocities per car slot.
RG_HD void SolveGroupSplit(const ManifoldEntry* entries, int entry_count, CarWorldRows& g, CarPairRows& p,
                           float inv_mass, Vec3* push, Vec3* turn) {
    GroupOrder o;
    BuildGroupOrder(entries, entry_count, g.ns, g, [&](int r) { return g.s[r].normal_row; }, p.ns, p,
                    [&](int r) { return p.s[r].normal_row; }, o);
    for (int iter = 0; iter < 10; ++iter) {
        bool changed = false;
        for (int k = 0; k < o.count; ++k) {
            const int r = o.row[k];
            if (o.kind[k] == 0) {
                const int c = g.n_car[g.s[r].normal_row];
                changed = SolveCarWorldSplitRow(g, g.s[r], inv_mass, push[c], turn[c]) || changed;
                continue;
            }
            CarPairSplitRow& row = p.s[r];
            if (row.rhs == 0.0f) continue;
            const CarPairRow& nr = p.n[row.normal_row];
            const int a = nr.a, b = nr.b;
            const float before = row.applied;
            float delta = row.rhs;
            delta -= (Dot(-nr.dir, push[b]) + Dot(nr.rxb, turn[b])) * nr.jac;
            delta -= (Dot(nr.dir, push[a]) + Dot(nr.rxa, turn[a])) * nr.jac;
            const float sum = row.applied + delta;
            if (sum < 0.0f) {
                delta = 0.0f - row.applied;
                row.applied = 0.0f;
            } else {
                row.applied = sum;
            }
            if (delta != 0.0f) {
                push[a] += (nr.dir * inv_mass) * delta;
                turn[a] += nr.anga * delta;
                push[b] += (-nr.dir * inv_mass) * delta;
                turn[b] += nr.angb * delta;
            }
            changed = changed || delta != 0.0f || FloatBits(before) != FloatBits(row.applied);
        }
        if (!changed) break;
    }
}

}  // namespace core
}  // namespace rg
