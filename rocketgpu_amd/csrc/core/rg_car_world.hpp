// Car hitbox against the arena mesh: contact points, the car's persistent manifolds, solver rows and their solve.
//
// Reference (the CPU build's path for a custom arena mesh, include/rocketgpu/detail/step/):
//   SolveCarWorldStaticContacts arena_ball.hpp:395-1049 -> AppendCarWorldMeshContactRowsBody
//   car_world_contacts.hpp:1270-2145 (the walk, consider_triangle, merge_retained_mesh_manifold, emit_candidates);
//   ComputeCarWorldTriangleHit :652-695; CarWorldTriangleQueryTest :1014-1065; the query frame :879-926;
//   row set-up arena_geometry.hpp:2404-2604 and :2767-2812; the velocity solve :3103-3316 and :5527-5866; the split
//   solve :3318-3500 and :5035-5276; RecordArenaWorldContact env_step.hpp:2908; ApplyGameCarWorldContactNormal
//   arena_ball.hpp:94-169.
//
// What it computes, in Bullet's terms: the car's compound shape (a box with a 0.04 bt margin) against the mesh's
// triangles whose boxes meet the car's box; each triangle within the contact-breaking threshold gives one point
// (GJK / EPA, rg_gjk.hpp), corrected at internal edges; the points go into the car's persistent manifold for that
// mesh (nearest-point replacement, deepest point kept, largest area when full); every manifold point is one
// non-penetration row, one friction row and, when penetrating, one split-impulse row.  The rows of one car touch only
// that car, so they are solved per car with the solver's iteration count and early exit.
#pragma once

#include "rg_car_tick.hpp"
#include "rg_gjk.hpp"

namespace rg {
namespace core {

constexpr int kCarWorldRowsPerMesh = 4;
// Rows held at once: one car's (two manifolds of four points), or those of the cars that touch each other.
constexpr int kMaxCarWorldRows = 8;

// Contact-breaking threshold of a car's hitbox compound: its bounding sphere (box center offset plus half diagonal)
// times Bullet's 0.02 factor (CarContactBreakingThresholdBt, arena_geometry.hpp:6787-6798).
RG_HD float CarBreakingThresholdBt(const CarConfig& cfg) {
    const Vec3 half = (BulletHitboxSizeUu(cfg) * 0.02f) * 0.5f;
    const float margin = Min(0.1f * Min(Min(half.x, half.y), half.z), 0.04f);
    const Vec3 mv(margin, margin, margin);
    const Vec3 hm = (half - mv) + mv;
    const Vec3 center = cfg.hitbox_offset * 0.02f;
    const Vec3 lo = center - hm, hi = center + hm;
    const Vec3 sphere_center = (lo + hi) * 0.5f;
    const float radius = Length(hi - lo) * 0.5f;
    return (radius + Length(sphere_center)) * 0.02f;
}

// The car body's Bullet transform (basis rows from the orientation columns, origin pos_bt) and its inverse on points.
RG_HD gjk::Xf CarBodyXf(const CarState& car) {
    Vec3 f, r, u;
    CarBasis(car, f, r, u);
    gjk::Xf t;
    t.r0 = Vec3(f.x, r.x, u.x);
    t.r1 = Vec3(f.y, r.y, u.y);
    t.r2 = Vec3(f.z, r.z, u.z);
    t.o = car.pos_bt;
    return t;
}
RG_HD Vec3 BodyToWorld(const gjk::Xf& body, const Vec3& local) { return gjk::Apply(body, local); }
RG_HD Vec3 WorldToBody(const gjk::Xf& body, const Vec3& p) {
    const Vec3 v = p - body.o;
    return Vec3(Dot(Vec3(body.r0.x, body.r1.x, body.r2.x), v), Dot(Vec3(body.r0.y, body.r1.y, body.r2.y), v),
                Dot(Vec3(body.r0.z, body.r1.z, body.r2.z), v));
}

// The hitbox's world box and frame for the triangle queries.
struct HitboxFrame {
    gjk::Xf body;
    Vec3 ax, ay, az;   // hitbox axes
    Vec3 center;       // hitbox center (bt)
    Vec3 half;         // half extents with the margin (bt)
    Vec3 extent;       // world half box
    Vec3 lo, hi;
};

RG_HD void MakeHitboxFrame(const CarState& car, const CarConfig& cfg, HitboxFrame& f) {
    f.body = CarBodyXf(car);
    CarBasis(car, f.ax, f.ay, f.az);
    f.center = gjk::Apply(f.body, cfg.hitbox_offset * (1.0f / 50.0f));
    f.half = HitboxHalfBt(cfg);
    const Vec3& h = f.half;
    f.extent = Vec3(Abs(f.ax.x) * h.x + Abs(f.ay.x) * h.y + Abs(f.az.x) * h.z,
                    Abs(f.ax.y) * h.x + Abs(f.ay.y) * h.y + Abs(f.az.y) * h.z,
                    Abs(f.ax.z) * h.x + Abs(f.ay.z) * h.y + Abs(f.az.z) * h.z);
    f.lo = f.center - f.extent;
    f.hi = f.center + f.extent;
}

// Triangle screen before the exact query: 0 = its box misses the hitbox's box, 1 = separated by more than the
// threshold along the triangle normal, a hitbox axis or an edge-axis cross product (no point can come from it),
// 2 = run the exact query.
RG_HD int TriangleScreen(const MeshTriangle& tri, const HitboxFrame& f, float threshold) {
    if (Min(Min(tri.a.z, tri.b.z), tri.c.z) > f.hi.z || Max(Max(tri.a.z, tri.b.z), tri.c.z) < f.lo.z ||
        Min(Min(tri.a.x, tri.b.x), tri.c.x) > f.hi.x || Max(Max(tri.a.x, tri.b.x), tri.c.x) < f.lo.x ||
        Min(Min(tri.a.y, tri.b.y), tri.c.y) > f.hi.y || Max(Max(tri.a.y, tri.b.y), tri.c.y) < f.lo.y) {
        return 0;
    }
    const Vec3 n = Normalize(tri.normal, Vec3(0.0f, 0.0f, 1.0f));
    const float center_proj = Dot(n, f.center);
    const float support = Abs(Dot(n, f.ax)) * f.half.x + Abs(Dot(n, f.ay)) * f.half.y + Abs(Dot(n, f.az)) * f.half.z;
    const float tri_proj = Dot(n, tri.a);
    if (tri_proj - (center_proj + support) > threshold || (center_proj - support) - tri_proj > threshold) return 1;
    const float reject = threshold + 1.0e-3f;
    const auto gap = [&](const Vec3& axis, float h) {
        const float c = Dot(axis, f.center);
        const float pa = Dot(axis, tri.a), pb = Dot(axis, tri.b), pc = Dot(axis, tri.c);
        return Max(Min(pa, Min(pb, pc)) - (c + h), (c - h) - Max(pa, Max(pb, pc)));
    };
    if (gap(f.ax, f.half.x) > reject || gap(f.ay, f.half.y) > reject || gap(f.az, f.half.z) > reject) return 1;
    const Vec3 e[3] = {tri.b - tri.a, tri.c - tri.b, tri.a - tri.c};
    const Vec3 axes[3] = {f.ax, f.ay, f.az};
    for (int i = 0; i < 3; ++i) {
        for (int j = 0; j < 3; ++j) {
            const Vec3 u = Cross(e[i], axes[j]);
            const float len = sqrtf(Dot(u, u));
            if (!(len > 1e-6f)) continue;
            const float c = Dot(u, f.center);
            const float r = Abs(Dot(u, f.ax)) * f.half.x + Abs(Dot(u, f.ay)) * f.half.y + Abs(Dot(u, f.az)) * f.half.z;
            const float pa = Dot(u, tri.a), pb = Dot(u, tri.b), pc = Dot(u, tri.c);
            if (Max(Min(pa, Min(pb, pc)) - (c + r), (c - r) - Max(pa, Max(pb, pc))) > reject * len) return 1;
        }
    }
    return 2;
}

// One triangle's contact: the GJK/EPA closest points, kept within the threshold; the normal replaced by the
// triangle's face normal and corrected at internal edges.  raw_*: before the internal-edge correction.
struct TriangleHit {
    Vec3 normal, point_b;
    Vec3 raw_normal, raw_point_b;
    float distance;
};

RG_HD bool CarTriangleHit(const CarConfig& cfg, const HitboxFrame& f, const MeshTriangle& tri, float threshold,
                          TriangleHit& out) {
    gjk::Xf box = f.body;
    box.o = f.center;
    const Vec3 raw_half = (BulletHitboxSizeUu(cfg) * 0.5f) * (1.0f / 50.0f);
    gjk::BoxTriHit h;
    if (!gjk::BoxTriangleContact(raw_half, box, tri.a, tri.b, tri.c, h) || h.distance > threshold) return false;
    const Vec3 e1 = tri.b - tri.a, e2 = tri.c - tri.a;
    const Vec3 fn(e1.y * e2.z - e1.z * e2.y, e1.z * e2.x - e1.x * e2.z, e1.x * e2.y - e1.y * e2.x);
    const float inv_len = 1.0f / sqrtf(fn.x * fn.x + fn.y * fn.y + fn.z * fn.z);
    Vec3 normal(fn.x * inv_len, fn.y * inv_len, fn.z * inv_len);
    out.raw_normal = normal;
    out.raw_point_b = h.point_b;
    Vec3 point_b = h.point_b;
    const Vec3 point_a = point_b + normal * h.distance;
    AdjustInternalEdgeContact(tri, point_a, h.distance, point_b, normal);
    out.normal = normal;
    out.point_b = point_b;
    out.distance = h.distance;
    return true;
}

// A point of the manifold under construction this tick.  src >= 0: found this tick; -2 - k: the refreshed
// retained point k.
struct WorldPoint {
    int32_t src;
    Vec3 normal;
    Vec3 local_a;
    Vec3 point_b;
    float distance;      // refreshed signed distance (bt)
    float add_distance;  // distance when it entered the manifold (the replacement rule ranks by it)
    Vec3 rel_bt;         // contact point on the car minus the car's origin (bt)
};

// Bullet's replacement slot for a fifth point: keep the deepest point, drop the one whose removal leaves the largest
// quadrilateral (sortCachedPoints).
RG_HD int ReplaceSlot(const WorldPoint* pts, const WorldPoint& np) {
    int deepest = -1;
    float max_pen = np.add_distance;
    for (int i = 0; i < 4; ++i) {
        if (pts[i].add_distance < max_pen) {
            deepest = i;
            max_pen = pts[i].add_distance;
        }
    }
    const auto area = [&](int a, int b, int c) {
        return LengthSq(Cross(np.local_a - pts[a].local_a, pts[b].local_a - pts[c].local_a));
    };
    const float r0 = deepest != 0 ? area(1, 3, 2) : 0.0f;
    const float r1 = deepest != 1 ? area(0, 3, 2) : 0.0f;
    const float r2 = deepest != 2 ? area(0, 3, 1) : 0.0f;
    const float r3 = deepest != 3 ? area(0, 2, 1) : 0.0f;
    int best = 0;
    float big = r0;
    if (r1 > big) {
        best = 1;
        big = r1;
    }
    if (r2 > big) {
        best = 2;
        big = r2;
    }
    if (r3 > big) best = 3;
    return best;
}

// The same rule on a stored manifold (distances at the last refresh), used when points merge into a live manifold.
RG_HD int ReplaceSlotStored(const CarWorldManifold& m, const CarWorldPoint& np) {
    int deepest = 4;
    float max_pen = np.distance;
    for (int i = 0; i < 4; ++i) {
        if (m.pts[i].distance < max_pen) {
            deepest = i;
            max_pen = m.pts[i].distance;
        }
    }
    const auto area = [&](int a, int b, int c) {
        return LengthSq(Cross(np.local_a - m.pts[a].local_a, m.pts[b].local_a - m.pts[c].local_a));
    };
    const float r0 = deepest != 0 ? area(1, 3, 2) : 0.0f;
    const float r1 = deepest != 1 ? area(0, 3, 2) : 0.0f;
    const float r2 = deepest != 2 ? area(0, 3, 1) : 0.0f;
    const float r3 = deepest != 3 ? area(0, 2, 1) : 0.0f;
    int best = 0;
    float big = r0;
    if (r1 > big) {
        best = 1;
        big = r1;
    }
    if (r2 > big) {
        best = 2;
        big = r2;
    }
    if (r3 > big) best = 3;
    return best;
}

RG_HD float StoredDistance(const gjk::Xf& body, const CarWorldPoint& p, Vec3& wa) {
    wa = BodyToWorld(body, p.local_a);
    return Dot(wa - p.point_b, p.normal);
}

// Refresh a stored manifold on the car's current pose: distances recomputed, points past the threshold (along the
// normal or sideways) dropped by swapping in the last point, from the last index down.
RG_HD void RefreshManifold(CarWorldManifold& m, const gjk::Xf& body, float threshold) {
    if (m.count == 0) return;
    Vec3 wa[4];
    for (int i = 0; i < m.count; ++i) m.pts[i].distance = StoredDistance(body, m.pts[i], wa[i]);
    for (int i = static_cast<int>(m.count) - 1; i >= 0; --i) {
        const CarWorldPoint& p = m.pts[i];
        const Vec3 proj = p.point_b - (wa[i] - p.normal * p.distance);
        if (p.distance > threshold || Dot(proj, proj) > threshold * threshold) {
            const int last = static_cast<int>(m.count) - 1;
            if (i != last) {
                m.pts[i] = m.pts[last];
                wa[i] = wa[last];
            }
            m.count = static_cast<uint8_t>(last);
        }
    }
}

RG_HD void MergeStoredPoint(CarWorldManifold& m, const CarWorldPoint& p, float threshold) {
    if (m.count == 0) {
        m.pts[0] = p;
        m.count = 1;
        return;
    }
    float shortest = threshold * threshold;
    int nearest = -1;
    for (int i = 0; i < m.count; ++i) {
        const Vec3 d = m.pts[i].local_a - p.local_a;
        const float dsq = Dot(d, d);
        if (dsq < shortest) {
            shortest = dsq;
            nearest = i;
        }
    }
    if (nearest >= 0) {
        m.pts[nearest] = p;
    } else if (m.count < 4) {
        m.pts[m.count++] = p;
    } else {
        m.pts[ReplaceSlotStored(m, p)] = p;
    }
}

// The working manifold rebuilt from a stored one: every point's world position on the current pose.
RG_HD int PointsFromManifold(const CarWorldManifold& m, const gjk::Xf& body, WorldPoint* pts) {
    for (int i = 0; i < m.count; ++i) {
        const CarWorldPoint& p = m.pts[i];
        Vec3 wa;
        StoredDistance(body, p, wa);
        WorldPoint& w = pts[i];
        w.normal = p.normal;
        w.distance = p.distance;
        w.local_a = p.local_a;
        w.point_b = p.point_b;
        w.rel_bt = wa - body.o;
    }
    return m.count;
}

RG_HD int FindManifold(CarWorldManifold* ms, int mesh) {
    for (int k = 0; k < kCarWorldManifoldsPerCar; ++k) {
        if (ms[k].count > 0 && ms[k].mesh == mesh) return k;
    }
    return -1;
}

// Rows of one car against the world, in bt units (the car's solver frame).
struct CarWorldRow {
    Vec3 normal;
    Vec3 rxn;        // rel x normal
    Vec3 ang;        // world inverse inertia * (rel x normal)
    float jac;       // 1 / (inv_mass + n . (ang x rel))
    float rhs;
    float applied;
};

struct CarWorldFrictionRow {
    Vec3 tangent;
    Vec3 rxt;
    Vec3 ang;
    float jac;
    float rhs;
    float applied;
    int normal_row;
};

// A split row shares its point's non-penetration row's direction terms and denominator.
struct CarWorldSplitRow {
    float rhs;
    float applied;
    int normal_row;
};

// Rows of cars against the meshes.  A car without car-car rows has its rows solved and dropped before the next car's
// are made; the rows of cars with car-car rows stay for the joint solve.
struct CarWorldRows {
    CarWorldRow n[kMaxCarWorldRows];
    CarWorldFrictionRow f[kMaxCarWorldRows];
    CarWorldSplitRow s[kMaxCarWorldRows];
    int8_t n_car[kMaxCarWorldRows], n_mesh[kMaxCarWorldRows];  // the car slot and mesh of each non-penetration row
    int nn, nf, ns;
    uint8_t full;        // a point found no room for its rows
    uint8_t hit;         // a manifold of the last car held a point this tick
    Vec3 latest_normal;  // the normal recorded as the last car's world contact
};

// Rows for a car's manifold points (emit_candidates): non-penetration with the mesh restitution (no velocity
// threshold), friction along the sliding direction (or Bullet's plane-space axis), and split rows for penetrating
// points.  Every point also records the car's world-surface contact.
RG_HD void EmitRows(CarState& car, const CarConfig& cfg, const CarParams& m, const WorldPoint* pts, int count,
                    float dt, CarWorldRows& out) {
    if (count <= 0) return;
    Vec3 i0, i1, i2;
    if (car.restored_inv_inertia_valid) {
        i0 = car.restored_inv_inertia_bt[0];
        i1 = car.restored_inv_inertia_bt[1];
        i2 = car.restored_inv_inertia_bt[2];
    } else {
        InvInertiaWorldRows(car.tick_start_basis_forward, car.tick_start_basis_right, car.tick_start_basis_up,
                            InvInertiaLocalBt(cfg, m.car_mass), i0, i1, i2);
    }
    const float inv_mass = 1.0f / m.car_mass;
    const Vec3 vel_ext = car.vel_bt + car.ext_force_impulse_bt;
    const Vec3 ang_ext = car.ang_vel + car.ext_torque_impulse;
    int emitted = 0;
    for (int c = 0; c < count && emitted < kCarWorldRowsPerMesh; ++c) {
        const WorldPoint& p = pts[c];
        const float pen_uu = Max(0.0f, -p.distance) * 50.0f;
        // The world-surface record of the contact (CarState::RecordWorldSurface with a body contact).
        car.last_world_surface_contact = car.tick_count_since_update;
        if (car.last_ball_contact_start == car.tick_count_since_update) car.last_ball_started_with_world = 1;
        if (car.last_ball_contact_start != ~0ULL && car.first_world_after_last_ball == ~0ULL &&
            car.tick_count_since_update > car.last_ball_contact_start) {
            car.first_world_after_last_ball = car.tick_count_since_update;
        }
        out.hit = 1;
        out.latest_normal = p.normal;
        ++emitted;
        if (out.nn >= kMaxCarWorldRows) {
            out.full = 1;
            continue;
        }
        const Vec3 ang = MulRows(i0, i1, i2, Cross(p.rel_bt, p.normal));
        const float jac = 1.0f / (inv_mass + Dot(p.normal, Cross(ang, p.rel_bt)));
        if (pen_uu > 0.0f && dt > 0.0f) {
            CarWorldSplitRow& s = out.s[out.ns++];
            s.normal_row = out.nn;
            s.rhs = ((-p.distance * 0.8f) / dt) * jac;
            s.applied = 0.0f;
        }
        // Sliding direction at the contact with the external impulses applied.
        const Vec3 vel_nd = vel_ext + Cross(ang_ext, p.rel_bt);
        const float vn = Dot(p.normal, vel_nd);
        const Vec3 lateral = vel_nd - p.normal * vn;
        const float lat2 = LengthSq(lateral);
        Vec3 tangent;
        if (lat2 > 1.1920928955078125e-7f) {
            tangent = lateral * (1.0f / sqrtf(lat2));
        } else if (Abs(p.normal.z) > 0.7071067811865475244008443621048490f) {
            const float a = p.normal.y * p.normal.y + p.normal.z * p.normal.z;
            const float k = 1.0f / sqrtf(a);
            tangent = Vec3(0.0f, -p.normal.z * k, p.normal.y * k);
        } else {
            const float a = p.normal.x * p.normal.x + p.normal.y * p.normal.y;
            const float k = 1.0f / sqrtf(a);
            tangent = Vec3(-p.normal.y * k, p.normal.x * k, 0.0f);
        }
        const Vec3 tang_ang = MulRows(i0, i1, i2, Cross(p.rel_bt, tangent));
        {
            CarWorldRow& r = out.n[out.nn];
            r.normal = p.normal;
            r.rxn = Cross(p.rel_bt, p.normal);
            r.ang = ang;
            r.jac = jac;
            const float rest_vn = Dot(p.normal, car.vel_bt + Cross(car.ang_vel, p.rel_bt));
            const float rhs_vn = Dot(p.normal, vel_ext) + Dot(r.rxn, ang_ext);
            const float restitution = Max(m.arena_restitution * -rest_vn, 0.0f);
            r.rhs = (restitution - rhs_vn) * r.jac;
            r.applied = 0.0f;
            const int normal_idx = out.nn++;
            if (m.arena_friction > 0.0f) {
                CarWorldFrictionRow& fr = out.f[out.nf++];
                fr.tangent = tangent;
                fr.rxt = Cross(p.rel_bt, tangent);
                fr.ang = tang_ang;
                fr.jac = 1.0f / (inv_mass + Dot(tangent, Cross(tang_ang, p.rel_bt)));
                const float v1 = Dot(tangent, vel_ext) + Dot(fr.rxt, car.ang_vel);
                fr.rhs = (0.0f - (v1 + 0.0f)) * fr.jac;
                fr.applied = 0.0f;
                fr.normal_row = normal_idx;
            }
        }
    }
}

// The car's contacts with the meshes whose broadphase pair with it exists (mesh_bits) or that still hold retained
// points.  Returns rows for the solver; updates the retained manifolds.
RG_HD void CarWorldContacts(CarState& car, const CarConfig& cfg, const CarParams& m, const ArenaMeshView& mesh,
                            uint32_t mesh_bits, CarWorldManifold* retained, float dt, CarWorldRows& out) {
    out.hit = 0;
    out.latest_normal = Vec3(0.0f, 0.0f, 1.0f);
    const int car_n0 = out.nn;
    uint32_t visit = mesh_bits;
    for (int k = 0; k < kCarWorldManifoldsPerCar; ++k) {
        if (retained[k].count > 0 && retained[k].mesh < 32) visit |= 1u << retained[k].mesh;
    }
    if (visit == 0u) return;
    const float threshold = CarBreakingThresholdBt(cfg);
    HitboxFrame f;
    MakeHitboxFrame(car, cfg, f);
    Vec3 last_normal(0.0f, 0.0f, 1.0f);
    for (int mi = 0; mi < mesh.mesh_count && mi < 32; ++mi) {
        if (((visit >> mi) & 1u) == 0u) continue;
        const MeshInfo& info = mesh.meshes[mi];
        if (info.node_count <= 0) continue;
        if (((mesh_bits >> mi) & 1u) == 0u) {
            const int k = FindManifold(retained, mi);
            if (k >= 0) retained[k].count = 0;
            continue;
        }
        WorldPoint pts[4];
        int n = 0;
        CarWorldPoint seeds[4];
        bool seeded = false;
        {
            const int k = FindManifold(retained, mi);
            if (k >= 0) {
                RefreshManifold(retained[k], f.body, threshold);
                n = PointsFromManifold(retained[k], f.body, pts);
                for (int q = 0; q < n; ++q) {
                    seeds[q] = retained[k].pts[q];
                    pts[q].src = -2 - q;
                    pts[q].add_distance = pts[q].distance;
                }
                seeded = true;
            }
        }
        constexpr float kMeshMargin = 1.0e-3f;
        if (f.lo.x <= info.aabb_max.x + kMeshMargin && f.hi.x >= info.aabb_min.x - kMeshMargin &&
            f.lo.y <= info.aabb_max.y + kMeshMargin && f.hi.y >= info.aabb_min.y - kMeshMargin &&
            f.lo.z <= info.aabb_max.z + kMeshMargin && f.hi.z >= info.aabb_min.z - kMeshMargin) {
            uint16_t qmin[3], qmax[3];
            QuantizePoint(info.q, f.center - f.extent, 0, qmin);
            QuantizePoint(info.q, f.center + f.extent, 1, qmax);
            WalkMeshBvh(
                mesh.nodes, info.node_start, info.node_start + info.node_count, qmin, qmax,
                [&](int32_t local) {
                    if (local < 0 || local >= info.tri_count) return;
                    const int32_t ti = info.tri_start + local;
                    const MeshTriangle& tri = mesh.tris[ti];
                    if (TriangleScreen(tri, f, threshold) != 2) return;
                    TriangleHit h;
                    if (!CarTriangleHit(cfg, f, tri, threshold, h)) return;
                    WorldPoint np;
                    np.src = ti;
                    np.normal = h.normal;
                    np.add_distance = h.distance;
                    np.point_b = h.point_b;
                    np.local_a = WorldToBody(f.body, h.raw_point_b + h.raw_normal * h.distance);
                    const Vec3 wa = BodyToWorld(f.body, np.local_a);
                    np.distance = Dot(wa - np.point_b, np.normal);
                    np.rel_bt = wa - f.body.o;
                    int merge = -1;
                    float shortest = threshold * threshold;
                    for (int q = 0; q < n; ++q) {
                        const Vec3 d = pts[q].local_a - np.local_a;
                        const float dsq = Dot(d, d);
                        if (dsq < shortest) {
                            shortest = dsq;
                            merge = q;
                        }
                    }
                    last_normal = h.raw_normal;
                    if (np.distance > threshold) {
                        if (merge >= 0) {
                            for (int q = merge; q + 1 < n; ++q) pts[q] = pts[q + 1];
                            --n;
                        }
                        return;
                    }
                    if (merge >= 0) {
                        pts[merge] = np;
                    } else if (n < 4) {
                        pts[n++] = np;
                    } else {
                        pts[ReplaceSlot(pts, np)] = np;
                    }
                },
                AcceptAll{});
        }
        // Store the manifold for the next tick and rebuild the working points from it.
        int k = FindManifold(retained, mi);
        if (k < 0 && n > 0) {
            for (int q = 0; q < kCarWorldManifoldsPerCar && k < 0; ++q) {
                if (retained[q].count == 0) k = q;
            }
            if (k >= 0) retained[k].mesh = static_cast<uint8_t>(mi);
        }
        if (k >= 0) {
            CarWorldManifold& rm = retained[k];
            RefreshManifold(rm, f.body, threshold);
            const bool push_all = rm.count == 0 || seeded;
            if (seeded) rm.count = 0;
            for (int q = 0; q < n; ++q) {
                if (seeded && pts[q].src <= -2) {
                    rm.pts[rm.count++] = seeds[-2 - pts[q].src];
                    continue;
                }
                CarWorldPoint sp;
                sp.local_a = pts[q].local_a;
                sp.point_b = pts[q].point_b;
                sp.normal = pts[q].normal;
                Vec3 wa;
                sp.distance = StoredDistance(f.body, sp, wa);
                if (push_all) {
                    rm.pts[rm.count++] = sp;
                } else {
                    MergeStoredPoint(rm, sp, threshold);
                }
            }
            n = PointsFromManifold(rm, f.body, pts);
        }
        const int nn0 = out.nn;
        EmitRows(car, cfg, m, pts, n, dt, out);
        for (int r = nn0; r < out.nn; ++r) out.n_mesh[r] = static_cast<int8_t>(mi);
    }
    if (out.hit) {
        out.latest_normal = last_normal;
        // The recorded normal: the normalized sum of the rows' normals (one mesh: a single ordering key).
        Vec3 sum(0.0f, 0.0f, 0.0f);
        for (int r = car_n0; r < out.nn; ++r) sum += out.n[r].normal;
        if (out.nn > car_n0) {
            const float len_sq = Dot(sum, sum);
            if (len_sq >= 1e-8f) {
                const Vec3 nn = sum * (1.0f / sqrtf(len_sq));
                out.latest_normal = Vec3(0.0f + nn.x, 0.0f + nn.y, 0.0f + nn.z);
            }
        }
    }
}

// Velocity solve of one car's rows (non-penetration rows from n0, friction rows from f0): non-penetration rows, then
// friction rows (each bounded by `friction` times its non-penetration row's impulse), in order, until an iteration
// changes nothing.  Returns the accumulated linear and angular velocity changes.
RG_HD void SolveCarWorldVelocity(CarWorldRows& r, int n0, int f0, float inv_mass, float friction, Vec3& d_lin,
                                 Vec3& d_ang) {
    d_lin = Vec3(0.0f, 0.0f, 0.0f);
    d_ang = Vec3(0.0f, 0.0f, 0.0f);
    for (int iter = 0; iter < 10; ++iter) {
        bool changed = false;
        for (int i = n0; i < r.nn; ++i) {
            CarWorldRow& row = r.n[i];
            const float dv = Dot(row.normal, d_lin) + Dot(row.rxn, d_ang);
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
            if (delta == 0.0f) {
                changed = changed || FloatBits(before) != FloatBits(row.applied);
                continue;
            }
            d_lin += (row.normal * inv_mass) * delta;
            d_ang += row.ang * delta;
            changed = true;
        }
        for (int i = f0; i < r.nf; ++i) {
            CarWorldFrictionRow& row = r.f[i];
            const float total = r.n[row.normal_row].applied;
            if (total <= 0.0f) continue;
            const float max_impulse = friction * total;
            float delta = row.rhs;
            const float dv = Dot(row.tangent, d_lin) + Dot(row.rxt, d_ang);
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
            if (delta == 0.0f) {
                changed = changed || FloatBits(before) != FloatBits(row.applied);
                continue;
            }
            d_lin += (row.tangent * inv_mass) * delta;
            d_ang += row.ang * delta;
            changed = true;
        }
        if (!changed) break;
    }
}

// One split row of a car against a mesh: returns whether it changed anything.
RG_HD bool SolveCarWorldSplitRow(const CarWorldRows& r, CarWorldSplitRow& row, float inv_mass, Vec3& push, Vec3& turn) {
    if (row.rhs == 0.0f) return false;
    const CarWorldRow& nr = r.n[row.normal_row];
    const float before = row.applied;
    float delta = row.rhs;
    const float dv_lin = Dot(nr.normal, push);
    const float dv_ang = Dot(nr.rxn, turn);
    delta = delta - (dv_lin + dv_ang) * nr.jac;
    const float sum = row.applied + delta;
    if (sum < 0.0f) {
        delta = 0.0f - row.applied;
        row.applied = 0.0f;
    } else {
        row.applied = sum;
    }
    if (delta != 0.0f) {
        push += (nr.normal * inv_mass) * delta;
        turn += nr.ang * delta;
    }
    return delta != 0.0f || FloatBits(before) != FloatBits(row.applied);
}

// Split-impulse solve of one car's penetrating rows (from s0): pseudo-velocities that push the car out of the mesh
// without adding momentum.
RG_HD void SolveCarWorldSplit(CarWorldRows& r, int s0, float inv_mass, Vec3& push, Vec3& turn) {
    push = Vec3(0.0f, 0.0f, 0.0f);
    turn = Vec3(0.0f, 0.0f, 0.0f);
    for (int iter = 0; iter < 10; ++iter) {
        bool changed = false;
        for (int i = s0; i < r.ns; ++i) changed = SolveCarWorldSplitRow(r, r.s[i], inv_mass, push, turn) || changed;
        if (!changed) break;
    }
}

}  // namespace core
}  // namespace rg
