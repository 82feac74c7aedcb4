// Arena mesh: device-side tables and the queries the step runs against them.
//
// The arena is a set of static triangle meshes in Bullet units (1 bt = 50 uu).  Each mesh carries a Bullet-style
// quantized AABB tree (btQuantizedBvh: 16-bit node boxes, escape indices, depth-first node order) built on the host
// by mesh_build.cpp.  Every query here visits leaves in depth-first node order, which is the order the reference's
// stackless walk (include/rocketgpu/detail/arena/mesh_bvh.hpp:594-681) and Bullet report them in; the contact and
// ray code depends on that order for its tie-breaks.
#pragma once

#include "rg_math.hpp"

namespace rg {
namespace core {

constexpr int kMaxArenaMeshes = 16;

// One triangle (128 B, two 64-byte lines).  Vertices and normal in bt; edge angles and flags are Bullet's
// btTriangleInfo (internal-edge data) computed at build time.
struct alignas(16) MeshTriangle {
    Vec3 a;
    float edge_angle_v0v1;
    Vec3 b;
    float edge_angle_v1v2;
    Vec3 c;
    float edge_angle_v2v0;
    Vec3 normal;   // Normalize(cross(b - a, c - a)), fallback +z
    uint32_t flags;  // bits 0-2: edge convex (v0v1, v1v2, v2v0); bits 3-5: swap normal B; bits 8-15: mesh index
    Vec3 min;
    float pad0;
    Vec3 max;
    float pad1;
    uint16_t qmin[3];
    uint16_t qmax[3];
    uint32_t pad2;
    uint32_t pad3[2];
};
static_assert(sizeof(MeshTriangle) == 128, "MeshTriangle is two cache lines");

constexpr uint32_t kTriEdgeConvex[3] = {1u << 0, 1u << 1, 1u << 2};
constexpr uint32_t kTriEdgeSwapNormalB[3] = {1u << 3, 1u << 4, 1u << 5};

// Quantized BVH node (16 B).  escape_or_tri >= 0: leaf with the mesh-local triangle index; < 0: internal node whose
// subtree spans -escape_or_tri nodes.
struct alignas(16) MeshBvhNode {
    uint16_t qmin[3];
    uint16_t qmax[3];
    int32_t escape_or_tri;
};
static_assert(sizeof(MeshBvhNode) == 16, "MeshBvhNode is one 16-byte load");

struct MeshQuantization {
    Vec3 bvh_min;
    Vec3 bvh_max;
    Vec3 scale;  // 65533 / extent
};

struct MeshInfo {
    Vec3 aabb_min;
    Vec3 aabb_max;
    MeshQuantization q;
    int32_t mesh_index;
    int32_t tri_start;
    int32_t tri_count;
    int32_t node_start;
    int32_t node_count;
    int32_t ray_order;  // meshes[k].ray_order: the mesh a ray visits k-th (by quantized max x, stable)
    int32_t pad[2];
};

// Everything one arena geometry needs on the device.  Pointers are device addresses (host addresses in the
// host-compiled core).
struct ArenaMeshView {
    const MeshTriangle* tris;
    const MeshBvhNode* nodes;
    const MeshInfo* meshes;
    int32_t mesh_count;
    int32_t tri_count;
};

// ------------------------------------------------------------------------------------------------------------
// Quantization (btQuantizedBvh::quantizeWithClamp; mesh_bvh.hpp:385-404).

RG_HD void QuantizePoint(const MeshQuantization& q, const Vec3& point, int is_max, uint16_t out[3]) {
    Vec3 p = Max3(point, q.bvh_min);
    p = Min3(p, q.bvh_max);
    const Vec3 v = Mul3(p - q.bvh_min, q.scale);
    if (is_max) {
        out[0] = static_cast<uint16_t>(static_cast<uint16_t>(v.x + 1.0f) | 1u);
        out[1] = static_cast<uint16_t>(static_cast<uint16_t>(v.y + 1.0f) | 1u);
        out[2] = static_cast<uint16_t>(static_cast<uint16_t>(v.z + 1.0f) | 1u);
    } else {
        out[0] = static_cast<uint16_t>(static_cast<uint16_t>(v.x) & 0xfffeu);
        out[1] = static_cast<uint16_t>(static_cast<uint16_t>(v.y) & 0xfffeu);
        out[2] = static_cast<uint16_t>(static_cast<uint16_t>(v.z) & 0xfffeu);
    }
}

RG_HD Vec3 UnquantizePoint(const MeshQuantization& q, const uint16_t in[3]) {
    return Vec3(q.bvh_min.x + static_cast<float>(in[0]) / q.scale.x, q.bvh_min.y + static_cast<float>(in[1]) / q.scale.y,
                q.bvh_min.z + static_cast<float>(in[2]) / q.scale.z);
}

RG_HD bool QuantizedOverlap(const uint16_t* amin, const uint16_t* amax, const uint16_t* bmin, const uint16_t* bmax) {
    return !(amin[0] > bmax[0] || amax[0] < bmin[0] || amin[1] > bmax[1] || amax[1] < bmin[1] || amin[2] > bmax[2] ||
             amax[2] < bmin[2]);
}

// Stackless depth-first walk of one mesh's tree: calls leaf(local_tri) for every leaf whose quantized box overlaps
// the query box and passes `accept`, in node order.
template <typename Leaf, typename Accept>
RG_HD void WalkMeshBvh(const MeshBvhNode* nodes, int begin, int end, const uint16_t* qmin, const uint16_t* qmax,
                       Leaf&& leaf, Accept&& accept) {
    int i = begin;
    while (i < end) {
        const MeshBvhNode node = nodes[i];
        const bool overlap = QuantizedOverlap(qmin, qmax, node.qmin, node.qmax);
        const bool is_leaf = node.escape_or_tri >= 0;
        if (is_leaf && overlap && accept(node.qmin, node.qmax)) leaf(node.escape_or_tri);
        i += (overlap || is_leaf) ? 1 : -node.escape_or_tri;
    }
}

struct AcceptAll {
    RG_HD bool operator()(const uint16_t*, const uint16_t*) const { return true; }
};

// How a query walks one mesh tree.  A walk policy runs, for every leaf the stackless walk visits (in node order),
// either serial_leaf(local_tri), or compute(local_tri) then commit(result): compute may run for several leaves at
// once on other lanes; commit runs in node order, and may be left out for a result whose keep() is false (its
// commit changes nothing).  Both forms give the same result.  SerialWalk is the one-lane
// walk; a cooperative device walk lives with the kernels.
struct SerialWalk {
    // Returns the walk's dependent memory rounds as a group walk counts them (0: not counted here).
    template <typename Accept, typename SerialLeaf, typename Compute, typename Commit>
    RG_HD static int Run(const MeshBvhNode* nodes, int begin, int end, const uint16_t* qmin, const uint16_t* qmax,
                         Accept&& accept, SerialLeaf&& serial_leaf, Compute&&, Commit&&) {
        WalkMeshBvh(nodes, begin, end, qmin, qmax, serial_leaf, accept);
        return 0;
    }
};

// ------------------------------------------------------------------------------------------------------------
// Ray cast (Bullet's btTriangleRaycastCallback over btBvhTriangleMeshShape::performRaycast; reference
// host_validation_and_contact_debug.hpp:1311-1378 and :1683-1944).

struct RayHit {
    float fraction = 1.0f;
    Vec3 normal = Vec3(0.0f, 0.0f, 1.0f);
    int32_t tri = -1;
    uint8_t hit = 0;
    // Hits nearer than sqrt(min_fraction_sq) of the segment are ignored (a wheel ray starts inside the car's own
    // hitbox; ArenaRayHit::cal_min_fraction_sq, arena_ball.hpp:3192-3204).  0: no limit.
    float min_fraction_sq = 0.0f;
};

// Slab test of a leaf box against the ray (mesh_bvh.hpp:522-591: Bullet's btRayAabb2 with a normalized direction).
struct RayLeafTest {
    const MeshQuantization* q;
    Vec3 source;
    Vec3 inv_dir;
    float lambda_max;

    RG_HD RayLeafTest(const MeshQuantization& quant, const Vec3& from, const Vec3& to) : q(&quant), source(from) {
        const Vec3 d = to - from;
        const float r = 1.0f / sqrtf(d.y * d.y + d.x * d.x + d.z * d.z);
        const Vec3 n(d.x * r, d.y * r, d.z * r);
        lambda_max = n.x * d.x + n.y * d.y + n.z * d.z;
        inv_dir = Vec3(n.x == 0.0f ? 1e18f : 1.0f / n.x, n.y == 0.0f ? 1e18f : 1.0f / n.y,
                       n.z == 0.0f ? 1e18f : 1.0f / n.z);
    }
    RG_HD bool operator()(const uint16_t* qmin, const uint16_t* qmax) const {
        const Vec3 lo = UnquantizePoint(*q, qmin);
        const Vec3 hi = UnquantizePoint(*q, qmax);
        const float tmin = ((inv_dir.x < 0.0f ? hi.x : lo.x) - source.x) * inv_dir.x;
        const float tmax = ((inv_dir.x < 0.0f ? lo.x : hi.x) - source.x) * inv_dir.x;
        const float tymin = ((inv_dir.y < 0.0f ? hi.y : lo.y) - source.y) * inv_dir.y;
        const float tymax = ((inv_dir.y < 0.0f ? lo.y : hi.y) - source.y) * inv_dir.y;
        if (!(tmin <= tymax) || !(tymin <= tmax)) return false;
        const float t0 = tymin <= tmin ? tmin : tymin;
        const float t1 = tmax <= tymax ? tmax : tymax;
        const float tzmin = ((inv_dir.z < 0.0f ? hi.z : lo.z) - source.z) * inv_dir.z;
        const float tzmax = ((inv_dir.z < 0.0f ? lo.z : hi.z) - source.z) * inv_dir.z;
        if (!(t0 <= tzmax) || !(tzmin <= t1)) return false;
        return (tzmin <= t0 ? t0 : tzmin) < lambda_max && 0.0f < (t1 <= tzmax ? t1 : tzmax);
    }
};

// One triangle against the ray from -> to; updates best when strictly closer.
RG_HD void RayTriangle(const MeshTriangle& tri, int32_t tri_idx, const Vec3& from, const Vec3& to, RayHit& best) {
    const Vec3 v10 = tri.b - tri.a;
    const Vec3 v20 = tri.c - tri.a;
    Vec3 n = Cross(v10, v20);
    const float plane = Dot(tri.a, n);
    const float dist_a = Dot(n, from) - plane;
    const float dist_b = Dot(n, to) - plane;
    if (dist_a * dist_b >= 0.0f) return;
    const float proj = dist_a - dist_b;
    if (proj == 0.0f) return;
    const float distance = dist_a / proj;
    if (distance >= best.fraction) return;
    const float tol = -0.0001f * LengthSq(n);
    const Vec3 p = (1.0f - distance) * from + distance * to;
    const Vec3 v0p = tri.a - p;
    const Vec3 v1p = tri.b - p;
    if (Dot(Cross(v0p, v1p), n) < tol) return;
    const Vec3 v2p = tri.c - p;
    if (Dot(Cross(v1p, v2p), n) < tol) return;
    if (Dot(Cross(v2p, v0p), n) < tol) return;
    n = Normalize(n, tri.normal);
    if (dist_a <= 0.0f) n = -n;
    if (distance < 0.0f || distance > 1.0f) return;
    if (best.min_fraction_sq > 0.0f && distance * distance < best.min_fraction_sq) return;
    if (!best.hit || distance < best.fraction) {
        best.hit = 1;
        best.fraction = distance;
        best.normal = n;
        best.tri = tri_idx;
    }
}

// One triangle against the ray, without RayTriangle's early rejection against the best hit so far (valid: the hit
// RayTriangle would record if the best hit were farther).  CommitRayCandidate then applies RayTriangle's
// comparisons with the best hit, in its order; the pair records exactly what RayTriangle records.
struct RayCandidate {
    float distance;
    Vec3 normal;
    int32_t tri;
    int32_t valid;
    RG_HD bool keep() const { return valid != 0; }  // CommitRayCandidate changes nothing otherwise
};

RG_HD RayCandidate RayTriangleCandidate(const MeshTriangle& tri, int32_t tri_idx, const Vec3& from, const Vec3& to) {
    RayCandidate c{0.0f, Vec3(0.0f, 0.0f, 0.0f), tri_idx, 0};
    const Vec3 v10 = tri.b - tri.a;
    const Vec3 v20 = tri.c - tri.a;
    Vec3 n = Cross(v10, v20);
    const float plane = Dot(tri.a, n);
    const float dist_a = Dot(n, from) - plane;
    const float dist_b = Dot(n, to) - plane;
    if (dist_a * dist_b >= 0.0f) return c;
    const float proj = dist_a - dist_b;
    if (proj == 0.0f) return c;
    const float distance = dist_a / proj;
    c.distance = distance;
    const float tol = -0.0001f * LengthSq(n);
    const Vec3 p = (1.0f - distance) * from + distance * to;
    const Vec3 v0p = tri.a - p;
    const Vec3 v1p = tri.b - p;
    if (Dot(Cross(v0p, v1p), n) < tol) return c;
    const Vec3 v2p = tri.c - p;
    if (Dot(Cross(v1p, v2p), n) < tol) return c;
    if (Dot(Cross(v2p, v0p), n) < tol) return c;
    n = Normalize(n, tri.normal);
    if (dist_a <= 0.0f) n = -n;
    if (distance < 0.0f || distance > 1.0f) return c;
    c.normal = n;
    c.valid = 1;
    return c;
}

RG_HD void CommitRayCandidate(const RayCandidate& c, RayHit& best) {
    if (!c.valid || c.distance >= best.fraction) return;
    if (best.min_fraction_sq > 0.0f && c.distance * c.distance < best.min_fraction_sq) return;
    if (!best.hit || c.distance < best.fraction) {
        best.hit = 1;
        best.fraction = c.distance;
        best.normal = c.normal;
        best.tri = c.tri;
    }
}

// Nearest hit of the segment from -> from + delta over every mesh (meshes in order, leaves in node order).
template <typename Walk = SerialWalk>
RG_HD RayHit CastArenaRay(const ArenaMeshView& m, const Vec3& from, const Vec3& delta, float min_fraction_sq = 0.0f) {
    RayHit best;
    best.min_fraction_sq = min_fraction_sq;
    const Vec3 to = from + delta;
    const Vec3 lo = Min3(from, to), hi = Max3(from, to);
    for (int k = 0; k < m.mesh_count; ++k) {
        const MeshInfo& info = m.meshes[m.meshes[k].ray_order];
        if (info.aabb_max.x < lo.x || info.aabb_min.x > hi.x || info.aabb_max.y < lo.y || info.aabb_min.y > hi.y ||
            info.aabb_max.z < lo.z || info.aabb_min.z > hi.z) {
            continue;
        }
        uint16_t qmin[3], qmax[3];
        QuantizePoint(info.q, lo, 0, qmin);
        QuantizePoint(info.q, hi, 1, qmax);
        const RayLeafTest test(info.q, from, to);
        const int32_t tri_start = info.tri_start;
        Walk::Run(
            m.nodes, info.node_start, info.node_start + info.node_count, qmin, qmax, test,
            [&](int32_t local) { RayTriangle(m.tris[tri_start + local], tri_start + local, from, to, best); },
            [&](int32_t local) { return RayTriangleCandidate(m.tris[tri_start + local], tri_start + local, from, to); },
            [&](const RayCandidate& c) { CommitRayCandidate(c, best); });
    }
    return best;
}

// Leaf filter of a sphere query: keeps the leaves whose (dequantized) box comes within `reach` of the center.  A
// triangle lies in its leaf's box, and SphereTriangleContact reports a contact only for a point of the triangle
// (computed to about 1e-4 bt) closer than r + threshold to the center, so with reach = r + threshold + 0.01 bt
// every leaf this drops would have given no contact: the walk's result is unchanged, without the triangle's load
// and test.
struct SphereNearLeaf {
    const MeshQuantization* q;
    Vec3 center;
    float reach_sq;

    RG_HD SphereNearLeaf(const MeshQuantization& quant, const Vec3& c, float reach)
        : q(&quant), center(c), reach_sq(reach * reach) {}
    RG_HD bool operator()(const uint16_t* qmin, const uint16_t* qmax) const {
        const Vec3 lo = UnquantizePoint(*q, qmin);
        const Vec3 hi = UnquantizePoint(*q, qmax);
        const float dx = Max(Max(lo.x - center.x, center.x - hi.x), 0.0f);
        const float dy = Max(Max(lo.y - center.y, center.y - hi.y), 0.0f);
        const float dz = Max(Max(lo.z - center.z, center.z - hi.z), 0.0f);
        return dx * dx + dy * dy + dz * dz <= reach_sq;
    }
};

// ------------------------------------------------------------------------------------------------------------
// Sphere against triangle (Bullet's SphereTriangleDetector::collide; reference arena_geometry.hpp:1905-1990).

RG_HD bool PointInTriangleBt(const Vec3& p, const Vec3& a, const Vec3& b, const Vec3& c, const Vec3& n) {
    const float r1 = Dot(Cross(b - a, n), p - a);
    const float r2 = Dot(Cross(c - b, n), p - b);
    const float r3 = Dot(Cross(a - c, n), p - c);
    return (r1 > 0.0f && r2 > 0.0f && r3 > 0.0f) || (r1 <= 0.0f && r2 <= 0.0f && r3 <= 0.0f);
}

RG_HD bool NearestEdgePoint(const Vec3& p, const Vec3& a, const Vec3& b, const Vec3& c, float max_dist_sq,
                            Vec3& out) {
    const Vec3 from[3] = {a, b, c};
    const Vec3 to[3] = {b, c, a};
    bool found = false;
    for (int i = 0; i < 3; ++i) {
        Vec3 diff = p - from[i];
        const Vec3 v = to[i] - from[i];
        float t = Dot(v, diff);
        if (t > 0.0f) {
            const float vv = Dot(v, v);
            if (t < vv) {
                t /= vv;
                diff -= v * t;
            } else {
                t = 1.0f;
                diff -= v;
            }
        } else {
            t = 0.0f;
        }
        const float d2 = Dot(diff, diff);
        if (d2 < max_dist_sq) {
            max_dist_sq = d2;
            out = from[i] + v * t;
            found = true;
        }
    }
    return found;
}

// Contact of a sphere (center c, radius r, breaking threshold th) with a triangle.  On contact: unit normal from the
// triangle toward the sphere center, the contact point on the triangle, the signed distance (negative when
// penetrating) and the penetration in uu.  Back faces are rejected, as the reference does for the ball.
RG_HD bool SphereTriangleContact(const Vec3& c, float r, float th, const Vec3& a, const Vec3& b, const Vec3& cc,
                                 Vec3& out_normal, Vec3& out_point, float& out_depth, float& out_pen_uu) {
    const float r_th = r + th;
    Vec3 n = Cross(b - a, cc - a);
    const float l2 = Dot(n, n);
    if (l2 < kSimdEpsilon * kSimdEpsilon) return false;
    n = n * (1.0f / sqrtf(l2));
    const float dist = Dot(c - a, n);
    if (dist < 0.0f) return false;
    if (dist >= r_th) return false;
    Vec3 point(0.0f, 0.0f, 0.0f);
    bool has = false;
    if (PointInTriangleBt(c, a, b, cc, n)) {
        has = true;
        point = c - n * dist;
    } else {
        has = NearestEdgePoint(c, a, b, cc, r_th * r_th, point);
    }
    if (!has) return false;
    const Vec3 to_center = c - point;
    const float d2 = Dot(to_center, to_center);
    if (d2 >= r_th * r_th) return false;
    if (d2 > kSimdEpsilon) {
        const float d = sqrtf(d2);
        out_normal = to_center * (1.0f / d);
        out_pen_uu = (r - d) * kBtToUu;
        out_depth = -(r - d);
    } else {
        out_normal = n;
        out_pen_uu = r * kBtToUu;
        out_depth = -r;
    }
    out_point = point;
    return true;
}

// ------------------------------------------------------------------------------------------------------------
// Internal-edge contact normal correction (Bullet's btAdjustInternalEdgeContacts with the game's mesh rules;
// reference raycast_wheels_and_contact_prep.hpp:5042-5177).

RG_HD float TriEdgeAngle(const MeshTriangle& t, int e) {
    return e == 0 ? t.edge_angle_v0v1 : (e == 1 ? t.edge_angle_v1v2 : t.edge_angle_v2v0);
}

RG_HD Vec3 ClosestOnSegment(const Vec3& p, const Vec3& a, const Vec3& b, float lo, float hi) {
    const Vec3 seg = b - a;
    const float len_sq = LengthSq(seg);
    if (len_sq <= 1.0e-12f) return a;
    const float t = Clamp(Dot(p - a, seg) / len_sq, lo, hi);
    return a + seg * t;
}

// Rotate v by the matrix of q (the game's setRotation columns).
RG_HD Vec3 MatrixRotate(const Quat& q, const Vec3& v) {
    Vec3 c0, c1, c2;
    GameQuatToColumns(q, c0, c1, c2);
    return Vec3(c0.x * v.x + c1.x * v.y + c2.x * v.z, c0.y * v.x + c1.y * v.y + c2.y * v.z,
                c0.z * v.x + c1.z * v.y + c2.z * v.z);
}

RG_HD bool ClampEdgeNormal(const Vec3& edge, const Vec3& tri_normal, const Vec3& contact_normal, float edge_angle,
                           Vec3& out) {
    const Vec3 edge_cross = Normalize(Cross(edge, tri_normal), tri_normal);
    const float cur = UcrtAtan2f(Dot(contact_normal, edge_cross), Dot(contact_normal, tri_normal));
    if (edge_angle < 0.0f) {
        if (cur < edge_angle) {
            out = MatrixRotate(QuatFromAxisAngle(edge, edge_angle - cur), contact_normal);
            return true;
        }
    } else if (cur > edge_angle) {
        out = MatrixRotate(QuatFromAxisAngle(edge, edge_angle - cur), contact_normal);
        return true;
    }
    return false;
}

RG_HD void AdjustInternalEdgeContact(const MeshTriangle& tri, const Vec3& point_on_a, float distance,
                                     Vec3& io_point_on_b, Vec3& io_normal) {
    constexpr float kLo = 1.0e-5f, kHi = 1.0f - 1.0e-5f;
    constexpr float kEdgeDistance = 0.1f;
    constexpr float kMaxAngle = 2.0f * Pi();
    const Vec3 tri_normal = tri.normal;
    const float contact_len_sq = LengthSq(io_normal);
    Vec3 contact_normal = tri_normal;
    if (contact_len_sq > 1e-12f) contact_normal = io_normal * (1.0f / sqrtf(contact_len_sq));
    const Vec3 contact = io_point_on_b;
    const Vec3 es[3] = {tri.a, tri.b, tri.c};
    const Vec3 ee[3] = {tri.b, tri.c, tri.a};

    int best_edge = -1;
    float best_dist = 1.0e30f;
    uint32_t on_edge = 0u;
    for (int e = 0; e < 3; ++e) {
        if (Abs(TriEdgeAngle(tri, e)) >= kMaxAngle) continue;
        const float d = Length(contact - ClosestOnSegment(contact, es[e], ee[e], kLo, kHi));
        if (d < 1.1920928955078125e-07f) on_edge |= 1u << e;
        if (d < best_dist) {
            best_dist = d;
            best_edge = e;
        }
    }
    if (best_edge < 0) return;

    bool near_edge = false;
    int concave_hits = 0;
    for (int e = 0; e < 3; ++e) {
        if (best_edge != e && (on_edge & (1u << e)) == 0u) continue;
        const float angle = TriEdgeAngle(tri, e);
        if (Abs(angle) >= kMaxAngle) continue;
        const float d = Length(contact - ClosestOnSegment(contact, es[e], ee[e], 0.0f, 1.0f));
        if (d >= kEdgeDistance) continue;
        near_edge = true;
        if (angle == 0.0f) {
            ++concave_hits;
            continue;
        }
        const bool convex = (tri.flags & kTriEdgeConvex[e]) != 0;
        const float swap = convex ? 1.0f : -1.0f;
        const Vec3 nA = tri_normal * swap;
        const Vec3 edge = es[e] - ee[e];
        Vec3 nb = QuatRotate(QuatFromAxisAngle(edge, angle), tri_normal);
        if ((tri.flags & kTriEdgeSwapNormalB[e]) != 0) nb = -nb;
        const Vec3 nB = nb * swap;
        if (Dot(contact_normal, nA) < 0.0f && Dot(contact_normal, nB) < 0.0f) {
            ++concave_hits;
            continue;
        }
        Vec3 clamped(0.0f, 0.0f, 1.0f);
        const bool did = ClampEdgeNormal(edge, tri_normal * swap, e == 0 ? contact_normal : io_normal, angle, clamped);
        if (did && Dot(clamped, tri_normal) > 0.0f) {
            io_normal = clamped;
            io_point_on_b = point_on_a - io_normal * distance;
        }
    }
    if (!near_edge) return;
    if (concave_hits > 0) {
        if (Dot(tri_normal, contact_normal) < 0.0f) return;
        io_normal = tri_normal;
        io_point_on_b = point_on_a - io_normal * distance;
    }
}

}  // namespace core
}  // namespace rg
