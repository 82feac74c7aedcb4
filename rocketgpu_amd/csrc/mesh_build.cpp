// Arena mesh preprocessing.  Reproduces, in this project's own code, the reference's custom-layout mesh pipeline
// (src/arena/mesh_loader.cpp:1263-1412): Bullet's btGenerateInternalEdgeInfo for adjacent triangles (:57-202),
// btQuantizedBvh quantization with the Bullet 2.82 bound fix (include/rocketgpu/detail/arena/mesh_bvh.hpp:321-383)
// and btQuantizedBvh::buildTree with its variance split (mesh_loader.cpp:210-396).  The node order of the tree fixes
// the order in which contacts are generated, so the build must match exactly.
#include "mesh_build.hpp"

#include "core/rg_broadphase.hpp"

#include <algorithm>
#include <cmath>
#include <limits>
#include <stdexcept>
#include <unordered_map>

namespace rg {
namespace {

using namespace core;

constexpr float kEqualVertexSq = 1.0e-4f * 1.0e-4f;

// btGenerateInternalEdgeInfo for the ordered pair (tri_a, tri_b): when the two share exactly one edge, record on
// tri_a the corrected angle of that edge and its convex / swap flags.
void InternalEdge(std::vector<MeshTriangle>& tris, int ia, int ib) {
    const MeshTriangle& A = tris[static_cast<size_t>(ia)];
    const MeshTriangle& B = tris[static_cast<size_t>(ib)];
    const Vec3 va[3] = {A.a, A.b, A.c};
    const Vec3 vb[3] = {B.a, B.b, B.c};
    int sa[2] = {-1, -1}, sb[2] = {-1, -1};
    int shared = 0;
    for (int i = 0; i < 3; ++i) {
        for (int j = 0; j < 3; ++j) {
            if (LengthSq(va[i] - vb[j]) < kEqualVertexSq) {
                sa[shared] = i;
                sb[shared] = j;
                if (++shared >= 3) return;
                break;
            }
        }
    }
    if (shared != 2) return;
    if (sa[0] == 0 && sa[1] == 2) {
        std::swap(sa[0], sa[1]);
        std::swap(sb[0], sb[1]);
    }
    const int sum_a = sa[0] + sa[1];
    const int other_a = 3 - sum_a;
    const int other_b = 3 - (sb[0] + sb[1]);
    Vec3 edge = va[sa[1]] - va[sa[0]];
    if (LengthSq(edge) <= 1.0e-12f) return;
    edge = Normalize(edge, Vec3(1.0f, 0.0f, 0.0f));
    const Vec3 normal_a = Normalize(Cross(va[1] - va[0], va[2] - va[0]), Vec3(0.0f, 0.0f, 1.0f));
    const Vec3 b0 = vb[sb[1]], b1 = vb[sb[0]], b2 = vb[other_b];
    const Vec3 normal_b = Normalize(Cross(b1 - b0, b2 - b0), Normalize(B.normal, Vec3(0.0f, 0.0f, 1.0f)));
    Vec3 cross_a = Normalize(Cross(edge, normal_a), normal_a);
    if (Dot(cross_a, va[other_a] - va[sa[0]]) < 0.0f) cross_a = -cross_a;
    Vec3 cross_b = Normalize(Cross(edge, normal_b), normal_b);
    if (Dot(cross_b, vb[other_b] - vb[sb[0]]) < 0.0f) cross_b = -cross_b;
    float corrected = 0.0f;
    bool convex = false;
    const Vec3 calc_edge = Cross(cross_a, cross_b);
    if (LengthSq(calc_edge) >= 0.0001f) {
        const Vec3 calc_edge_unit = Normalize(calc_edge, edge);
        const Vec3 calc_normal_a = Normalize(Cross(calc_edge_unit, cross_a), normal_a);
        const float angle2 = UcrtAtan2f(Dot(cross_b, calc_normal_a), Dot(cross_b, cross_a));
        const float ang4 = Pi() - angle2;
        convex = Dot(normal_a, cross_b) < 0.0f;
        corrected = convex ? ang4 : -ang4;
    }
    MeshTriangle& T = tris[static_cast<size_t>(ia)];
    const auto set = [&](float& angle, int e, const Vec3& axis) {
        Vec3 computed_b = QuatRotate(QuatFromAxisAngle(axis, -corrected), normal_a);
        if (Dot(computed_b, normal_b) < 0.0f) T.flags |= kTriEdgeSwapNormalB[e];
        angle = -corrected;
        if (convex) T.flags |= kTriEdgeConvex[e];
    };
    switch (sum_a) {
    case 1: set(T.edge_angle_v0v1, 0, va[0] - va[1]); break;
    case 2: set(T.edge_angle_v2v0, 2, va[2] - va[0]); break;
    case 3: set(T.edge_angle_v1v2, 1, va[1] - va[2]); break;
    default: break;
    }
}

// Quantization of one mesh's box (SetArenaMeshBvhQuantization with the Bullet 2.82 bounds, margin 1).
MeshQuantization MakeQuantization(const Vec3& mn, const Vec3& mx) {
    MeshQuantization q;
    const Vec3 clamp(1.0f, 1.0f, 1.0f);
    q.bvh_min = mn - clamp;
    q.bvh_max = mx + clamp;
    Vec3 size = q.bvh_max - q.bvh_min;
    q.scale = Div3(Vec3(65533.0f, 65533.0f, 65533.0f), size);
    uint16_t qmin[3];
    const Vec3 cur_min = q.bvh_min;
    QuantizePoint(q, cur_min, 0, qmin);
    const Vec3 un_min(static_cast<float>(qmin[0]) / q.scale.x, static_cast<float>(qmin[1]) / q.scale.y,
                      static_cast<float>(qmin[2]) / q.scale.z);
    q.bvh_min = Min3(cur_min, cur_min + un_min - clamp);
    const Vec3 new_min = q.bvh_min;
    const Vec3 cur_max = q.bvh_max;
    const float qx = (cur_max.x - new_min.x) * q.scale.x;
    const float qy = (cur_max.y - new_min.y) * q.scale.y;
    const float qz = (cur_max.z - new_min.z) * q.scale.z;
    const uint16_t qmax[3] = {static_cast<uint16_t>(static_cast<uint16_t>(static_cast<int>(qx + 1.0f)) | 1),
                              static_cast<uint16_t>(static_cast<uint16_t>(static_cast<int>(qy + 1.0f)) | 1),
                              static_cast<uint16_t>(static_cast<uint16_t>(static_cast<int>(qz + 1.0f)) | 1)};
    const Vec3 un_max(static_cast<float>(qmax[0]) / q.scale.x, static_cast<float>(qmax[1]) / q.scale.y,
                      static_cast<float>(qmax[2]) / q.scale.z);
    q.bvh_max = Max3(cur_max, new_min + un_max + clamp);
    size = q.bvh_max - q.bvh_min;
    q.scale = Div3(Vec3(65533.0f, 65533.0f, 65533.0f), size);
    return q;
}

// btQuantizedBvh::buildTree over one mesh's triangles; appends nodes in depth-first order.
class BvhBuilder {
public:
    BvhBuilder(const MeshQuantization& q, std::vector<MeshBvhNode>& leaves) : q_(q), leaves_(leaves) {}

    std::vector<MeshBvhNode> Build() {
        nodes_.assign(2 * leaves_.size(), MeshBvhNode{});
        cur_ = 0;
        Recurse(0, static_cast<int>(leaves_.size()));
        nodes_.resize(static_cast<size_t>(cur_));
        return nodes_;
    }

private:
    Vec3 Center(int i) const {
        const MeshBvhNode& n = leaves_[static_cast<size_t>(i)];
        return (UnquantizePoint(q_, n.qmax) + UnquantizePoint(q_, n.qmin)) * 0.5f;
    }
    Vec3 Mean(int s, int e) const {
        Vec3 m(0.0f, 0.0f, 0.0f);
        for (int i = s; i < e; ++i) m = m + Center(i);
        return m * (1.0f / static_cast<float>(e - s));
    }
    int SplitAxis(int s, int e) const {
        const Vec3 m = Mean(s, e);
        Vec3 var(0.0f, 0.0f, 0.0f);
        for (int i = s; i < e; ++i) {
            const Vec3 d = Center(i) - m;
            var = var + Mul3(d, d);
        }
        var = var * (1.0f / static_cast<float>(e - s - 1));
        return var.x < var.y ? (var.y < var.z ? 2 : 1) : (var.x < var.z ? 2 : 0);
    }
    int SplitIndex(int s, int e, int axis) {
        const Vec3 m = Mean(s, e);
        const float split = (&m.x)[axis];
        int idx = s;
        for (int i = s; i < e; ++i) {
            const Vec3 c = Center(i);
            if ((&c.x)[axis] > split) {
                std::swap(leaves_[static_cast<size_t>(i)], leaves_[static_cast<size_t>(idx)]);
                ++idx;
            }
        }
        const int n = e - s;
        const int balanced = n / 3;
        if (idx <= s + balanced || idx >= e - 1 - balanced) idx = s + (n >> 1);
        return idx;
    }
    void Recurse(int s, int e) {
        const int start_node = cur_;
        if (e - s == 1) {
            nodes_[static_cast<size_t>(cur_++)] = leaves_[static_cast<size_t>(s)];
            return;
        }
        const int axis = SplitAxis(s, e);
        const int split = SplitIndex(s, e, axis);
        MeshBvhNode& node = nodes_[static_cast<size_t>(cur_)];
        QuantizePoint(q_, q_.bvh_max, 0, node.qmin);
        QuantizePoint(q_, q_.bvh_min, 1, node.qmax);
        for (int i = s; i < e; ++i) {
            const MeshBvhNode& leaf = leaves_[static_cast<size_t>(i)];
            uint16_t lmin[3], lmax[3];
            QuantizePoint(q_, UnquantizePoint(q_, leaf.qmin), 0, lmin);
            QuantizePoint(q_, UnquantizePoint(q_, leaf.qmax), 1, lmax);
            for (int a = 0; a < 3; ++a) {
                if (node.qmin[a] > lmin[a]) node.qmin[a] = lmin[a];
                if (node.qmax[a] < lmax[a]) node.qmax[a] = lmax[a];
            }
        }
        ++cur_;
        Recurse(s, split);
        Recurse(split, e);
        nodes_[static_cast<size_t>(start_node)].escape_or_tri = -(cur_ - start_node);
    }

    const MeshQuantization& q_;
    std::vector<MeshBvhNode>& leaves_;
    std::vector<MeshBvhNode> nodes_;
    int cur_ = 0;
};

}  // namespace

std::vector<MeshSource> GroupLayoutMeshes(const std::vector<std::vector<Vec3>>& vertices_uu,
                                          const std::vector<std::vector<int32_t>>& indices) {
    if (vertices_uu.size() != indices.size() || vertices_uu.empty()) {
        throw std::invalid_argument("a custom arena needs at least one mesh with matching vertex and index lists");
    }
    const size_t n = vertices_uu.size();
    const size_t groups = std::min<size_t>(n, kMaxArenaMeshes);
    std::vector<MeshSource> out(groups);
    for (size_t i = 0; i < n; ++i) {
        MeshSource& m = out[i * groups / n];
        const int32_t base = static_cast<int32_t>(m.vertices_bt.size());
        for (const Vec3& v : vertices_uu[i]) {
            m.vertices_bt.push_back(Vec3(v.x * kUuToBt, v.y * kUuToBt, v.z * kUuToBt));
        }
        for (int32_t idx : indices[i]) m.indices.push_back(idx < 0 ? idx : idx + base);
    }
    return out;
}

ArenaMeshHost BuildArenaMesh(const std::vector<MeshSource>& sources) {
    if (sources.empty() || sources.size() > static_cast<size_t>(kMaxArenaMeshes)) {
        throw std::invalid_argument("a custom arena needs 1 to 16 meshes");
    }
    ArenaMeshHost out;
    for (size_t mi = 0; mi < sources.size(); ++mi) {
        const MeshSource& src = sources[mi];
        const int n = static_cast<int>(src.indices.size() / 3);
        if (n <= 0 || src.indices.size() % 3 != 0) {
            throw std::invalid_argument("a custom arena mesh needs a whole number of triangles");
        }
        const size_t base = out.tris.size();
        const float inf = std::numeric_limits<float>::infinity();
        Vec3 mesh_min(inf, inf, inf), mesh_max(-inf, -inf, -inf);
        for (int t = 0; t < n; ++t) {
            Vec3 v[3];
            for (int k = 0; k < 3; ++k) {
                const int32_t idx = src.indices[static_cast<size_t>(t) * 3 + k];
                if (idx < 0 || static_cast<size_t>(idx) >= src.vertices_bt.size()) {
                    throw std::invalid_argument("custom arena mesh triangle index out of range");
                }
                v[k] = src.vertices_bt[static_cast<size_t>(idx)];
                if (!std::isfinite(v[k].x) || !std::isfinite(v[k].y) || !std::isfinite(v[k].z)) {
                    throw std::invalid_argument("custom arena mesh vertex is not finite");
                }
            }
            MeshTriangle tri{};
            tri.a = v[0];
            tri.b = v[1];
            tri.c = v[2];
            tri.edge_angle_v0v1 = tri.edge_angle_v1v2 = tri.edge_angle_v2v0 = 2.0f * Pi();
            const Vec3 ab(tri.b.x - tri.a.x, tri.b.y - tri.a.y, tri.b.z - tri.a.z);
            const Vec3 ac(tri.c.x - tri.a.x, tri.c.y - tri.a.y, tri.c.z - tri.a.z);
            tri.normal = Normalize(Cross(ab, ac), Vec3(0.0f, 0.0f, 1.0f));
            tri.min = Vec3(Min(tri.a.x, Min(tri.b.x, tri.c.x)), Min(tri.a.y, Min(tri.b.y, tri.c.y)),
                           Min(tri.a.z, Min(tri.b.z, tri.c.z)));
            tri.max = Vec3(Max(tri.a.x, Max(tri.b.x, tri.c.x)), Max(tri.a.y, Max(tri.b.y, tri.c.y)),
                           Max(tri.a.z, Max(tri.b.z, tri.c.z)));
            tri.flags = static_cast<uint32_t>(mi) << 8;
            mesh_min = Min3(mesh_min, tri.min);
            mesh_max = Max3(mesh_max, tri.max);
            out.tris.push_back(tri);
        }
        // Adjacency: triangles sharing a vertex (vertices bucketed on a 1e-3 bt grid, 3x3x3 neighbourhood), each
        // triangle's neighbours visited in increasing index order.
        constexpr float kCell = 1.0e-3f;
        const auto cell = [&](const Vec3& p, int64_t c[3]) {
            c[0] = static_cast<int64_t>(std::floor(p.x / kCell));
            c[1] = static_cast<int64_t>(std::floor(p.y / kCell));
            c[2] = static_cast<int64_t>(std::floor(p.z / kCell));
        };
        const auto key = [](int64_t x, int64_t y, int64_t z) {
            return (static_cast<uint64_t>(x & 0x1fffff) << 42) | (static_cast<uint64_t>(y & 0x1fffff) << 21) |
                   static_cast<uint64_t>(z & 0x1fffff);
        };
        std::unordered_map<uint64_t, std::vector<int32_t>> buckets;
        for (int t = 0; t < n; ++t) {
            const MeshTriangle& tri = out.tris[base + static_cast<size_t>(t)];
            for (const Vec3* p : {&tri.a, &tri.b, &tri.c}) {
                int64_t c[3];
                cell(*p, c);
                std::vector<int32_t>& l = buckets[key(c[0], c[1], c[2])];
                if (l.empty() || l.back() != t) l.push_back(t);
            }
        }
        std::vector<int32_t> cand;
        for (int ta = 0; ta < n; ++ta) {
            const MeshTriangle& A = out.tris[base + static_cast<size_t>(ta)];
            cand.clear();
            for (const Vec3* p : {&A.a, &A.b, &A.c}) {
                int64_t c[3];
                cell(*p, c);
                for (int dx = -1; dx <= 1; ++dx)
                    for (int dy = -1; dy <= 1; ++dy)
                        for (int dz = -1; dz <= 1; ++dz) {
                            const auto it = buckets.find(key(c[0] + dx, c[1] + dy, c[2] + dz));
                            if (it != buckets.end()) cand.insert(cand.end(), it->second.begin(), it->second.end());
                        }
            }
            std::sort(cand.begin(), cand.end());
            cand.erase(std::unique(cand.begin(), cand.end()), cand.end());
            for (const int32_t tb : cand) {
                if (tb == ta) continue;
                const MeshTriangle& B = out.tris[base + static_cast<size_t>(tb)];
                if (A.max.x < B.min.x || A.min.x > B.max.x || A.max.y < B.min.y || A.min.y > B.max.y ||
                    A.max.z < B.min.z || A.min.z > B.max.z) {
                    continue;
                }
                InternalEdge(out.tris, static_cast<int>(base) + ta, static_cast<int>(base) + tb);
            }
        }
        for (int t = 0; t < n; ++t) {
            MeshTriangle& w = out.tris[base + static_cast<size_t>(t)];
            if (w.edge_angle_v0v1 == 2.0f * Pi()) w.edge_angle_v0v1 = 0.0f;
            if (w.edge_angle_v1v2 == 2.0f * Pi()) w.edge_angle_v1v2 = 0.0f;
            if (w.edge_angle_v2v0 == 2.0f * Pi()) w.edge_angle_v2v0 = 0.0f;
        }
        const MeshQuantization q = MakeQuantization(mesh_min, mesh_max);
        std::vector<MeshBvhNode> leaves(static_cast<size_t>(n));
        for (int t = 0; t < n; ++t) {
            MeshTriangle& tri = out.tris[base + static_cast<size_t>(t)];
            QuantizePoint(q, tri.min, 0, tri.qmin);
            QuantizePoint(q, tri.max, 1, tri.qmax);
            Vec3 lo = tri.min, hi = tri.max;
            constexpr float kMinDim = 0.002f, kMinHalf = 0.001f;
            if (hi.x - lo.x < kMinDim) {
                hi.x += kMinHalf;
                lo.x -= kMinHalf;
            }
            if (hi.y - lo.y < kMinDim) {
                hi.y += kMinHalf;
                lo.y -= kMinHalf;
            }
            if (hi.z - lo.z < kMinDim) {
                hi.z += kMinHalf;
                lo.z -= kMinHalf;
            }
            MeshBvhNode& leaf = leaves[static_cast<size_t>(t)];
            QuantizePoint(q, lo, 0, leaf.qmin);
            QuantizePoint(q, hi, 1, leaf.qmax);
            leaf.escape_or_tri = t;
        }
        BvhBuilder builder(q, leaves);
        const std::vector<MeshBvhNode> nodes = builder.Build();
        MeshInfo info{};
        info.aabb_min = mesh_min;
        info.aabb_max = mesh_max;
        info.q = q;
        info.mesh_index = static_cast<int32_t>(mi);
        info.tri_start = static_cast<int32_t>(base);
        info.tri_count = n;
        info.node_start = static_cast<int32_t>(out.nodes.size());
        info.node_count = static_cast<int32_t>(nodes.size());
        out.meshes.push_back(info);
        out.nodes.insert(out.nodes.end(), nodes.begin(), nodes.end());
    }
    // Ray visit order: meshes by quantized max x, ties in mesh order (mesh_loader.cpp:1082-1098).
    std::vector<int32_t> order(out.meshes.size());
    for (size_t i = 0; i < order.size(); ++i) order[i] = static_cast<int32_t>(i);
    std::stable_sort(order.begin(), order.end(), [&](int32_t a, int32_t b) {
        return SweepQuantize(out.meshes[static_cast<size_t>(a)].aabb_max.x, 1) <
               SweepQuantize(out.meshes[static_cast<size_t>(b)].aabb_max.x, 1);
    });
    for (size_t k = 0; k < order.size(); ++k) out.meshes[k].ray_order = order[k];
    return out;
}

}  // namespace rg
