// Per-arena broadphase: Bullet's btAxisSweep3 (16-bit sweep-and-prune with incremental edge sorting), its overlapping
// pair list and the dispatcher's manifold list, kept across ticks because the manifold order decides the solver's
// row order.  Semantics follow the reference's BuildBulletManifoldList (include/rocketgpu/detail/step/
// arena_geometry.hpp:4162-5033), which reproduces the game's Bullet world; this is a re-implementation over a compact
// per-arena record sized for at most kMaxArenaMeshes static meshes and 9 dynamic bodies (ball + 8 cars).
#pragma once

#include "rg_mesh.hpp"

namespace rg {
namespace core {

constexpr int kMaxBodies = 9;                          // body 0: ball, body 1 + i: car slot i
constexpr int kSweepMaxHandles = kMaxArenaMeshes + kMaxBodies;  // handle 0 is the sentinel
constexpr int kSweepMaxEdges = 2 * kSweepMaxHandles + 2;
constexpr int kMaxManifoldEntries = 96;                // the reference's kMaxContactOrder
constexpr int kPredictiveStaticId = 99;
constexpr float kSweepHalfExtentBt = 10485.740234375f;

// A pair / manifold key: kind in bits 0-1 (0 ball-static, 1 ball-car, 2 car-static, 3 car-car), body_a + 1 in
// bits 2-5, body_b + 1 in bits 6-9, static id (mesh index or 99 for a predictive contact) in bits 16-31.
RG_HD constexpr uint32_t PackPairKey(int kind, int body_a, int body_b, int static_id) {
    return (static_cast<uint32_t>(kind) & 3u) | ((static_cast<uint32_t>(body_a + 1) & 15u) << 2) |
           ((static_cast<uint32_t>(body_b + 1) & 15u) << 6) | ((static_cast<uint32_t>(static_id) & 0xffffu) << 16);
}
RG_HD constexpr int KeyKind(uint32_t k) { return static_cast<int>(k & 3u); }
RG_HD constexpr int KeyBodyA(uint32_t k) { return static_cast<int>((k >> 2) & 15u) - 1; }
RG_HD constexpr int KeyBodyB(uint32_t k) { return static_cast<int>((k >> 6) & 15u) - 1; }
RG_HD constexpr int KeyStatic(uint32_t k) { return static_cast<int>(static_cast<int16_t>(k >> 16)); }

struct ManifoldEntry {
    uint32_t key;
    int32_t island;
};

struct SweepState {
    uint8_t handle_count;
    uint8_t static_count;
    uint8_t next_fresh;
    uint8_t free_count;
    uint8_t free_list[kMaxBodies];
    uint8_t body_handle[kMaxBodies];
    uint8_t obj_count;
    uint8_t obj[kMaxBodies];
    uint16_t edge_pos[3][kSweepMaxEdges];
    uint8_t edge_handle[3][kSweepMaxEdges];
    uint8_t min_edge[kSweepMaxHandles + 1][3];
    uint8_t max_edge[kSweepMaxHandles + 1][3];
    int16_t static_id[kSweepMaxHandles + 1];  // >= 0 static mesh index, < 0: -1 - body
    // Each body's swept box at its last update (bt: min xyz, max xyz), valid while its bit is set; wheel rays are
    // cast only against the bodies whose box they reach.
    uint16_t body_aabb_valid;
    float body_aabb[kMaxBodies][6];
};

struct BroadphaseState {
    SweepState sweep;
    uint8_t pair_count;
    uint8_t manifold_count;
    uint8_t predictive_count;
    uint8_t pad;
    uint32_t pairs[kMaxManifoldEntries];
    uint32_t manifolds[kMaxManifoldEntries];
    uint32_t predictive[kMaxBodies];
};

// Broadphase proxy of one body this tick: the swept AABB (SAP update) and the unswept child box (insertion and
// pair filtering).
struct BodyProxy {
    Vec3 aabb_min, aabb_max;
    Vec3 child_min, child_max;
    uint8_t valid;
};

RG_HD uint16_t SweepQuantize(float v, int is_max) {
    const float q = 65535.0f / (kSweepHalfExtentBt + kSweepHalfExtentBt);
    float s = (v + kSweepHalfExtentBt) * q;
    s = s >= 0.0f ? (s < 65535.0f ? s : 65535.0f) : 0.0f;
    return static_cast<uint16_t>((static_cast<int>(s) & 0xfffe) | is_max);
}

RG_HD bool AabbOverlap(const Vec3& amin, const Vec3& amax, const Vec3& bmin, const Vec3& bmax) {
    return !(amax.x < bmin.x || amin.x > bmax.x || amax.y < bmin.y || amin.y > bmax.y || amax.z < bmin.z ||
             amin.z > bmax.z);
}

RG_HD int FindKey(const uint32_t* list, int n, uint32_t key) {
    for (int i = 0; i < n; ++i) {
        if (list[i] == key) return i;
    }
    return -1;
}

RG_HD void SwapRemove(uint32_t* list, uint8_t& n, int i) {
    if (i >= 0) list[i] = list[--n];
}

// Incremental sweep-and-prune over the arena's handles.  `bodies` is the number of dynamic bodies this tick.
struct Sweep {
    BroadphaseState& bp;
    SweepState& h;
    int body_count;

    RG_HD Sweep(BroadphaseState& state, int bodies) : bp(state), h(state.sweep), body_count(bodies) {}

    RG_HD bool PairKeyFor(int ha, int hb, uint32_t& key) const {
        if (ha > hb) {
            const int t = ha;
            ha = hb;
            hb = t;
        }
        if (h.static_id[hb] >= 0) return false;
        int body_b = -1 - h.static_id[hb];
        if (body_b < 0 || body_b >= body_count) return false;
        if (h.static_id[ha] >= 0) {
            key = PackPairKey(body_b == 0 ? 0 : 2, body_b - 1, -1, h.static_id[ha]);
        } else {
            int body_a = -1 - h.static_id[ha];
            if (body_a > body_b) {
                const int t = body_a;
                body_a = body_b;
                body_b = t;
            }
            key = PackPairKey(body_a == 0 ? 1 : 3, body_a == 0 ? body_b - 1 : body_a - 1, body_a == 0 ? -1 : body_b - 1,
                              -1);
        }
        return true;
    }
    RG_HD void AddPair(int ha, int hb) {
        uint32_t key = 0;
        if (PairKeyFor(ha, hb, key) && FindKey(bp.pairs, bp.pair_count, key) < 0 &&
            bp.pair_count < kMaxManifoldEntries) {
            bp.pairs[bp.pair_count++] = key;
        }
    }
    RG_HD void RemovePair(int ha, int hb) {
        uint32_t key = 0;
        if (!PairKeyFor(ha, hb, key)) return;
        const int at = FindKey(bp.pairs, bp.pair_count, key);
        if (at >= 0) {
            SwapRemove(bp.pairs, bp.pair_count, at);
            SwapRemove(bp.manifolds, bp.manifold_count, FindKey(bp.manifolds, bp.manifold_count, key));
        }
    }
    // Do handles a and b overlap on the two axes other than `axis` (edge-index comparison)?
    RG_HD bool Overlap2D(int a, int b, int axis) const {
        const int axis1 = (1 << axis) & 3;
        const int axis2 = (1 << axis1) & 3;
        return !(h.max_edge[a][axis1] < h.min_edge[b][axis1] || h.max_edge[b][axis1] < h.min_edge[a][axis1] ||
                 h.max_edge[a][axis2] < h.min_edge[b][axis2] || h.max_edge[b][axis2] < h.min_edge[a][axis2]);
    }
    RG_HD uint8_t& EdgeRef(int handle, int axis, bool is_max) {
        return is_max ? h.max_edge[handle][axis] : h.min_edge[handle][axis];
    }
    // Move the edge at `index` toward lower positions; `update` adds/removes pairs as edges cross.
    RG_HD void SortDown(int axis, int index, bool is_max, bool update) {
        const uint8_t owner = h.edge_handle[axis][index];
        const uint16_t pos = h.edge_pos[axis][index];
        while (index > 0 && pos < h.edge_pos[axis][index - 1]) {
            const uint8_t prev = h.edge_handle[axis][index - 1];
            const uint16_t prev_pos = h.edge_pos[axis][index - 1];
            const bool prev_is_max = (prev_pos & 1) != 0;
            if (update && prev_is_max != is_max && Overlap2D(owner, prev, axis)) {
                if (is_max) {
                    RemovePair(owner, prev);
                } else {
                    AddPair(owner, prev);
                }
            }
            ++EdgeRef(prev, axis, prev_is_max);
            --EdgeRef(owner, axis, is_max);
            h.edge_pos[axis][index] = prev_pos;
            h.edge_handle[axis][index] = prev;
            --index;
        }
        h.edge_pos[axis][index] = pos;
        h.edge_handle[axis][index] = owner;
    }
    RG_HD void SortUp(int axis, int index, bool is_max, bool update) {
        const uint8_t owner = h.edge_handle[axis][index];
        const uint16_t pos = h.edge_pos[axis][index];
        const int len = 2 * h.handle_count + 2;
        while (index + 1 < len && h.edge_handle[axis][index + 1] != 0 && pos >= h.edge_pos[axis][index + 1]) {
            const uint8_t next = h.edge_handle[axis][index + 1];
            const uint16_t next_pos = h.edge_pos[axis][index + 1];
            const bool next_is_max = (next_pos & 1) != 0;
            if (update && next_is_max != is_max && Overlap2D(owner, next, axis)) {
                if (is_max) {
                    AddPair(owner, next);
                } else {
                    RemovePair(owner, next);
                }
            }
            --EdgeRef(next, axis, next_is_max);
            ++EdgeRef(owner, axis, is_max);
            h.edge_pos[axis][index] = next_pos;
            h.edge_handle[axis][index] = next;
            ++index;
        }
        h.edge_pos[axis][index] = pos;
        h.edge_handle[axis][index] = owner;
    }
    // Insert a handle with box [mn, mx] (already quantized edge values), appended at the top then sorted down;
    // pairs are created on the z axis only (Bullet's addHandle).
    RG_HD int AddHandle(const uint16_t qmin[3], const uint16_t qmax[3], int16_t id, bool update_pairs) {
        const int handle = h.free_count > 0 ? h.free_list[--h.free_count] : h.next_fresh++;
        const int limit = 2 * ++h.handle_count;
        h.static_id[handle] = id;
        for (int axis = 0; axis < 3; ++axis) {
            h.max_edge[0][axis] = static_cast<uint8_t>(h.max_edge[0][axis] + 2);
            h.edge_pos[axis][limit + 1] = h.edge_pos[axis][limit - 1];
            h.edge_handle[axis][limit + 1] = h.edge_handle[axis][limit - 1];
            h.edge_pos[axis][limit - 1] = qmin[axis];
            h.edge_handle[axis][limit - 1] = static_cast<uint8_t>(handle);
            h.edge_pos[axis][limit] = qmax[axis];
            h.edge_handle[axis][limit] = static_cast<uint8_t>(handle);
            h.min_edge[handle][axis] = static_cast<uint8_t>(limit - 1);
            h.max_edge[handle][axis] = static_cast<uint8_t>(limit);
        }
        for (int axis = 0; axis < 3; ++axis) {
            SortDown(axis, h.min_edge[handle][axis], false, update_pairs && axis == 2);
            SortDown(axis, h.max_edge[handle][axis], true, update_pairs && axis == 2);
        }
        return handle;
    }
    RG_HD void RemoveBody(int body) {
        const int handle = h.body_handle[body];
        if (handle == 0) return;
        // Drop every pair (and its manifold) that involves the body.
        for (int i = 0; i < bp.pair_count; ++i) {
            const uint32_t key = bp.pairs[i];
            const int kind = KeyKind(key);
            const int a = KeyBodyA(key) + 1, b = KeyBodyB(key) + 1;
            const bool has = kind == 0 ? body == 0 : kind == 1 ? (body == 0 || body == a) : (body == a || (kind == 3 && body == b));
            if (!has) continue;
            const int m = FindKey(bp.manifolds, bp.manifold_count, key);
            if (m >= 0) bp.manifolds[m] = bp.manifolds[--bp.manifold_count];
        }
        for (int i = 0; i < bp.pair_count;) {
            const uint32_t key = bp.pairs[i];
            const int kind = KeyKind(key);
            const int a = KeyBodyA(key) + 1, b = KeyBodyB(key) + 1;
            const bool has = kind == 0 ? body == 0 : kind == 1 ? (body == 0 || body == a) : (body == a || (kind == 3 && body == b));
            if (has) {
                bp.pairs[i] = bp.pairs[--bp.pair_count];
            } else {
                ++i;
            }
        }
        const int limit = 2 * h.handle_count;
        for (int axis = 0; axis < 3; ++axis) h.max_edge[0][axis] = static_cast<uint8_t>(h.max_edge[0][axis] - 2);
        for (int axis = 0; axis < 3; ++axis) {
            for (int is_max = 1; is_max >= 0; --is_max) {
                int index = is_max ? h.max_edge[handle][axis] : h.min_edge[handle][axis];
                h.edge_pos[axis][index] = 0xffff;
                while (h.edge_handle[axis][index + 1] != 0 && h.edge_pos[axis][index] >= h.edge_pos[axis][index + 1]) {
                    const int next = h.edge_handle[axis][index + 1];
                    const bool next_is_max = (h.edge_pos[axis][index + 1] & 1) != 0;
                    --EdgeRef(next, axis, next_is_max);
                    ++EdgeRef(handle, axis, is_max != 0);
                    const uint16_t p = h.edge_pos[axis][index];
                    h.edge_pos[axis][index] = h.edge_pos[axis][index + 1];
                    h.edge_handle[axis][index] = h.edge_handle[axis][index + 1];
                    h.edge_pos[axis][index + 1] = p;
                    h.edge_handle[axis][index + 1] = static_cast<uint8_t>(handle);
                    ++index;
                }
            }
            h.edge_handle[axis][limit - 1] = 0;
            h.edge_pos[axis][limit - 1] = 0xffff;
        }
        h.free_list[h.free_count++] = static_cast<uint8_t>(handle);
        h.handle_count = static_cast<uint8_t>(h.handle_count - 1);
        h.body_handle[body] = 0;
        h.body_aabb_valid = static_cast<uint16_t>(h.body_aabb_valid & ~(1u << body));
        for (int i = 0; i < h.obj_count; ++i) {
            if (h.obj[i] == body) {
                h.obj[i] = h.obj[--h.obj_count];
                break;
            }
        }
    }
};

// Fresh sweep with the arena's static meshes (in mesh order), as Bullet's world is built.
RG_HD void SweepInitStatics(BroadphaseState& bp, const ArenaMeshView& mesh) {
    SweepState& h = bp.sweep;
    h.next_fresh = 1;
    h.free_count = 0;
    h.handle_count = 0;
    h.obj_count = 0;
    for (int b = 0; b < kMaxBodies; ++b) h.body_handle[b] = 0;
    h.body_aabb_valid = 0;
    for (int axis = 0; axis < 3; ++axis) {
        h.edge_pos[axis][0] = 0;
        h.edge_handle[axis][0] = 0;
        h.edge_pos[axis][1] = 0xffff;
        h.edge_handle[axis][1] = 0;
        h.min_edge[0][axis] = 0;
        h.max_edge[0][axis] = 1;
    }
    Sweep s(bp, 0);
    for (int m = 0; m < mesh.mesh_count; ++m) {
        const MeshInfo& info = mesh.meshes[m];
        if (info.mesh_index < 0 || info.tri_count <= 0 || h.handle_count + kMaxBodies >= kSweepMaxHandles + 1) continue;
        uint16_t qmin[3], qmax[3];
        const float mn[3] = {info.aabb_min.x, info.aabb_min.y, info.aabb_min.z};
        const float mx[3] = {info.aabb_max.x, info.aabb_max.y, info.aabb_max.z};
        for (int a = 0; a < 3; ++a) {
            qmin[a] = SweepQuantize(mn[a], 0);
            qmax[a] = SweepQuantize(mx[a], 1);
        }
        s.AddHandle(qmin, qmax, static_cast<int16_t>(info.mesh_index), false);
    }
    h.static_count = h.handle_count;
}

// Bullet's btAlignedObjectArray::quickSortInternal over manifold indices keyed by island (not stable).
RG_HD void QuickSortByIsland(int8_t* order, const int8_t* island, int lo, int hi) {
    constexpr int kMaxRanges = 32;
    int8_t rlo[kMaxRanges], rhi[kMaxRanges];
    int depth = 0;
    rlo[depth] = static_cast<int8_t>(lo);
    rhi[depth] = static_cast<int8_t>(hi);
    ++depth;
    while (depth > 0) {
        --depth;
        int i = rlo[depth], j = rhi[depth];
        const int cur_lo = i, cur_hi = j;
        const int pivot = island[order[(cur_lo + cur_hi) / 2]];
        do {
            while (island[order[i]] < pivot) ++i;
            while (pivot < island[order[j]]) --j;
            if (i <= j) {
                const int8_t t = order[i];
                order[i] = order[j];
                order[j] = t;
                ++i;
                --j;
            }
        } while (i <= j);
        const bool right_first = cur_hi - i < j - cur_lo;
        for (int side = 0; side < 2; ++side) {
            const bool right = (side == 0) != right_first;
            const int s = right ? i : cur_lo;
            const int e = right ? cur_hi : j;
            if (s < e && depth < kMaxRanges) {
                rlo[depth] = static_cast<int8_t>(s);
                rhi[depth] = static_cast<int8_t>(e);
                ++depth;
            }
        }
    }
}

}  // namespace core
}  // namespace rg
