// Host-side arena mesh preprocessing: triangle records with Bullet's internal-edge data, per-mesh quantization and
// the Bullet-style quantized AABB tree.  Runs once per arena geometry at BatchSim construction; the result is
// uploaded to HBM and shared read-only by every arena that uses the geometry.
#pragma once

#include <cstdint>
#include <vector>

#include "core/rg_mesh.hpp"

namespace rg {

struct MeshSource {
    std::vector<Vec3> vertices_bt;
    std::vector<int32_t> indices;
};

struct ArenaMeshHost {
    std::vector<core::MeshTriangle> tris;
    std::vector<core::MeshBvhNode> nodes;
    std::vector<core::MeshInfo> meshes;
};

// Groups of at most 16 meshes, triangles in source order; throws std::invalid_argument on malformed input.
ArenaMeshHost BuildArenaMesh(const std::vector<MeshSource>& sources);

// Merge a layout's meshes (vertices in uu) into at most 16 groups the way the reference's custom layouts are merged
// (src/rocketgpu.cpp:752-759), converting to bt.
std::vector<MeshSource> GroupLayoutMeshes(const std::vector<std::vector<Vec3>>& vertices_uu,
                                          const std::vector<std::vector<int32_t>>& indices);

}  // namespace rg
