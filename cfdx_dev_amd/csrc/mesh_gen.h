// Synthetic unstructured hex meshes (host side). See mesh_gen.cpp.
#pragma once

#include <cstddef>
#include <cstdint>
#include <vector>

namespace cfdxhip {

enum MeshKind : uint32_t { MESH_BOX = 0, MESH_EXTRUDED_2D = 1, MESH_CHANNEL_STEP = 2 };
enum PatchKind : uint8_t { PATCH_WALL = 0, PATCH_EMPTY = 1, PATCH_SYMMETRY = 2 };

struct MeshSpec {
    uint32_t kind = MESH_BOX;
    uint32_t nx = 1, ny = 1, nz = 1;
    double lx = 1.0, ly = 1.0, lz = 1.0;
    double jitter = 0.0;          // fraction of the local spacing, interior vertices
    uint32_t wall_layers = 0;     // clustered layers at each y wall (box)
    double first_aspect = 0.0;    // first-cell dx/dy at a y wall (0: uniform)
    double growth = 0.0;          // wall-normal growth ratio (channel)
    double step_height_fraction = 0.0;
    double step_length_fraction = 0.0;
    uint32_t permute = 0;
    uint64_t seed = 0;
};

struct HostMesh {
    std::vector<double> px, py, pz;
    std::vector<uint32_t> foff, fvtx, owner;
    std::vector<int32_t> neigh;
    std::vector<uint32_t> coff, cfaces, patch_of_face;
    std::vector<uint8_t> patch_kind;
    std::vector<uint32_t> cell_ijk;
    uint64_t n_cells = 0, n_internal_faces = 0, n_patches = 0;
};

HostMesh generate_mesh(const MeshSpec& spec);

}  // namespace cfdxhip
