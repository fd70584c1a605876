// mesh.h - host preparation of a triangle mesh for the GPU distance kernel (boundary.hip: k_mesh_field0).
#pragma once

#include <vector>

#include "../../include/vfd.h"

namespace vfd {

struct MeshPrep {
    std::vector<float> tri;      // 16 floats per triangle, the MeshDev::tri records (Morton order of the centroids)
    std::vector<float> normals;  // 21 floats per triangle, the MeshDev::normals records (same order)
    std::vector<float> spheres;  // 4 floats per cluster of kMeshCluster stored triangles: centre, radius
    uint32_t firstPos = 0;       // stored position of the mesh's triangle 0
    float lo[3], hi[3];          // BoundingBox(vertices) of the transformed vertices (always holds the origin)
};

// Transforms the vertices (transform = column-major 4x4, NULL = identity), checks the triangle indices and fills
// `out`. Returns a VFD status; the message goes to vfd_last_error.
int prepare_mesh(const vfd_mesh& mesh, const float* transform, MeshPrep& out);

}  // namespace vfd
