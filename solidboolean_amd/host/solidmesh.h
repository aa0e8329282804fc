// SolidMesh: a closed triangle mesh prepared for the boolean, with the
// reference's interface (src/solidmesh.h:29-74).  The mesh does not own its
// vertex and triangle arrays; the caller keeps them alive.  prepare() builds
// the per-triangle normals and boxes and the bounding-volume hierarchy on the
// GPU (libsbgpu: sbg_mesh_create + sbg_mesh_prepare); the host-side accessors
// copy results back on first use.
#pragma once
#include <cstddef>
#include <future>
#include <memory>
#include <utility>
#include <vector>

#include "axisalignedboundingbox.h"
#include "axisalignedboundingboxtree.h"
#include "sbgpu.h"
#include "vector3.h"

namespace sbhost {
class ReferenceTree;
}

class SolidMesh {
public:
    SolidMesh();
    ~SolidMesh();
    SolidMesh(const SolidMesh &) = delete;
    SolidMesh &operator=(const SolidMesh &) = delete;

    void setVertices(const std::vector<Vector3> *vertices) { m_vertices = vertices; }
    void setTriangles(const std::vector<std::vector<size_t>> *triangles) { m_triangles = triangles; }
    const std::vector<Vector3> *vertices() const { return m_vertices; }
    const std::vector<std::vector<size_t>> *triangles() const { return m_triangles; }
    const std::vector<Vector3> *triangleNormals() const;
    const AxisAlignedBoudingBoxTree *axisAlignedBoundingBoxTree() const;
    const std::vector<AxisAlignedBoudingBox> *triangleAxisAlignedBoundingBoxes() const;

    // Does nothing without triangles (as the reference); preparing again
    // rebuilds from the current arrays.
    void prepare();

    // the device mesh (null before prepare())
    sbg_mesh gpuMesh() const { return m_gpu; }
    // The order tables of the reference's own tree over this mesh (the order
    // its pair search reports pairs in).  prepare() starts building them on a
    // host thread beside the GPU work; this waits for that build.  Null
    // before prepare().
    const sbhost::ReferenceTree *referenceTree() const;

private:
    const std::vector<Vector3> *m_vertices = nullptr;
    const std::vector<std::vector<size_t>> *m_triangles = nullptr;
    sbg_mesh m_gpu = nullptr;
    mutable std::vector<Vector3> *m_triangleNormals = nullptr;
    mutable std::vector<AxisAlignedBoudingBox> *m_boxes = nullptr;
    mutable AxisAlignedBoudingBoxTree *m_tree = nullptr;
    mutable std::future<std::unique_ptr<sbhost::ReferenceTree>> m_refTreeBuild;
    mutable std::unique_ptr<sbhost::ReferenceTree> m_refTree;
    void release();
};
