// SolidMesh: a closed triangle mesh prepared for the boolean, with the
// reference's interface (src/solidmesh.h:29-74).  The mesh does not own its
// vertex and triangle arrays; the caller keeps them alive.  prepare() builds
// the per-triangle normals and boxes and the bounding-volume hierarchy on the
// GPU (libsbgpu: sbg_mesh_create + sbg_mesh_prepare); the host-side accessors
// copy results back on first use.
#pragma once
#include <cstddef>
#include <utility>
#include <vector>

#include "axisalignedboundingbox.h"
#include "sbgpu.h"
#include "vector3.h"

class SolidMesh;

// The mesh's bounding-volume hierarchy.  In the reference this is a host
// pointer tree (src/axisalignedboundingboxtree.h); here it is a linear BVH in
// device memory, so instead of walking nodes a caller asks it for the
// candidate pairs against another mesh's tree.
class AxisAlignedBoudingBoxTree {
public:
    explicit AxisAlignedBoudingBoxTree(const SolidMesh *mesh) : m_mesh(mesh) {}
    // every (a in this mesh, b in other) whose triangle boxes closed-overlap,
    // sorted by (a, b) (the set of axisalignedboundingboxtree.h:54-95)
    void test(const AxisAlignedBoudingBoxTree *other, std::vector<std::pair<size_t, size_t>> *pairs) const;
    const SolidMesh *mesh() const { return m_mesh; }

private:
    const SolidMesh *m_mesh;
};

class SolidMesh {
public:
    SolidMesh() = default;
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

private:
    const std::vector<Vector3> *m_vertices = nullptr;
    const std::vector<std::vector<size_t>> *m_triangles = nullptr;
    sbg_mesh m_gpu = nullptr;
    mutable std::vector<Vector3> *m_triangleNormals = nullptr;
    mutable std::vector<AxisAlignedBoudingBox> *m_boxes = nullptr;
    mutable AxisAlignedBoudingBoxTree *m_tree = nullptr;
    void release();
};
