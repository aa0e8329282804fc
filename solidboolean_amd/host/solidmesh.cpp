#include "solidmesh.h"

#include <cstdint>
#include <stdexcept>

#include "referencetree.h"
#include "sbhost_log.h"

SolidMesh::SolidMesh() = default;

SolidMesh::~SolidMesh() { release(); }

void SolidMesh::release()
{
    if (m_refTreeBuild.valid())
        m_refTreeBuild.wait();
    m_refTreeBuild = {};
    m_refTree.reset();
    delete m_triangleNormals;
    m_triangleNormals = nullptr;
    delete m_boxes;
    m_boxes = nullptr;
    delete m_tree;
    m_tree = nullptr;
    if (m_gpu) {
        sbhost::GpuLock lock(sbhost::gpuMutex());
        sbg_mesh_destroy(m_gpu);
    }
    m_gpu = nullptr;
}

void SolidMesh::prepare()
{
    if (!m_triangles)
        return;  // as src/solidmesh.cpp:44-45
    if (!m_vertices)
        throw std::invalid_argument("SolidMesh::prepare: vertices not set");
    release();
    const size_t nv = m_vertices->size();
    const size_t nt = m_triangles->size();
    std::vector<double> xyz(3 * nv);
    for (size_t i = 0; i < nv; ++i) {
        xyz[3 * i] = (*m_vertices)[i].x();
        xyz[3 * i + 1] = (*m_vertices)[i].y();
        xyz[3 * i + 2] = (*m_vertices)[i].z();
    }
    std::vector<uint32_t> tri(3 * nt);
    for (size_t i = 0; i < nt; ++i) {
        const auto &t = (*m_triangles)[i];
        if (t.size() < 3)
            throw std::invalid_argument("SolidMesh::prepare: triangle with fewer than 3 indices");
        for (int k = 0; k < 3; ++k) {
            if (t[k] > 0xffffffffu)
                throw std::invalid_argument("SolidMesh::prepare: vertex index beyond 32 bits");
            tri[3 * i + k] = (uint32_t)t[k];
        }
    }
    {
        sbhost::GpuLock lock(sbhost::gpuMutex());
        sbhost::check(sbg_mesh_create(sbhost::gpu(), xyz.data(), nv, tri.data(), nt, &m_gpu), "sbg_mesh_create");
        sbhost::check(sbg_mesh_prepare(m_gpu), "sbg_mesh_prepare");
    }
    // the reference's tree, for the pair order, on a host thread while the
    // GPU prepares and searches (the arrays are the caller's and stay alive)
    const std::vector<Vector3> *v = m_vertices;
    const std::vector<std::vector<size_t>> *t = m_triangles;
    m_refTreeBuild = std::async(std::launch::async, [v, t]() {
        return std::unique_ptr<sbhost::ReferenceTree>(new sbhost::ReferenceTree(*v, *t));
    });
}

const sbhost::ReferenceTree *SolidMesh::referenceTree() const
{
    if (!m_refTree && m_refTreeBuild.valid())
        m_refTree = m_refTreeBuild.get();
    return m_refTree.get();
}

const std::vector<Vector3> *SolidMesh::triangleNormals() const
{
    if (!m_gpu)
        return nullptr;
    if (!m_triangleNormals) {
        const size_t nt = m_triangles->size();
        std::vector<double> n(3 * nt);
        {
            sbhost::GpuLock lock(sbhost::gpuMutex());
            sbhost::check(sbg_mesh_normals(m_gpu, n.data()), "sbg_mesh_normals");
        }
        auto *v = new std::vector<Vector3>(nt);
        for (size_t i = 0; i < nt; ++i)
            (*v)[i] = Vector3(n[3 * i], n[3 * i + 1], n[3 * i + 2]);
        m_triangleNormals = v;
    }
    return m_triangleNormals;
}

const std::vector<AxisAlignedBoudingBox> *SolidMesh::triangleAxisAlignedBoundingBoxes() const
{
    if (!m_gpu)
        return nullptr;
    if (!m_boxes) {
        // boxes built from the triangle's vertices exactly as prepare() does on
        // the reference (src/solidmesh.cpp:57-62), so centers match too
        auto *b = new std::vector<AxisAlignedBoudingBox>(m_triangles->size());
        for (size_t i = 0; i < b->size(); ++i) {
            for (size_t k = 0; k < 3; ++k)
                (*b)[i].update((*m_vertices)[(*m_triangles)[i][k]]);
            (*b)[i].updateCenter();
        }
        m_boxes = b;
    }
    return m_boxes;
}

const AxisAlignedBoudingBoxTree *SolidMesh::axisAlignedBoundingBoxTree() const
{
    if (!m_gpu)
        return nullptr;
    if (!m_tree)
        m_tree = new AxisAlignedBoudingBoxTree(this);
    return m_tree;
}
