// Host-side FP64 axis-aligned box with the reference's public surface
// (src/axisalignedboundingbox.h:28-121): update from points, center as
// sum / (float)count, closed-interval overlap.  The boxes the GPU path uses are
// built on the device; this type is what SolidMesh::
// triangleAxisAlignedBoundingBoxes() hands back to host callers.
#pragma once
#include <cstddef>
#include <limits>

#include "vector3.h"

class AxisAlignedBoudingBox {
public:
    void update(const Vector3 &v)
    {
        for (size_t i = 0; i < 3; ++i) {
            if (v[i] > m_max[i])
                m_max[i] = v[i];
            if (v[i] < m_min[i])
                m_min[i] = v[i];
            m_sum[i] += v[i];
        }
        ++m_num;
    }
    const Vector3 &center() const { return m_center; }
    void updateCenter()
    {
        if (m_num == 0)
            return;
        m_center = m_sum /= (float)m_num;
    }
    const Vector3 &lowerBound() const { return m_min; }
    const Vector3 &upperBound() const { return m_max; }
    Vector3 &lowerBound() { return m_min; }
    Vector3 &upperBound() { return m_max; }
    bool intersectWith(const AxisAlignedBoudingBox &o) const
    {
        for (size_t i = 0; i < 3; ++i)
            if (!(m_min[i] <= o.m_max[i] && m_max[i] >= o.m_min[i]))
                return false;
        return true;
    }

private:
    Vector3 m_min{std::numeric_limits<double>::max(), std::numeric_limits<double>::max(),
                  std::numeric_limits<double>::max()};
    Vector3 m_max{std::numeric_limits<double>::lowest(), std::numeric_limits<double>::lowest(),
                  std::numeric_limits<double>::lowest()};
    Vector3 m_sum;
    size_t m_num = 0;
    Vector3 m_center;
};
