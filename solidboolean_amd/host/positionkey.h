// Integer grid key of a position: each coordinate times 1e5, truncated toward
// zero (reference: src/positionkey.cpp:25-37), ordered lexicographically
// (:39-54).  Used for welding intersection points on the host.
#pragma once
#include "vector3.h"

class PositionKey {
public:
    explicit PositionKey(const Vector3 &v) : PositionKey(v.x(), v.y(), v.z()) {}
    PositionKey(double x, double y, double z)
        : m_x((long)(x * (double)kFactor)), m_y((long)(y * (double)kFactor)), m_z((long)(z * (double)kFactor))
    {
    }
    bool operator<(const PositionKey &o) const
    {
        if (m_x != o.m_x)
            return m_x < o.m_x;
        if (m_y != o.m_y)
            return m_y < o.m_y;
        return m_z < o.m_z;
    }
    bool operator==(const PositionKey &o) const { return m_x == o.m_x && m_y == o.m_y && m_z == o.m_z; }

private:
    static constexpr long kFactor = 100000;
    long m_x, m_y, m_z;
};
