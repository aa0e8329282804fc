// FP64 3-vector with the public surface of the reference's Vector3
// (src/vector3.h:43-311) that callers of SolidMesh / SolidBoolean and the host
// retriangulator use.  Every routine rounds as the reference's does: the host
// library is compiled with -ffp-contract=off and the expressions keep the
// reference's operand order (normal: :155-176, normalized: :100-109,
// project: :227-252, intersectSegmentAndPlane: :264-280).
#pragma once
#include <cmath>
#include <cstddef>
#include <limits>
#include <ostream>
#include <string>
#include <vector>

#include "vector2.h"

class Vector3 {
public:
    Vector3() : m_data{0.0, 0.0, 0.0} {}
    Vector3(double x, double y, double z) : m_data{x, y, z} {}

    double &operator[](size_t i) { return m_data[i]; }
    const double &operator[](size_t i) const { return m_data[i]; }
    const double &x() const { return m_data[0]; }
    const double &y() const { return m_data[1]; }
    const double &z() const { return m_data[2]; }
    void setX(double v) { m_data[0] = v; }
    void setY(double v) { m_data[1] = v; }
    void setZ(double v) { m_data[2] = v; }
    const double *constData() const { return m_data; }
    void setData(double x, double y, double z)
    {
        m_data[0] = x;
        m_data[1] = y;
        m_data[2] = z;
    }

    double lengthSquared() const { return x() * x() + y() * y() + z() * z(); }
    double length() const { return std::sqrt(x() * x() + y() * y() + z() * z()); }

    Vector3 normalized() const
    {
        const double len = std::sqrt(x() * x() + y() * y() + z() * z());
        if (sbIsZero(len))
            return Vector3();
        return Vector3(x() / len, y() / len, z() / len);
    }
    void normalize()
    {
        const double len = std::sqrt(x() * x() + y() * y() + z() * z());
        if (sbIsZero(len))
            return;
        m_data[0] = x() / len;
        m_data[1] = y() / len;
        m_data[2] = z() / len;
    }

    static Vector3 crossProduct(const Vector3 &a, const Vector3 &b)
    {
        return Vector3(a.y() * b.z() - a.z() * b.y(), a.z() * b.x() - a.x() * b.z(), a.x() * b.y() - a.y() * b.x());
    }
    static double dotProduct(const Vector3 &a, const Vector3 &b) { return a.x() * b.x() + a.y() * b.y() + a.z() * b.z(); }

    // unit normal of triangle (a, b, c); the zero vector when degenerate
    static Vector3 normal(const Vector3 &a, const Vector3 &b, const Vector3 &c)
    {
        const double baX = b.x() - a.x(), baY = b.y() - a.y(), baZ = b.z() - a.z();
        const double caX = c.x() - a.x(), caY = c.y() - a.y(), caZ = c.z() - a.z();
        const double cx = baY * caZ - baZ * caY;
        const double cy = baZ * caX - baX * caZ;
        const double cz = baX * caY - baY * caX;
        const double len = std::sqrt(cx * cx + cy * cy + cz * cz);
        if (sbIsZero(len))
            return Vector3();
        return Vector3(cx / len, cy / len, cz / len);
    }

    bool isZero() const { return sbIsZero(m_data[0]) && sbIsZero(m_data[1]) && sbIsZero(m_data[2]); }

    Vector3 &operator+=(const Vector3 &o)
    {
        m_data[0] += o.x();
        m_data[1] += o.y();
        m_data[2] += o.z();
        return *this;
    }
    Vector3 &operator*=(double s)
    {
        m_data[0] *= s;
        m_data[1] *= s;
        m_data[2] *= s;
        return *this;
    }
    Vector3 &operator/=(double s)
    {
        m_data[0] /= s;
        m_data[1] /= s;
        m_data[2] /= s;
        return *this;
    }

    // 2-D coordinates of points in the plane frame (axis, normal x axis) at origin
    static void project(const std::vector<Vector3> &in, std::vector<Vector2> *out, const Vector3 &normal,
                        const Vector3 &axis, const Vector3 &origin = Vector3());

    // segment p0 -> p1 against the plane through `onPlane` with normal n
    static bool intersectSegmentAndPlane(const Vector3 &p0, const Vector3 &p1, const Vector3 &onPlane,
                                         const Vector3 &n, Vector3 *hit);

    static bool sbIsZero(double v) { return std::abs(v) <= std::numeric_limits<double>::epsilon(); }

private:
    double m_data[3];
};

inline Vector3 operator*(double s, const Vector3 &v) { return Vector3(s * v.x(), s * v.y(), s * v.z()); }
inline Vector3 operator*(const Vector3 &v, double s) { return Vector3(s * v.x(), s * v.y(), s * v.z()); }
inline Vector3 operator+(const Vector3 &a, const Vector3 &b) { return Vector3(a.x() + b.x(), a.y() + b.y(), a.z() + b.z()); }
inline Vector3 operator-(const Vector3 &a, const Vector3 &b) { return Vector3(a.x() - b.x(), a.y() - b.y(), a.z() - b.z()); }
inline Vector3 operator-(const Vector3 &v) { return Vector3(-v.x(), -v.y(), -v.z()); }
inline Vector3 operator/(const Vector3 &v, double s) { return Vector3(v.x() / s, v.y() / s, v.z() / s); }

inline void Vector3::project(const std::vector<Vector3> &in, std::vector<Vector2> *out, const Vector3 &normal,
                             const Vector3 &axis, const Vector3 &origin)
{
    const Vector3 perp = crossProduct(normal, axis);
    for (const Vector3 &p : in) {
        const Vector3 d = p - origin;
        out->push_back(Vector2(dotProduct(d, axis), dotProduct(d, perp)));
    }
}

inline bool Vector3::intersectSegmentAndPlane(const Vector3 &p0, const Vector3 &p1, const Vector3 &onPlane,
                                              const Vector3 &n, Vector3 *hit)
{
    const Vector3 u = p1 - p0;
    const Vector3 w = p0 - onPlane;
    const double d = dotProduct(n, u);
    const double num = dotProduct(-n, w);
    if (std::abs(d) <= std::numeric_limits<double>::epsilon())
        return false;
    const double s = num / d;
    if (s < 0 || s > 1 || std::isnan(s) || std::isinf(s))
        return false;
    if (hit)
        *hit = p0 + s * u;
    return true;
}

inline std::ostream &operator<<(std::ostream &os, const Vector3 &v)
{
    os << v.x() << ',' << v.y() << ',' << v.z();
    return os;
}
