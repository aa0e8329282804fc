// FP64 2-vector with the predicates the host retriangulator needs, rounding as
// the reference's Vector2 does (src/vector2.h: isOnLine :212-215, PNPOLY
// isInPolygon :201-210).
#pragma once
#include <cmath>
#include <cstddef>
#include <limits>
#include <vector>

class Vector2 {
public:
    Vector2() : m_data{0.0, 0.0} {}
    Vector2(double x, double y) : m_data{x, y} {}
    double &operator[](size_t i) { return m_data[i]; }
    const double &operator[](size_t i) const { return m_data[i]; }
    const double &x() const { return m_data[0]; }
    const double &y() const { return m_data[1]; }
    double lengthSquared() const { return x() * x() + y() * y(); }

    // |cross(b - a, p - a)| within DBL_EPSILON (absolute)
    bool isOnLine(const Vector2 &a, const Vector2 &b) const
    {
        return std::abs(std::abs((y() - a.y()) * (b.x() - a.x()) - (x() - a.x()) * (b.y() - a.y()))) <=
               std::numeric_limits<double>::epsilon();
    }

    // even-odd crossing test against the polygon polygonVertices[polygonIndices]
    bool isInPolygon(const std::vector<Vector2> &v, const std::vector<size_t> &idx) const
    {
        bool inside = false;
        const int n = (int)idx.size();
        for (int i = 0, j = n - 1; i < n; j = i++) {
            const Vector2 &pi = v[idx[i]];
            const Vector2 &pj = v[idx[j]];
            if (((pi.y() > y()) != (pj.y() > y())) &&
                (x() < (pj.x() - pi.x()) * (y() - pi.y()) / (pj.y() - pi.y()) + pi.x()))
                inside = !inside;
        }
        return inside;
    }

private:
    double m_data[2];
};

inline Vector2 operator-(const Vector2 &a, const Vector2 &b) { return Vector2(a.x() - b.x(), a.y() - b.y()); }
inline Vector2 operator+(const Vector2 &a, const Vector2 &b) { return Vector2(a.x() + b.x(), a.y() + b.y()); }
