// OBJ reading and writing; see objfile.h.  The reading side follows
// tinyobjloader as the reference vendors it (test/tiny_obj_loader.h): line
// splitting (safeGetline), the decimal reader (tryParseDouble, :814-939, and
// parseReal, :941-950), face indices (parseTriple, :1078-1128, and fixIndex,
// :748-770) and the triangulation of exportGroupsToShape (:1343-1573) with
// pnpoly (:1329-1339).  Its float arithmetic is kept operation for operation:
// which diagonal a polygon is cut along depends on it.
#include "objfile.h"

#include <cfloat>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <utility>

namespace sbobj {

namespace {

inline bool isDigit(char c) { return c >= '0' && c <= '9'; }
inline bool isBlank(char c) { return c == ' ' || c == '\t'; }
inline bool endsLine(char c) { return c == '\r' || c == '\n' || c == '\0'; }

// The decimal reader: an optional sign, integer digits folded as m = 10 m + d,
// fraction digits added as d * 10^-k (k = 1, 2, ...; 10^-k from a table of
// double constants for k < 8, from pow beyond), an optional exponent e applied
// as ldexp(m * 5^e, e).  Reads greedily from s up to `end`; false when no
// number starts at s.  On false *out is untouched.
bool readDecimal(const char *s, const char *end, double *out)
{
    if (s >= end)
        return false;
    static const double kTenth[8] = {1.0, 0.1, 0.01, 0.001, 0.0001, 0.00001, 0.000001, 0.0000001};
    const char *p = s;
    bool negative = false, dotFirst = false;
    if (*p == '+' || *p == '-') {
        negative = *p == '-';
        ++p;
        dotFirst = p != end && *p == '.';
    } else if (*p == '.') {
        dotFirst = true;
    } else if (!isDigit(*p)) {
        return false;
    }
    double m = 0.0;
    int e = 0;
    if (!dotFirst) {
        int digits = 0;
        while (p != end && isDigit(*p)) {
            m *= 10;
            m += (int)(*p - '0');
            ++p;
            ++digits;
        }
        if (digits == 0)
            return false;
    }
    if (p != end) {
        bool toExponent = false;
        if (*p == '.') {
            ++p;
            for (int k = 1; p != end && isDigit(*p); ++k, ++p)
                m += (int)(*p - '0') * (k < 8 ? kTenth[k] : std::pow(10.0, -k));
            toExponent = p != end && (*p == 'e' || *p == 'E');
        } else {
            toExponent = *p == 'e' || *p == 'E';
        }
        if (toExponent) {
            ++p;
            bool expNegative = false;
            if (p != end && (*p == '+' || *p == '-')) {
                expNegative = *p == '-';
                ++p;
            } else if (!isDigit(*p)) {
                return false;
            }
            int digits = 0;
            while (p != end && isDigit(*p)) {
                e = e * 10 + (int)(*p - '0');
                ++p;
                ++digits;
            }
            if (expNegative)
                e = -e;
            if (digits == 0)
                return false;
        }
    }
    const double magnitude = e ? std::ldexp(m * std::pow(5.0, e), e) : m;
    *out = (negative ? -1 : 1) * magnitude;
    return true;
}

// One whitespace-separated coordinate; `fallback` when none can be read.
float readCoordinate(const char *&p, double fallback)
{
    p += std::strspn(p, " \t");
    const char *end = p + std::strcspn(p, " \t\r");
    double value = fallback;
    readDecimal(p, end, &value);
    p = end;
    return (float)value;
}

// A face index: 1-based, or negative counting back from the `count` items
// read so far; zero (also what a non-number reads as) is an error.
bool faceIndex(const char *p, int count, int *out)
{
    const int i = std::atoi(p);
    if (i == 0)
        return false;
    *out = i > 0 ? i - 1 : count + i;
    return true;
}

// "v", "v/t", "v//n" or "v/t/n"; only the vertex index is kept, but a bad
// texture or normal index fails the file as the vertex index does.
bool readCorner(const char *&p, int nv, int nvt, int nvn, int *vertex)
{
    if (!faceIndex(p, nv, vertex))
        return false;
    p += std::strcspn(p, "/ \t\r");
    if (*p != '/')
        return true;
    ++p;
    int unused = 0;
    if (*p == '/') {
        ++p;
        if (!faceIndex(p, nvn, &unused))
            return false;
        p += std::strcspn(p, "/ \t\r");
        return true;
    }
    if (!faceIndex(p, nvt, &unused))
        return false;
    p += std::strcspn(p, "/ \t\r");
    if (*p != '/')
        return true;
    ++p;
    if (!faceIndex(p, nvn, &unused))
        return false;
    p += std::strcspn(p, "/ \t\r");
    return true;
}

// W. R. Franklin's crossing test: is (tx, ty) inside the polygon (x, y)?
bool insidePolygon(int n, const float *x, const float *y, float tx, float ty)
{
    bool in = false;
    for (int i = 0, j = n - 1; i < n; j = i++)
        if (((y[i] > ty) != (y[j] > ty)) && (tx < (x[j] - x[i]) * (ty - y[i]) / (y[j] - y[i]) + x[i]))
            in = !in;
    return in;
}

// Appends the triangles of one face (indices into the float coordinates v).
void cutFace(const std::vector<int> &face, const std::vector<float> &v, std::vector<int> &out)
{
    const size_t n = face.size();
    if (n < 3)
        return;
    auto valid = [&](int vi, size_t axis) { return vi >= 0 && 3 * (size_t)vi + axis < v.size(); };
    auto at = [&](int vi, size_t axis) { return v[3 * (size_t)vi + axis]; };
    // the plane to work in: drop the axis the first proper corner's normal
    // leans on most (by default x, i.e. work in y and z)
    size_t ax0 = 1, ax1 = 2;
    for (size_t k = 0; k < n; ++k) {
        const int a = face[k], b = face[(k + 1) % n], c = face[(k + 2) % n];
        if (!valid(a, 2) || !valid(b, 2) || !valid(c, 2))
            continue;
        const float e0x = at(b, 0) - at(a, 0), e0y = at(b, 1) - at(a, 1), e0z = at(b, 2) - at(a, 2);
        const float e1x = at(c, 0) - at(b, 0), e1y = at(c, 1) - at(b, 1), e1z = at(c, 2) - at(b, 2);
        const float nx = std::fabs(e0y * e1z - e0z * e1y);
        const float ny = std::fabs(e0z * e1x - e0x * e1z);
        const float nz = std::fabs(e0x * e1y - e0y * e1x);
        if (nx > FLT_EPSILON || ny > FLT_EPSILON || nz > FLT_EPSILON) {
            if (!(nx > ny && nx > nz)) {
                ax0 = 0;
                if (nz > nx && nz > ny)
                    ax1 = 1;
            }
            break;
        }
    }
    // twice-halved signed area in that plane (its sign is the winding)
    float area = 0;
    for (size_t k = 0; k < n; ++k) {
        const int a = face[k], b = face[(k + 1) % n];
        if (!valid(a, ax0) || !valid(a, ax1) || !valid(b, ax0) || !valid(b, ax1))
            continue;
        area += (at(a, ax0) * at(b, ax1) - at(a, ax1) * at(b, ax0)) * 0.5f;
    }
    // ear clipping: the corner at `guess` + 1 is cut off when it turns the
    // polygon's way and no other remaining vertex lies inside it; a full round
    // without a cut gives up on what is left
    std::vector<int> rest = face;
    size_t guess = 0, tries = n, lastSize = n;
    while (rest.size() > 3 && tries > 0) {
        const size_t m = rest.size();
        if (guess >= m)
            guess -= m;
        if (lastSize != m) {
            lastSize = m;
            tries = m;
        } else {
            --tries;
        }
        int corner[3];
        float px[3], py[3];
        for (size_t k = 0; k < 3; ++k) {
            corner[k] = rest[(guess + k) % m];
            const bool ok = valid(corner[k], ax0) && valid(corner[k], ax1);
            px[k] = ok ? at(corner[k], ax0) : 0.0f;
            py[k] = ok ? at(corner[k], ax1) : 0.0f;
        }
        const float turn = (px[1] - px[0]) * (py[2] - py[1]) - (py[1] - py[0]) * (px[2] - px[1]);
        if (turn * area < 0.0f) {
            ++guess;
            continue;
        }
        bool covers = false;
        for (size_t k = 3; k < m && !covers; ++k) {
            const int o = rest[(guess + k) % m];
            if (!valid(o, ax0) || !valid(o, ax1))
                continue;
            covers = insidePolygon(3, px, py, at(o, ax0), at(o, ax1));
        }
        if (covers) {
            ++guess;
            continue;
        }
        out.insert(out.end(), corner, corner + 3);
        rest.erase(rest.begin() + (long)((guess + 1) % m));
    }
    if (rest.size() == 3)
        out.insert(out.end(), rest.begin(), rest.end());
}

}  // namespace

bool parseObj(const std::string &text, std::vector<Vector3> &vertices, std::vector<std::vector<size_t>> &triangles,
              std::string *error)
{
    std::vector<float> v;
    int nvt = 0, nvn = 0;
    std::vector<int> tri;
    // faces are cut when their group ends (a "g" or "o" line, or the end of
    // the file), with the vertices read by then
    std::vector<std::vector<int>> pending;
    auto flush = [&]() {
        for (const auto &f : pending)
            cutFace(f, v, tri);
        pending.clear();
    };
    size_t lineNo = 0;
    for (size_t pos = 0; pos < text.size();) {
        // a line ends at "\n", "\r\n" or a lone "\r"
        size_t stop = pos;
        while (stop < text.size() && text[stop] != '\n' && text[stop] != '\r')
            ++stop;
        const std::string line = text.substr(pos, stop - pos);
        pos = stop;
        if (pos < text.size()) {
            if (text[pos] == '\r' && pos + 1 < text.size() && text[pos + 1] == '\n')
                ++pos;
            ++pos;
        }
        ++lineNo;
        const char *p = line.c_str();
        p += std::strspn(p, " \t");
        if (p[0] == 'v' && isBlank(p[1])) {
            p += 2;
            for (int k = 0; k < 3; ++k)
                v.push_back(readCoordinate(p, 0.0));
        } else if (p[0] == 'v' && p[1] == 't' && isBlank(p[2])) {
            ++nvt;
        } else if (p[0] == 'v' && p[1] == 'n' && isBlank(p[2])) {
            ++nvn;
        } else if ((p[0] == 'g' || p[0] == 'o') && isBlank(p[1])) {
            flush();
        } else if (p[0] == 'f' && isBlank(p[1])) {
            p += 2;
            p += std::strspn(p, " \t");
            std::vector<int> face;
            while (!endsLine(*p)) {
                int vi = 0;
                if (!readCorner(p, (int)(v.size() / 3), nvt, nvn, &vi)) {
                    if (error)
                        *error = "bad face index on line " + std::to_string(lineNo);
                    return false;
                }
                face.push_back(vi);
                p += std::strspn(p, " \t\r");
            }
            pending.push_back(std::move(face));
        }
    }
    flush();
    const size_t nv = v.size() / 3;
    for (int i : tri)
        if (i < 0 || (size_t)i >= nv) {
            if (error)
                *error = "face index " + std::to_string(i + 1) + " outside the " + std::to_string(nv) + " vertices";
            return false;
        }
    vertices.resize(nv);
    for (size_t i = 0; i < nv; ++i)
        vertices[i] = Vector3(v[3 * i], v[3 * i + 1], v[3 * i + 2]);
    triangles.clear();
    triangles.reserve(tri.size() / 3);
    for (size_t i = 0; i + 2 < tri.size(); i += 3)
        triangles.push_back({(size_t)tri[i], (size_t)tri[i + 1], (size_t)tri[i + 2]});
    return true;
}

bool loadObj(const std::string &path, std::vector<Vector3> &vertices, std::vector<std::vector<size_t>> &triangles,
             std::string *error)
{
    std::ifstream in(path, std::ios::binary);
    if (!in) {
        if (error)
            *error = "cannot open " + path;
        return false;
    }
    std::ostringstream text;
    text << in.rdbuf();
    return parseObj(text.str(), vertices, triangles, error);
}

bool exportObject(const std::string &path, const std::vector<Vector3> &vertices,
                  const std::vector<std::vector<size_t>> &faces)
{
    FILE *fp = std::fopen(path.c_str(), "wb");
    if (!fp)
        return false;
    for (const Vector3 &p : vertices)
        std::fprintf(fp, "v %f %f %f\n", p.x(), p.y(), p.z());
    for (const auto &f : faces) {
        std::fputs(f.size() == 2 ? "l" : "f", fp);
        for (size_t i : f)
            std::fprintf(fp, " %zu", i + 1);
        std::fputc('\n', fp);
    }
    return std::fclose(fp) == 0;
}

}  // namespace sbobj
