// Wavefront OBJ input and output the way the reference's test driver does
// them (test/main.cpp:9-28 exportObject, test/main.cpp:31-69 loadObj), the
// reading side being tinyobjloader's LoadObj with triangulation on
// (test/tiny_obj_loader.h, vendored by the reference), restated here.
#pragma once

#include <cstddef>
#include <string>
#include <vector>

#include "vector3.h"

namespace sbobj {

// Every vertex ("v" lines) and every face ("f" lines) of the file, faces cut
// into triangles, in file order.  A coordinate is read by tinyobjloader's
// decimal procedure into a double and narrowed to float, its coordinate type;
// a polygon is cut by its ear clipping, which works in float on the two axes
// of the polygon's first non-degenerate corner, when its group ends (a "g" or
// "o" line, or the end of the file) with the vertices read by then.  Other
// statements (vt, vn, s, usemtl, mtllib, ...) are skipped: material libraries
// are not read, so "usemtl" never ends a group here.  Groups do not reorder
// faces.
// Returns false with a message when the file cannot be read, when a face index
// is zero or not a number (tinyobjloader's own failures), or when an index
// points outside the vertices (which tinyobjloader only warns about, and
// which the reference driver would then read out of bounds).
bool loadObj(const std::string &path, std::vector<Vector3> &vertices, std::vector<std::vector<size_t>> &triangles,
             std::string *error = nullptr);

// Same, from the text of a file.
bool parseObj(const std::string &text, std::vector<Vector3> &vertices, std::vector<std::vector<size_t>> &triangles,
              std::string *error = nullptr);

// "v %f %f %f" per vertex, then per face "f" (or "l" for a two-index face)
// with 1-based indices, as test/main.cpp:9-28 writes them.  False when the
// file cannot be opened.
bool exportObject(const std::string &path, const std::vector<Vector3> &vertices,
                  const std::vector<std::vector<size_t>> &faces);

}  // namespace sbobj
