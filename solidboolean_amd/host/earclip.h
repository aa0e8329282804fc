// Ear-clipping triangulation of a polygon with holes.
//
// Follows the published algorithm of mapbox/earcut (ISC; the reference vendors
// its C++ port at thirdparty/earcut.hpp, "parity with Earcut v2.2.2"): a
// circular vertex list in a fixed winding, holes bridged to the outer ring
// from left to right with Eberly's visible-vertex search, ear slicing with a
// z-order index above 80 input points, and the three fallback passes (filter
// collinear points, cure local self-intersections, split along a valid
// diagonal).  As in the vendored copy, a signed triangle area within
// DBL_EPSILON of zero counts as zero (earcut.hpp:694-699), which decides
// which collinear points are filtered and therefore how many triangles come
// out.
#pragma once
#include <array>
#include <cstddef>
#include <vector>

namespace sbhost {

using Ring = std::vector<std::array<double, 2>>;

// rings[0] is the outer boundary, the rest are holes.  Returns vertex indices
// (into the concatenation of all rings), three per triangle.
std::vector<size_t> earclip(const std::vector<Ring> &rings);

}  // namespace sbhost
