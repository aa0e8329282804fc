/*
 * The steps of this triangulator follow mapbox/earcut and its C++ port, which
 * carry these notices:
 *
 * ISC License
 *
 * Copyright (c) 2015, Mapbox
 *
 * Permission to use, copy, modify, and/or distribute this software for any purpose
 * with or without fee is hereby granted, provided that the above copyright notice
 * and this permission notice appear in all copies.
 *
 * THE SOFTWARE IS PROVIDED "AS IS" AND THE AUTHOR DISCLAIMS ALL WARRANTIES WITH
 * REGARD TO THIS SOFTWARE INCLUDING ALL IMPLIED WARRANTIES OF MERCHANTABILITY AND
 * FITNESS. IN NO EVENT SHALL THE AUTHOR BE LIABLE FOR ANY SPECIAL, DIRECT,
 * INDIRECT, OR CONSEQUENTIAL DAMAGES OR ANY DAMAGES WHATSOEVER RESULTING FROM LOSS
 * OF USE, DATA OR PROFITS, WHETHER IN AN ACTION OF CONTRACT, NEGLIGENCE OR OTHER
 * TORTIOUS ACTION, ARISING OUT OF OR IN CONNECTION WITH THE USE OR PERFORMANCE OF
 * THIS SOFTWARE.
 *
 * MIT License for modified changes
 * Copyright (c) 2016-2021 Jeremy HU <jeremy-at-dust3d dot org>. All rights reserved.
 */
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
