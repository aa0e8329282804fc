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
// Ear-clipping triangulation (see earclip.h for the algorithm and its source).
#include "earclip.h"

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdint>
#include <deque>
#include <limits>

namespace sbhost {
namespace {

struct Vtx {
    size_t i;  // index into the concatenated input rings
    double x, y;
    Vtx *prev = nullptr, *next = nullptr;
    int32_t z = 0;
    Vtx *prevZ = nullptr, *nextZ = nullptr;
    bool steiner = false;
    Vtx(size_t i_, double x_, double y_) : i(i_), x(x_), y(y_) {}
};

class Clipper {
public:
    std::vector<size_t> out;

    void run(const std::vector<Ring> &rings)
    {
        if (rings.empty())
            return;
        int budget = 80;
        for (size_t r = 0; budget >= 0 && r < rings.size(); ++r)
            budget -= (int)rings[r].size();
        Vtx *outer = ring_list(rings[0], true);
        if (!outer || outer->prev == outer->next)
            return;
        if (rings.size() > 1)
            outer = bridge_holes(rings, outer);
        zindex_ = budget < 0;
        if (zindex_) {
            minX_ = maxX_ = outer->x;
            minY_ = maxY_ = outer->y;
            for (Vtx *p = outer->next; p != outer; p = p->next) {
                minX_ = std::min(minX_, p->x);
                minY_ = std::min(minY_, p->y);
                maxX_ = std::max(maxX_, p->x);
                maxY_ = std::max(maxY_, p->y);
            }
            invSize_ = std::max(maxX_ - minX_, maxY_ - minY_);
            invSize_ = invSize_ != 0.0 ? 1.0 / invSize_ : 0.0;
        }
        slice(outer, 0);
    }

private:
    std::deque<Vtx> pool_;
    size_t base_ = 0;  // index of the first vertex of the ring being linked
    bool zindex_ = false;
    double minX_ = 0, minY_ = 0, maxX_ = 0, maxY_ = 0, invSize_ = 0;

    // signed area, clamped to 0 within DBL_EPSILON (earcut.hpp:694-699)
    static double tri_area(const Vtx *p, const Vtx *q, const Vtx *r)
    {
        double v = (q->y - p->y) * (r->x - q->x) - (q->x - p->x) * (r->y - q->y);
        return std::abs(v) <= DBL_EPSILON ? 0.0 : v;
    }
    static bool same(const Vtx *a, const Vtx *b) { return a->x == b->x && a->y == b->y; }
    static int sgn(double v) { return (0.0 < v) - (v < 0.0); }

    Vtx *push(size_t i, double x, double y, Vtx *last)
    {
        pool_.emplace_back(i, x, y);
        Vtx *p = &pool_.back();
        if (!last) {
            p->prev = p->next = p;
        } else {
            p->next = last->next;
            p->prev = last;
            last->next->prev = p;
            last->next = p;
        }
        return p;
    }

    static void unlink(Vtx *p)
    {
        p->next->prev = p->prev;
        p->prev->next = p->next;
        if (p->prevZ)
            p->prevZ->nextZ = p->nextZ;
        if (p->nextZ)
            p->nextZ->prevZ = p->prevZ;
    }

    // circular list of one ring in the requested winding
    Vtx *ring_list(const Ring &pts, bool clockwise)
    {
        const size_t n = pts.size();
        double s = 0.0;
        for (size_t i = 0, j = n ? n - 1 : 0; i < n; j = i++)
            s += (pts[j][0] - pts[i][0]) * (pts[i][1] + pts[j][1]);
        Vtx *last = nullptr;
        if (clockwise == (s > 0)) {
            for (size_t i = 0; i < n; ++i)
                last = push(base_ + i, pts[i][0], pts[i][1], last);
        } else {
            for (size_t i = n; i-- > 0;)
                last = push(base_ + i, pts[i][0], pts[i][1], last);
        }
        if (last && same(last, last->next)) {
            unlink(last);
            last = last->next;
        }
        base_ += n;
        return last;
    }

    // drop duplicate and collinear vertices between start and end
    Vtx *drop_flat(Vtx *start, Vtx *end = nullptr)
    {
        if (!end)
            end = start;
        Vtx *p = start;
        bool again;
        do {
            again = false;
            if (!p->steiner && (same(p, p->next) || tri_area(p->prev, p, p->next) == 0)) {
                unlink(p);
                p = end = p->prev;
                if (p == p->next)
                    break;
                again = true;
            } else {
                p = p->next;
            }
        } while (again || p != end);
        return end;
    }

    void emit(const Vtx *a, const Vtx *b, const Vtx *c)
    {
        out.push_back(a->i);
        out.push_back(b->i);
        out.push_back(c->i);
    }

    void slice(Vtx *ear, int pass)
    {
        if (!ear)
            return;
        if (!pass && zindex_)
            build_zindex(ear);
        Vtx *stop = ear;
        while (ear->prev != ear->next) {
            Vtx *prev = ear->prev;
            Vtx *next = ear->next;
            if (zindex_ ? is_ear_z(ear) : is_ear(ear)) {
                emit(prev, ear, next);
                unlink(ear);
                ear = next->next;  // skipping a vertex gives fewer slivers
                stop = next->next;
                continue;
            }
            ear = next;
            if (ear == stop) {
                if (!pass) {
                    slice(drop_flat(ear), 1);
                } else if (pass == 1) {
                    ear = cure_crossings(drop_flat(ear));
                    slice(ear, 2);
                } else if (pass == 2) {
                    split_slice(ear);
                }
                break;
            }
        }
    }

    static bool in_triangle(double ax, double ay, double bx, double by, double cx, double cy, double px, double py)
    {
        return (cx - px) * (ay - py) - (ax - px) * (cy - py) >= 0 && (ax - px) * (by - py) - (bx - px) * (ay - py) >= 0 &&
               (bx - px) * (cy - py) - (cx - px) * (by - py) >= 0;
    }

    bool is_ear(Vtx *ear) const
    {
        const Vtx *a = ear->prev, *b = ear, *c = ear->next;
        if (tri_area(a, b, c) >= 0)
            return false;  // reflex
        for (Vtx *p = ear->next->next; p != ear->prev; p = p->next)
            if (in_triangle(a->x, a->y, b->x, b->y, c->x, c->y, p->x, p->y) && tri_area(p->prev, p, p->next) >= 0)
                return false;
        return true;
    }

    int32_t zcode(double x_, double y_) const
    {
        int32_t x = static_cast<int32_t>(32767.0 * (x_ - minX_) * invSize_);
        int32_t y = static_cast<int32_t>(32767.0 * (y_ - minY_) * invSize_);
        x = (x | (x << 8)) & 0x00FF00FF;
        x = (x | (x << 4)) & 0x0F0F0F0F;
        x = (x | (x << 2)) & 0x33333333;
        x = (x | (x << 1)) & 0x55555555;
        y = (y | (y << 8)) & 0x00FF00FF;
        y = (y | (y << 4)) & 0x0F0F0F0F;
        y = (y | (y << 2)) & 0x33333333;
        y = (y | (y << 1)) & 0x55555555;
        return x | (y << 1);
    }

    bool is_ear_z(Vtx *ear) const
    {
        const Vtx *a = ear->prev, *b = ear, *c = ear->next;
        if (tri_area(a, b, c) >= 0)
            return false;
        const double x0 = std::min(a->x, std::min(b->x, c->x)), y0 = std::min(a->y, std::min(b->y, c->y));
        const double x1 = std::max(a->x, std::max(b->x, c->x)), y1 = std::max(a->y, std::max(b->y, c->y));
        const int32_t zlo = zcode(x0, y0), zhi = zcode(x1, y1);
        for (Vtx *p = ear->nextZ; p && p->z <= zhi; p = p->nextZ)
            if (p != ear->prev && p != ear->next && in_triangle(a->x, a->y, b->x, b->y, c->x, c->y, p->x, p->y) &&
                tri_area(p->prev, p, p->next) >= 0)
                return false;
        for (Vtx *p = ear->prevZ; p && p->z >= zlo; p = p->prevZ)
            if (p != ear->prev && p != ear->next && in_triangle(a->x, a->y, b->x, b->y, c->x, c->y, p->x, p->y) &&
                tri_area(p->prev, p, p->next) >= 0)
                return false;
        return true;
    }

    void build_zindex(Vtx *start)
    {
        Vtx *p = start;
        do {
            p->z = p->z ? p->z : zcode(p->x, p->y);
            p->prevZ = p->prev;
            p->nextZ = p->next;
            p = p->next;
        } while (p != start);
        p->prevZ->nextZ = nullptr;
        p->prevZ = nullptr;
        zsort(p);
    }

    // bottom-up merge sort of the z list (stable, as Tatham's list sort)
    static Vtx *zsort(Vtx *list)
    {
        int run = 1;
        for (;;) {
            Vtx *p = list, *tail = nullptr;
            list = nullptr;
            int merges = 0;
            while (p) {
                ++merges;
                Vtx *q = p;
                int psize = 0;
                for (int k = 0; k < run; ++k) {
                    ++psize;
                    q = q->nextZ;
                    if (!q)
                        break;
                }
                int qsize = run;
                while (psize > 0 || (qsize > 0 && q)) {
                    Vtx *e;
                    if (psize == 0) {
                        e = q;
                        q = q->nextZ;
                        --qsize;
                    } else if (qsize == 0 || !q) {
                        e = p;
                        p = p->nextZ;
                        --psize;
                    } else if (p->z <= q->z) {
                        e = p;
                        p = p->nextZ;
                        --psize;
                    } else {
                        e = q;
                        q = q->nextZ;
                        --qsize;
                    }
                    if (tail)
                        tail->nextZ = e;
                    else
                        list = e;
                    e->prevZ = tail;
                    tail = e;
                }
                p = q;
            }
            tail->nextZ = nullptr;
            if (merges <= 1)
                return list;
            run *= 2;
        }
    }

    static bool on_segment(const Vtx *p, const Vtx *q, const Vtx *r)
    {
        return q->x <= std::max(p->x, r->x) && q->x >= std::min(p->x, r->x) && q->y <= std::max(p->y, r->y) &&
               q->y >= std::min(p->y, r->y);
    }

    static bool crosses(const Vtx *p1, const Vtx *q1, const Vtx *p2, const Vtx *q2)
    {
        const int o1 = sgn(tri_area(p1, q1, p2)), o2 = sgn(tri_area(p1, q1, q2));
        const int o3 = sgn(tri_area(p2, q2, p1)), o4 = sgn(tri_area(p2, q2, q1));
        if (o1 != o2 && o3 != o4)
            return true;
        if (o1 == 0 && on_segment(p1, p2, q1))
            return true;
        if (o2 == 0 && on_segment(p1, q2, q1))
            return true;
        if (o3 == 0 && on_segment(p2, p1, q2))
            return true;
        if (o4 == 0 && on_segment(p2, q1, q2))
            return true;
        return false;
    }

    static bool locally_inside(const Vtx *a, const Vtx *b)
    {
        return tri_area(a->prev, a, a->next) < 0 ? tri_area(a, b, a->next) >= 0 && tri_area(a, a->prev, b) >= 0
                                                 : tri_area(a, b, a->prev) < 0 || tri_area(a, a->next, b) < 0;
    }

    Vtx *cure_crossings(Vtx *start)
    {
        Vtx *p = start;
        do {
            Vtx *a = p->prev, *b = p->next->next;
            if (!same(a, b) && crosses(a, p, p->next, b) && locally_inside(a, b) && locally_inside(b, a)) {
                emit(a, p, b);
                unlink(p);
                unlink(p->next);
                p = start = b;
            }
            p = p->next;
        } while (p != start);
        return drop_flat(p);
    }

    static bool hits_polygon(const Vtx *a, const Vtx *b)
    {
        const Vtx *p = a;
        do {
            if (p->i != a->i && p->next->i != a->i && p->i != b->i && p->next->i != b->i && crosses(p, p->next, a, b))
                return true;
            p = p->next;
        } while (p != a);
        return false;
    }

    static bool middle_inside(const Vtx *a, const Vtx *b)
    {
        const Vtx *p = a;
        bool inside = false;
        const double px = (a->x + b->x) / 2, py = (a->y + b->y) / 2;
        do {
            if (((p->y > py) != (p->next->y > py)) && p->next->y != p->y &&
                (px < (p->next->x - p->x) * (py - p->y) / (p->next->y - p->y) + p->x))
                inside = !inside;
            p = p->next;
        } while (p != a);
        return inside;
    }

    static bool valid_diagonal(const Vtx *a, const Vtx *b)
    {
        return a->next->i != b->i && a->prev->i != b->i && !hits_polygon(a, b) &&
               ((locally_inside(a, b) && locally_inside(b, a) && middle_inside(a, b) &&
                 (tri_area(a->prev, a, b->prev) != 0.0 || tri_area(a, b->prev, b) != 0.0)) ||
                (same(a, b) && tri_area(a->prev, a, a->next) > 0 && tri_area(b->prev, b, b->next) > 0));
    }

    // join a and b by two coincident diagonals; returns the copy of b
    Vtx *cut(Vtx *a, Vtx *b)
    {
        pool_.emplace_back(a->i, a->x, a->y);
        Vtx *a2 = &pool_.back();
        pool_.emplace_back(b->i, b->x, b->y);
        Vtx *b2 = &pool_.back();
        Vtx *an = a->next, *bp = b->prev;
        a->next = b;
        b->prev = a;
        a2->next = an;
        an->prev = a2;
        b2->next = a2;
        a2->prev = b2;
        bp->next = b2;
        b2->prev = bp;
        return b2;
    }

    void split_slice(Vtx *start)
    {
        Vtx *a = start;
        do {
            for (Vtx *b = a->next->next; b != a->prev; b = b->next) {
                if (a->i != b->i && valid_diagonal(a, b)) {
                    Vtx *c = cut(a, b);
                    a = drop_flat(a, a->next);
                    c = drop_flat(c, c->next);
                    slice(a, 0);
                    slice(c, 0);
                    return;
                }
            }
            a = a->next;
        } while (a != start);
    }

    static Vtx *leftmost(Vtx *start)
    {
        Vtx *p = start, *l = start;
        do {
            if (p->x < l->x || (p->x == l->x && p->y < l->y))
                l = p;
            p = p->next;
        } while (p != start);
        return l;
    }

    Vtx *bridge_holes(const std::vector<Ring> &rings, Vtx *outer)
    {
        std::vector<Vtx *> queue;
        for (size_t r = 1; r < rings.size(); ++r) {
            Vtx *list = ring_list(rings[r], false);
            if (list) {
                if (list == list->next)
                    list->steiner = true;
                queue.push_back(leftmost(list));
            }
        }
        std::sort(queue.begin(), queue.end(), [](const Vtx *a, const Vtx *b) { return a->x < b->x; });
        for (Vtx *h : queue) {
            Vtx *m = hole_bridge(h, outer);
            if (m) {
                Vtx *b = cut(m, h);
                drop_flat(m, m->next);
                drop_flat(b, b->next);
            }
            outer = drop_flat(outer, outer->next);
        }
        return outer;
    }

    static bool sector_contains(const Vtx *m, const Vtx *p)
    {
        return tri_area(m->prev, m, p->prev) < 0 && tri_area(p->next, m, m->next) < 0;
    }

    // Eberly's search for a vertex of the outer ring visible from the hole's
    // leftmost vertex
    static Vtx *hole_bridge(Vtx *hole, Vtx *outer)
    {
        Vtx *p = outer;
        const double hx = hole->x, hy = hole->y;
        double qx = -std::numeric_limits<double>::infinity();
        Vtx *m = nullptr;
        do {
            if (hy <= p->y && hy >= p->next->y && p->next->y != p->y) {
                const double x = p->x + (hy - p->y) * (p->next->x - p->x) / (p->next->y - p->y);
                if (x <= hx && x > qx) {
                    qx = x;
                    if (x == hx) {
                        if (hy == p->y)
                            return p;
                        if (hy == p->next->y)
                            return p->next;
                    }
                    m = p->x < p->next->x ? p : p->next;
                }
            }
            p = p->next;
        } while (p != outer);
        if (!m)
            return nullptr;
        if (hx == qx)
            return m;
        const Vtx *stop = m;
        double tanMin = std::numeric_limits<double>::infinity();
        const double mx = m->x, my = m->y;
        p = m;
        do {
            if (hx >= p->x && p->x >= mx && hx != p->x &&
                in_triangle(hy < my ? hx : qx, hy, mx, my, hy < my ? qx : hx, hy, p->x, p->y)) {
                const double t = std::abs(hy - p->y) / (hx - p->x);
                if (locally_inside(p, hole) && (t < tanMin || (t == tanMin && (p->x > m->x || sector_contains(m, p))))) {
                    m = p;
                    tanMin = t;
                }
            }
            p = p->next;
        } while (p != stop);
        return m;
    }
};

}  // namespace

std::vector<size_t> earclip(const std::vector<Ring> &rings)
{
    Clipper c;
    c.run(rings);
    return std::move(c.out);
}

}  // namespace sbhost
