// Plane geometry of the cut cells, header-only and host-only (plain C++):
// the library's case set-up (case_setup.cpp) and the C++ host layer's mesh
// (cpp/CartesianMesh.cpp) both use it, so a host-built table equals the
// library's own bit for bit.
//
//  - clipped_area:       area of a polygon within an axis-aligned rectangle
//                        (Sutherland-Hodgman clipping, the rectangle's four
//                        half-planes in turn; any simple polygon);
//  - segment_hits_rect:  does a segment touch a closed rectangle
//                        (Liang-Barsky);
//  - point_in_polygon:   even-odd crossing rule;
//  - fluid_bbox:         bounding box of (rectangle minus body);
//  - fluid_mask:         which sub-boxes of an 8x8 split of a box touch fluid.
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <vector>

namespace dsmc {
namespace geom {

struct V2 { double x, y; };

// Area of (polygon intersect axis-aligned rectangle), Sutherland-Hodgman.
inline double clipped_area(const std::vector<V2> &poly, double x0, double y0, double x1, double y1)
{
    std::vector<V2> cur = poly, nxt;
    auto clip = [&](auto inside, auto cross) {
        nxt.clear();
        size_t n = cur.size();
        for (size_t i = 0; i < n; ++i) {
            V2 a = cur[i], b = cur[(i + 1) % n];
            bool ia = inside(a), ib = inside(b);
            if (ia) nxt.push_back(a);
            if (ia != ib) nxt.push_back(cross(a, b));
        }
        cur.swap(nxt);
    };
    auto xcut = [](double X) {
        return [X](V2 a, V2 b) { double t = (X - a.x) / (b.x - a.x); return V2{ X, a.y + t * (b.y - a.y) }; };
    };
    auto ycut = [](double Y) {
        return [Y](V2 a, V2 b) { double t = (Y - a.y) / (b.y - a.y); return V2{ a.x + t * (b.x - a.x), Y }; };
    };
    clip([&](V2 p) { return p.x >= x0; }, xcut(x0));
    if (cur.empty()) return 0.0;
    clip([&](V2 p) { return p.x <= x1; }, xcut(x1));
    if (cur.empty()) return 0.0;
    clip([&](V2 p) { return p.y >= y0; }, ycut(y0));
    if (cur.empty()) return 0.0;
    clip([&](V2 p) { return p.y <= y1; }, ycut(y1));
    if (cur.size() < 3) return 0.0;
    // shoelace relative to the cell corner (avoids cancellation)
    double a = 0.0;
    for (size_t i = 0; i < cur.size(); ++i) {
        const V2 &p = cur[i], &q = cur[(i + 1) % cur.size()];
        double px = p.x - x0, py = p.y - y0, qx = q.x - x0, qy = q.y - y0;
        a += px * qy - qx * py;
    }
    return std::fabs(0.5 * a);
}

// Does segment a-b touch the closed rectangle [x0,x1]x[y0,y1]?  (Liang-Barsky)
inline bool segment_hits_rect(V2 a, V2 b, double x0, double y0, double x1, double y1)
{
    double t0 = 0.0, t1 = 1.0;
    double dx = b.x - a.x, dy = b.y - a.y;
    const double p[4] = { -dx, dx, -dy, dy };
    const double q[4] = { a.x - x0, x1 - a.x, a.y - y0, y1 - a.y };
    for (int k = 0; k < 4; ++k) {
        if (p[k] == 0.0) {
            if (q[k] < 0.0) return false;
        } else {
            double r = q[k] / p[k];
            if (p[k] < 0.0) { if (r > t1) return false; if (r > t0) t0 = r; }
            else            { if (r < t0) return false; if (r < t1) t1 = r; }
        }
    }
    return t0 <= t1;
}

inline bool point_in_polygon(const std::vector<V2> &poly, double px, double py)
{
    bool in = false;
    const size_t n = poly.size();
    for (size_t k = 0; k < n; ++k) {
        const V2 &a = poly[k], &b = poly[(k + 1) % n];
        if ((a.y > py) != (b.y > py)) {
            double xi = a.x + (py - a.y) * (b.x - a.x) / (b.y - a.y);
            if (px < xi) in = !in;
        }
    }
    return in;
}

// Bounding box of (cell rectangle minus body): the rectangle corners outside
// the body, the crossings of body edges with the rectangle boundary and the
// body vertices inside the rectangle.  Grown by one part in 1e5 of the cell
// and clipped to the cell.  Used to place fill particles in cut cells.
inline void fluid_bbox(const std::vector<V2> &poly, double x0, double y0, double x1, double y1, float *out)
{
    double bx0 = 1e300, by0 = 1e300, bx1 = -1e300, by1 = -1e300;
    auto add = [&](double x, double y) {
        bx0 = std::min(bx0, x); by0 = std::min(by0, y);
        bx1 = std::max(bx1, x); by1 = std::max(by1, y);
    };
    const double cx[4] = { x0, x1, x1, x0 }, cy[4] = { y0, y0, y1, y1 };
    for (int k = 0; k < 4; ++k)
        if (!point_in_polygon(poly, cx[k], cy[k])) add(cx[k], cy[k]);
    const size_t n = poly.size();
    for (size_t k = 0; k < n; ++k) {
        const V2 &a = poly[k], &b = poly[(k + 1) % n];
        if (a.x >= x0 && a.x <= x1 && a.y >= y0 && a.y <= y1) add(a.x, a.y);
        // crossings with the four rectangle sides
        for (int s = 0; s < 4; ++s) {
            const bool vert = (s & 1) == 0;             // s = 0: x = x0, 2: x = x1; 1: y = y0, 3: y = y1
            const double L = s == 0 ? x0 : s == 2 ? x1 : s == 1 ? y0 : y1;
            const double pa = vert ? a.x : a.y, pb = vert ? b.x : b.y;
            if ((pa - L) * (pb - L) > 0.0 || pa == pb) continue;
            const double t = (L - pa) / (pb - pa);
            const double q = vert ? a.y + t * (b.y - a.y) : a.x + t * (b.x - a.x);
            const double lo = vert ? y0 : x0, hi = vert ? y1 : x1;
            if (q < lo || q > hi) continue;
            if (vert) add(L, q); else add(q, L);
        }
    }
    if (bx0 > bx1) { bx0 = x0; bx1 = x1; by0 = y0; by1 = y1; }
    const double gx = 1e-5 * (x1 - x0), gy = 1e-5 * (y1 - y0);
    out[0] = (float)std::max(x0, bx0 - gx);
    out[1] = (float)std::max(y0, by0 - gy);
    out[2] = (float)std::min(x1, bx1 + gx);
    out[3] = (float)std::min(y1, by1 + gy);
}

// Which sub-boxes of the 8x8 split of a fluid bounding box touch fluid
// (are not entirely inside the body).  The fill draws a marked sub-box, then
// a point in it, then rejects points inside the body: uniform over the fluid,
// with an acceptance rate that does not collapse for thin fluid slivers.
inline uint64_t fluid_mask(const std::vector<V2> &poly, const float *fb)
{
    const double x0 = fb[0], y0 = fb[1], w = ((double)fb[2] - x0) / 8.0, h = ((double)fb[3] - y0) / 8.0;
    uint64_t m = 0;
    for (int sy = 0; sy < 8; ++sy)
        for (int sx = 0; sx < 8; ++sx) {
            const double a0 = x0 + sx * w, b0 = y0 + sy * h;
            const double solid = clipped_area(poly, a0, b0, a0 + w, b0 + h);
            if (solid < w * h * (1.0 - 1e-9)) m |= 1ull << (sx + 8 * sy);
        }
    return m ? m : ~0ull;
}

// Fluid fraction of every cell c = j + i*ny of an nx x ny grid of dx x dy
// cells against a counter-clockwise body polygon: 1 - (polygon area inside
// the cell) / (dx dy), exactly 1 for cells the polygon does not reach and 0
// below 1e-9 (the three wall-vertex cases of the reference's
// CartesianMesh.cpp:112-168, computed exactly by clipping).
inline void cell_fluid_fractions(const std::vector<V2> &poly, int nx, int ny, double dx, double dy,
                                 std::vector<double> &frac)
{
    frac.assign((size_t)nx * ny, 1.0);
    double bx0 = 1e300, bx1 = -1e300, by0 = 1e300, by1 = -1e300;
    for (const V2 &v : poly) { bx0 = std::min(bx0, v.x); bx1 = std::max(bx1, v.x); by0 = std::min(by0, v.y); by1 = std::max(by1, v.y); }
    const double cell_area = dx * dy;
    const int i0 = std::max(0, (int)std::floor(bx0 / dx) - 1), i1 = std::min(nx - 1, (int)std::floor(bx1 / dx) + 1);
    const int j0 = std::max(0, (int)std::floor(by0 / dy) - 1), j1 = std::min(ny - 1, (int)std::floor(by1 / dy) + 1);
    for (int i = i0; i <= i1; ++i)
        for (int j = j0; j <= j1; ++j) {
            const double ain = clipped_area(poly, i * dx, j * dy, (i + 1) * dx, (j + 1) * dy);
            double fr = 1.0 - ain / cell_area;
            if (fr > 1.0 - 1e-9) continue;  // uncut: keep exactly 1.0
            if (fr < 1e-9) fr = 0.0;
            frac[(size_t)j + (size_t)i * ny] = fr;
        }
}

}  // namespace geom
}  // namespace dsmc
