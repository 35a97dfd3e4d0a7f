#include "CartesianMesh.h"

#include <cmath>
#include <stdexcept>

#include "../csrc/host/geometry.h"

namespace dsmc_host {

CartesianMesh::CartesianMesh(int N1, int N2, double L1, double L2)
    : m_N1(N1), m_N2(N2), m_L1(L1), m_L2(L2), m_dx(L1 / N1), m_dy(L2 / N2)
{
    if (N1 < 1 || N2 < 1 || !(L1 > 0) || !(L2 > 0)) throw std::invalid_argument("CartesianMesh: bad grid");
    m_frac.assign((size_t)N1 * N2, 1.0);
    m_cut.assign((size_t)N1 * N2, 0);
}

int CartesianMesh::getIndex(double x, double y) const
{
    int i = (int)(x / m_dx), j = (int)(y / m_dy);
    i = i < 0 ? 0 : (i > m_N1 - 1 ? m_N1 - 1 : i);
    j = j < 0 ? 0 : (j > m_N2 - 1 ? m_N2 - 1 : j);
    return j + i * m_N2;
}

void CartesianMesh::cutcell(const std::vector<double> &outline)
{
    m_frac.assign((size_t)m_N1 * m_N2, 1.0);
    m_cut.assign((size_t)m_N1 * m_N2, 0);
    m_seg.clear();
    const size_t n = outline.size() / 2;
    if (n == 0) return;
    if (n < 3 || outline.size() % 2) throw std::invalid_argument("cutcell: the outline needs >= 3 points");
    std::vector<dsmc::geom::V2> poly(n);
    for (size_t k = 0; k < n; ++k) poly[k] = dsmc::geom::V2{ outline[2 * k], outline[2 * k + 1] };
    dsmc::geom::cell_fluid_fractions(poly, m_N1, m_N2, m_dx, m_dy, m_frac);
    for (size_t c = 0; c < m_frac.size(); ++c) m_cut[c] = (m_frac[c] > 0.0 && m_frac[c] < 1.0) ? 1 : 0;
    // one segment per edge, normal (ey, -ex)/|e|: out of a counter-clockwise
    // outline, into the fluid (Circle.cpp:33-36)
    m_seg.resize(6 * n);
    for (size_t k = 0; k < n; ++k) {
        const dsmc::geom::V2 &a = poly[k], &b = poly[(k + 1) % n];
        const double ex = b.x - a.x, ey = b.y - a.y, L = std::hypot(ex, ey);
        if (!(L > 0.0)) throw std::invalid_argument("cutcell: degenerate edge");
        double *s = &m_seg[6 * k];
        s[0] = a.x; s[1] = a.y; s[2] = b.x; s[3] = b.y;
        s[4] = ey / L; s[5] = -ex / L;
    }
}

}  // namespace dsmc_host
