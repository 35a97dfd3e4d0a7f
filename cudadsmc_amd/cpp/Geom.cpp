#include "Geom.h"

#include <cmath>
#include <stdexcept>

namespace dsmc_host {

std::vector<double> circle_polygon(int n, double cx, double cy, double r)
{
    if (n < 3) throw std::invalid_argument("a circle needs at least 3 points");
    std::vector<double> xy(2 * (size_t)n);
    for (int k = 0; k < n; ++k) {
        const double a = 2.0 * M_PI * (double)k / (double)n;
        xy[2 * (size_t)k] = cx + r * std::cos(a);
        xy[2 * (size_t)k + 1] = cy + r * std::sin(a);
    }
    return xy;
}

std::vector<double> wedge_polygon(double apex_x, double cy, double length, double half_angle_deg)
{
    const double h = length * std::tan(half_angle_deg * M_PI / 180.0);
    return { apex_x, cy, apex_x + length, cy - h, apex_x + length, cy + h };
}

std::vector<double> body_of(const Param &p)
{
    if (p.geometry == "cylinder") return circle_polygon(p.n_points, p.Center_x, p.Center_y, p.Radius);
    if (p.geometry == "wedge") return wedge_polygon(p.wedge_apex_x, p.Center_y, p.wedge_length, p.wedge_half_angle_deg);
    if (p.geometry == "none") return {};
    throw std::invalid_argument("unknown geometry '" + p.geometry + "'");
}

}  // namespace dsmc_host
