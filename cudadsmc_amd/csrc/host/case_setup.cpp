// Host-side case construction (see case_setup.h).
#include "case_setup.h"

#include <algorithm>
#include <cmath>
#include <sstream>

namespace dsmc {

namespace {

const double kPi = 3.14159265358979323846;

// Maxwellian number flux through a face per unit area (Bird 1994, eq. 4.22):
// n vmp / (2 sqrt(pi)) [exp(-a^2) + sqrt(pi) a (1 + erf(a))], a = u.n / vmp.
double maxwell_flux(double n, double vmp, double a)
{
    return n * vmp / (2.0 * std::sqrt(kPi)) *
           (std::exp(-a * a) + std::sqrt(kPi) * a * (1.0 + std::erf(a)));
}

}  // namespace

std::string Param::derive(const dsmc_params &p)
{
    in = p;
    if (!(p.boltz > 0 && p.mass > 0 && p.diam > 0 && p.Tref > 0)) return "gas constants must be positive";
    if (!(p.Vtl > 0.5 && p.Vtl < 1.5)) return "Vtl (omega) must lie in (0.5, 1.5)";
    if (!(p.L1 > 0 && p.L2 > 0 && p.L3 > 0)) return "domain extents must be positive";
    if (p.N1 < 1 || p.N2 < 1) return "N1, N2 must be >= 1";
    if ((long long)p.N1 * p.N2 > (1LL << 28)) return "too many cells";
    if (!(p.n_inf > 0 && p.T > 0)) return "n_inf and T must be positive";
    if (!(p.tau > 0)) return "tau must be positive";
    if (!(p.Fn > 0) && !(p.ppc > 0)) return "either Fn or ppc must be positive";
    for (int f = 0; f < 4; ++f) {
        if (p.face_bc[f] < DSMC_BC_STREAM || p.face_bc[f] > DSMC_BC_DIFFUSE) return "bad face_bc";
        if (p.face_bc[f] == DSMC_BC_DIFFUSE && !(p.face_T[f] > 0)) return "diffuse face needs face_T > 0";
    }
    if (p.wall_bc != DSMC_BC_DIFFUSE && p.wall_bc != DSMC_BC_SPECULAR) return "wall_bc must be DIFFUSE or SPECULAR";
    if (p.wall_bc == DSMC_BC_DIFFUSE && !(p.T_wall > 0)) return "diffuse body wall needs T_wall > 0";

    dx = p.L1 / p.N1;
    dy = p.L2 / p.N2;
    Vcell = dx * dy * p.L3;
    if (p.Fn > 0) {
        Fn = p.Fn;
        ppc_eff = p.n_inf * Vcell / Fn;
    } else {
        Fn = p.n_inf * Vcell / p.ppc;
        ppc_eff = p.ppc;
    }
    // Run.cpp:32-35 (Vstd, Vmax); VHS_coe = Gamma(5/2 - omega) with a correct
    // Gamma function (defect D5: the reference GammaFun returns 0.6257 for 1.69).
    Vstd = std::sqrt(2.0 * p.boltz * p.T / p.mass);
    Vmax = 2.0 * std::sqrt(8.0 / kPi) * Vstd;
    VHS_coe = std::tgamma(2.5 - p.Vtl);
    // sigma_T c_r = pi d^2 (2 k Tref / (m_r c_r^2))^(omega-1/2) / Gamma(5/2-omega) * c_r
    //             = K1 (c_r^2)^(1-omega),  m_r = m/2   (GPUCell.cu:251, Bird eq. 4.63)
    K1 = kPi * p.diam * p.diam * std::pow(4.0 * p.boltz * p.Tref / p.mass, p.Vtl - 0.5) / VHS_coe;
    e_exp = 1.0 - p.Vtl;
    ntc_coef = Fn * p.tau / 2.0;
    sigcmax0 = Vmax * kPi * p.diam * p.diam;  // GPUCell.cu:224
    if (p.wall_bc == DSMC_BC_DIFFUSE) {
        body_vmp = std::sqrt(2.0 * p.boltz * p.T_wall / p.mass);
        body_sig = std::sqrt(p.boltz * p.T_wall / p.mass);
    }
    for (int f = 0; f < 4; ++f) {
        if (p.face_bc[f] == DSMC_BC_DIFFUSE) {
            face_vmp[f] = std::sqrt(2.0 * p.boltz * p.face_T[f] / p.mass);
            face_sig[f] = std::sqrt(p.boltz * p.face_T[f] / p.mass);
        }
    }
    inj_vmp = Vstd;
    inj_sig = std::sqrt(p.boltz * p.T / p.mass);
    const double nrm[4][2] = { { 1, 0 }, { -1, 0 }, { 0, 1 }, { 0, -1 } };
    const double len[4] = { p.L2, p.L2, p.L1, p.L1 };
    for (int f = 0; f < 4; ++f) {
        inj_a[f] = (p.V_jet * nrm[f][0] + p.V_y * nrm[f][1]) / inj_vmp;
        inj_e[f] = 0.0;
        if (p.face_bc[f] == DSMC_BC_STREAM)
            inj_e[f] = maxwell_flux(p.n_inf, inj_vmp, inj_a[f]) * len[f] * p.L3 * p.tau / Fn;
    }
    seed32 = (uint32_t)(p.seed ^ (p.seed >> 32));
    return "";
}

namespace {

struct V2 { double x, y; };

// Area of (polygon intersect axis-aligned rectangle), Sutherland-Hodgman.
double clipped_area(const std::vector<V2> &poly, double x0, double y0, double x1, double y1)
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
bool segment_hits_rect(V2 a, V2 b, double x0, double y0, double x1, double y1)
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

bool point_in_polygon(const std::vector<V2> &poly, double px, double py)
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
void fluid_bbox(const std::vector<V2> &poly, double x0, double y0, double x1, double y1, float *out)
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
uint64_t fluid_mask(const std::vector<V2> &poly, const float *fb)
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

}  // namespace

std::string build_cut_tables(const Param &P, const dsmc_body *body, CutTables &T)
{
    const int nx = P.in.N1, ny = P.in.N2;
    const int C = nx * ny;
    T = CutTables{};
    T.ncell = C;
    T.margin = (float)std::min(P.dx, P.dy);
    T.vol.assign(C, P.Vcell);
    T.frac.assign(C, 1.0);
    T.vol_ntc.assign(C, (float)P.Vcell);
    T.cut.assign(C, 0);
    T.fbox.assign(4 * (size_t)C, 0.0f);
    T.fmask.assign(C, ~0ull);
    T.cseg_start.assign(C + 1, 0);
    if (!body || body->n_points == 0) return "";
    if (body->n_points < 3 || !body->xy) return "body polygon needs >= 3 points";

    const int np = body->n_points;
    std::vector<V2> poly(np);
    double area2 = 0.0;
    for (int k = 0; k < np; ++k) {
        poly[k] = V2{ body->xy[2 * k], body->xy[2 * k + 1] };
        if (!std::isfinite(poly[k].x) || !std::isfinite(poly[k].y)) return "non-finite body point";
        // a body reaching a face would cover part of it: inflow there would
        // start inside the body
        if (poly[k].x < 0.0 || poly[k].x > P.in.L1 || poly[k].y < 0.0 || poly[k].y > P.in.L2)
            return "body polygon must lie inside the domain [0, L1] x [0, L2]";
    }
    for (int k = 0; k < np; ++k) {
        const V2 &a = poly[k], &b = poly[(k + 1) % np];
        area2 += a.x * b.y - b.x * a.y;
    }
    if (!(area2 > 0.0)) return "body polygon must be counter-clockwise with positive area";

    // One wall segment per polygon edge; the normal (ey, -ex)/|e| points out of
    // a counter-clockwise polygon, i.e. into the fluid (Circle.cpp:33-36).
    T.nseg = np;
    T.seg.resize(6 * (size_t)np);
    for (int k = 0; k < np; ++k) {
        const V2 &a = poly[k], &b = poly[(k + 1) % np];
        double ex = b.x - a.x, ey = b.y - a.y, L = std::hypot(ex, ey);
        if (!(L > 0.0)) return "degenerate body edge";
        float *s = &T.seg[6 * (size_t)k];
        s[0] = (float)a.x; s[1] = (float)a.y; s[2] = (float)b.x; s[3] = (float)b.y;
        s[4] = (float)(ey / L); s[5] = (float)(-ex / L);
    }

    // Per-cell reachable-segment lists: cell rectangle grown by the margin
    // (a particle whose flight is <= margin cannot reach anything else).
    const double grow = (double)T.margin * 1.001 + 1e-9 * std::max(P.in.L1, P.in.L2);
    std::vector<std::vector<int>> lists(C);
    for (int k = 0; k < np; ++k) {
        const V2 &a = poly[k], &b = poly[(k + 1) % np];
        double xlo = std::min(a.x, b.x) - grow, xhi = std::max(a.x, b.x) + grow;
        double ylo = std::min(a.y, b.y) - grow, yhi = std::max(a.y, b.y) + grow;
        int i0 = std::max(0, (int)std::floor(xlo / P.dx) - 1), i1 = std::min(nx - 1, (int)std::floor(xhi / P.dx) + 1);
        int j0 = std::max(0, (int)std::floor(ylo / P.dy) - 1), j1 = std::min(ny - 1, (int)std::floor(yhi / P.dy) + 1);
        for (int i = i0; i <= i1; ++i)
            for (int j = j0; j <= j1; ++j) {
                double x0 = i * P.dx - grow, x1 = (i + 1) * P.dx + grow;
                double y0 = j * P.dy - grow, y1 = (j + 1) * P.dy + grow;
                if (segment_hits_rect(a, b, x0, y0, x1, y1)) lists[cell_id(i, j, ny)].push_back(k);
            }
    }
    for (int c = 0; c < C; ++c) {
        T.cseg_start[c + 1] = T.cseg_start[c] + (int)lists[c].size();
        for (int k : lists[c]) T.cseg_idx.push_back(k);
    }

    // Fluid volume of every cell: cell area minus its overlap with the body
    // (the three wall-vertex cases of CartesianMesh.cpp:112-168, computed
    // exactly by polygon clipping).
    double bx0 = 1e300, bx1 = -1e300, by0 = 1e300, by1 = -1e300;
    for (const V2 &v : poly) { bx0 = std::min(bx0, v.x); bx1 = std::max(bx1, v.x); by0 = std::min(by0, v.y); by1 = std::max(by1, v.y); }
    const double cell_area = P.dx * P.dy;
    int i0 = std::max(0, (int)std::floor(bx0 / P.dx) - 1), i1 = std::min(nx - 1, (int)std::floor(bx1 / P.dx) + 1);
    int j0 = std::max(0, (int)std::floor(by0 / P.dy) - 1), j1 = std::min(ny - 1, (int)std::floor(by1 / P.dy) + 1);
    for (int i = i0; i <= i1; ++i)
        for (int j = j0; j <= j1; ++j) {
            double ain = clipped_area(poly, i * P.dx, j * P.dy, (i + 1) * P.dx, (j + 1) * P.dy);
            int c = cell_id(i, j, ny);
            double fr = 1.0 - ain / cell_area;
            if (fr > 1.0 - 1e-9) continue;  // uncut: keep exactly 1.0
            if (fr < 1e-9) fr = 0.0;
            T.frac[c] = fr;
            T.vol[c] = fr * P.Vcell;
            T.vol_ntc[c] = (float)std::max(T.vol[c], 0.1 * P.Vcell);
            T.cut[c] = fr > 0.0 ? 1 : 0;
            if (fr > 0.0) {
                ++T.ncut;
                fluid_bbox(poly, i * P.dx, j * P.dy, (i + 1) * P.dx, (j + 1) * P.dy, &T.fbox[4 * (size_t)c]);
                T.fmask[c] = fluid_mask(poly, &T.fbox[4 * (size_t)c]);
            } else {
                T.vol_ntc[c] = 0.0f;
            }
        }
    return "";
}

}  // namespace dsmc
