// Host-side case construction (see case_setup.h).
#include "case_setup.h"

#include "geometry.h"

#include <algorithm>
#include <cmath>
#include <sstream>

namespace dsmc {

using geom::V2;

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
    if (p.zeta != 0 && p.zeta != 2) return "zeta must be 0 (monatomic) or 2 (diatomic, Larsen-Borgnakke)";
    if (p.zeta == 2 && !(p.rot_prob > 0 && p.rot_prob <= 1)) return "rot_prob (1/Z_rot) must lie in (0, 1]";

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
    // Larsen-Borgnakke (zeta = 2): translational fraction Beta(5/2-omega, 2)
    // as U1^(1/(5/2-omega)) U2^(1/(7/2-omega)); rotational energy per unit
    // mass in equilibrium at T has mean kT/m
    sample_blocks = 6;
    if (p.zeta == 2) {
        lb_ia = 1.0 / (2.5 - p.Vtl);
        lb_ib = 1.0 / (3.5 - p.Vtl);
        body_ekt = p.wall_bc == DSMC_BC_DIFFUSE ? p.boltz * p.T_wall / p.mass : 0.0;
        for (int f = 0; f < 4; ++f)
            face_ekt[f] = p.face_bc[f] == DSMC_BC_DIFFUSE ? p.boltz * p.face_T[f] / p.mass : 0.0;
        inj_ekt = p.boltz * p.T / p.mass;
        sample_blocks = 7;
    }
    seed32 = (uint32_t)(p.seed ^ (p.seed >> 32));
    return "";
}

namespace {

// Everything a case needs from its body outline `poly` (counter-clockwise,
// checked by the callers): the wall segments -- from `seg_in` (6 per edge, as
// a host mesh layer computed them) or from the outline's edges -- the
// per-cell reach lists, the fluid fractions -- from `frac_in` (and the cut
// flags `cut_in`) or by clipping -- and the fill helpers of the cut cells.
std::string tables_from_outline(const Param &P, const std::vector<V2> &poly, const double *seg_in,
                                const double *frac_in, const uint8_t *cut_in, CutTables &T)
{
    const int nx = P.in.N1, ny = P.in.N2, C = nx * ny, np = (int)poly.size();
    for (const V2 &v : poly) {
        if (!std::isfinite(v.x) || !std::isfinite(v.y)) return "non-finite body point";
        // a body reaching a face would cover part of it: inflow there would
        // start inside the body
        if (v.x < 0.0 || v.x > P.in.L1 || v.y < 0.0 || v.y > P.in.L2)
            return "body polygon must lie inside the domain [0, L1] x [0, L2]";
    }
    double area2 = 0.0;
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
        float *s = &T.seg[6 * (size_t)k];
        if (seg_in) {
            for (int q = 0; q < 6; ++q) s[q] = (float)seg_in[6 * (size_t)k + q];
            continue;
        }
        const V2 &a = poly[k], &b = poly[(k + 1) % np];
        double ex = b.x - a.x, ey = b.y - a.y, L = std::hypot(ex, ey);
        if (!(L > 0.0)) return "degenerate body edge";
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
                if (geom::segment_hits_rect(a, b, x0, y0, x1, y1)) lists[cell_id(i, j, ny)].push_back(k);
            }
    }
    for (int c = 0; c < C; ++c) {
        T.cseg_start[c + 1] = T.cseg_start[c] + (int)lists[c].size();
        for (int k : lists[c]) T.cseg_idx.push_back(k);
    }

    // Fluid fraction of every cell (the three wall-vertex cases of
    // CartesianMesh.cpp:112-168, computed exactly by polygon clipping), or
    // the host mesh's own; then volumes and the fill helpers of cut cells.
    if (frac_in) {
        for (int c = 0; c < C; ++c) {
            const double fr = frac_in[c];
            if (!(fr >= 0.0 && fr <= 1.0)) return "host tables: fluid fractions must lie in [0, 1]";
            const uint8_t ct = cut_in ? cut_in[c] : (uint8_t)(fr > 0.0 && fr < 1.0);
            if (ct > 1 || (ct == 1) != (fr > 0.0 && fr < 1.0))
                return "host tables: cut flags must mark exactly the cells with 0 < fraction < 1";
            T.frac[c] = fr;
        }
    } else {
        geom::cell_fluid_fractions(poly, nx, ny, P.dx, P.dy, T.frac);
    }
    for (int i = 0; i < nx; ++i)
        for (int j = 0; j < ny; ++j) {
            const int c = cell_id(i, j, ny);
            const double fr = T.frac[c];
            if (fr == 1.0) continue;    // uncut
            T.vol[c] = fr * P.Vcell;
            T.vol_ntc[c] = (float)std::max(T.vol[c], 0.1 * P.Vcell);
            T.cut[c] = fr > 0.0 ? 1 : 0;
            if (fr > 0.0) {
                ++T.ncut;
                geom::fluid_bbox(poly, i * P.dx, j * P.dy, (i + 1) * P.dx, (j + 1) * P.dy, &T.fbox[4 * (size_t)c]);
                T.fmask[c] = geom::fluid_mask(poly, &T.fbox[4 * (size_t)c]);
            } else {
                T.vol_ntc[c] = 0.0f;
            }
        }
    return "";
}

void empty_tables(const Param &P, CutTables &T)
{
    const int C = P.in.N1 * P.in.N2;
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
}

}  // namespace

std::string build_cut_tables(const Param &P, const dsmc_body *body, CutTables &T)
{
    empty_tables(P, T);
    if (!body || body->n_points == 0) return "";
    if (body->n_points < 3 || !body->xy) return "body polygon needs >= 3 points";
    std::vector<V2> poly(body->n_points);
    for (int k = 0; k < body->n_points; ++k) poly[k] = V2{ body->xy[2 * k], body->xy[2 * k + 1] };
    return tables_from_outline(P, poly, nullptr, nullptr, nullptr, T);
}

std::string build_cut_tables_from_host(const Param &P, const dsmc_cut_tables *h, CutTables &T)
{
    empty_tables(P, T);
    if (!h) return "null host tables";
    if (h->nseg == 0) {
        if (h->frac)
            for (int c = 0; c < T.ncell; ++c)
                if (h->frac[c] != 1.0) return "host tables: without wall segments every fluid fraction is 1";
        return "";
    }
    if (h->nseg < 3 || !h->seg) return "host tables: a wall outline needs >= 3 segments";
    // the segments must form one closed outline in order, each with the unit
    // normal (ey, -ex)/|e| into the fluid
    const int n = h->nseg;
    const double L = std::max(P.in.L1, P.in.L2);
    std::vector<V2> poly(n);
    for (int k = 0; k < n; ++k) {
        const double *s = h->seg + 6 * (size_t)k, *t = h->seg + 6 * (size_t)((k + 1) % n);
        for (int q = 0; q < 6; ++q)
            if (!std::isfinite(s[q])) return "host tables: non-finite segment value";
        if (std::hypot(s[2] - t[0], s[3] - t[1]) > 1e-9 * L)
            return "host tables: segment " + std::to_string(k) + " does not end where the next one starts";
        const double ex = s[2] - s[0], ey = s[3] - s[1], len = std::hypot(ex, ey);
        if (!(len > 0.0)) return "host tables: degenerate segment " + std::to_string(k);
        if (std::fabs(s[4] - ey / len) > 1e-6 || std::fabs(s[5] + ex / len) > 1e-6)
            return "host tables: segment " + std::to_string(k) + " needs the unit normal (ey, -ex)/|e| into the fluid";
        poly[k] = V2{ s[0], s[1] };
    }
    return tables_from_outline(P, poly, h->seg, h->frac, h->cut, T);
}

}  // namespace dsmc
