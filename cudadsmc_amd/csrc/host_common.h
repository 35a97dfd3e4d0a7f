// Host helpers shared by the single-case handle (dsmc_abi.hip) and the
// batched handle (dsmc_batch.hip): kernel constants of a case, the default
// particle capacity and the macroscopic fields from the sample sums.
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <string>

#include "dsmc_kernels.h"
#include "host/case_setup.h"

namespace dsmc {

// Message of the last failed call on this thread (dsmc_last_error).
std::string &last_error_slot();

inline int fail(int code, const std::string &msg)
{
    last_error_slot() = msg;
    return code;
}

#define HIPCHK(expr)                                                                              \
    do {                                                                                          \
        hipError_t e_ = (expr);                                                                   \
        if (e_ != hipSuccess)                                                                     \
            return ::dsmc::fail(DSMC_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

template <class T>
hipError_t dalloc(T **p, size_t n)
{
    return hipMalloc((void **)p, sizeof(T) * std::max<size_t>(n, 1));
}

// Validates launch settings (dsmc_set_launch): grids of 1..65536 blocks,
// k_cell tiles of CELL_P_MIN..CELL_CAP particles, 1..64 steps per graph.
inline int check_launch(const dsmc_launch *c)
{
    auto ok = [](int v, int lo, int hi) { return v >= lo && v <= hi; };
    if (!ok(c->move_blocks, 1, 65536) || !ok(c->scatter_blocks, 1, 65536) || !ok(c->cell_blocks, 1, 65536))
        return fail(DSMC_ERR_ARG, "launch: block counts must be 1..65536");
    if (!ok(c->cell_tile, CELL_P_MIN, CELL_CAP))
        return fail(DSMC_ERR_ARG, "launch: cell_tile must be " + std::to_string(CELL_P_MIN) + ".." +
                                      std::to_string(CELL_CAP));
    if (!ok(c->graph_steps, 1, 64)) return fail(DSMC_ERR_ARG, "launch: graph_steps must be 1..64");
    return DSMC_OK;
}

// Particle capacity: the requested one, else 4x the fill estimate.
inline int64_t default_capacity(const Param &P, const CutTables &T)
{
    if (P.in.capacity > 0) return P.in.capacity;
    double fill_est = 0;
    for (int c = 0; c < T.ncell; ++c) fill_est += P.ppc_eff * T.frac[c];
    return (int64_t)std::ceil(4.0 * fill_est) + 65536;
}

// Scalar constants of the kernels (device table pointers are set by the caller).
inline DevCase make_devcase(const Param &P, const CutTables &T, int64_t cap)
{
    const dsmc_params *p = &P.in;
    DevCase cs{};
    cs.nx = p->N1; cs.ny = p->N2; cs.ncell = T.ncell; cs.nseg = T.nseg;
    cs.L1 = (float)p->L1; cs.L2 = (float)p->L2; cs.dx = (float)P.dx; cs.dy = (float)P.dy;
    cs.dt = (float)p->tau;
    cs.ntc_coef = (float)P.ntc_coef; cs.K1 = (float)P.K1; cs.e_exp = (float)P.e_exp; cs.margin = T.margin;
    cs.body_bc = p->wall_bc; cs.body_vmp = (float)P.body_vmp; cs.body_sig = (float)P.body_sig;
    for (int f = 0; f < 4; ++f) {
        cs.face_bc[f] = p->face_bc[f];
        cs.face_vmp[f] = (float)P.face_vmp[f];
        cs.face_sig[f] = (float)P.face_sig[f];
        cs.inj_e[f] = P.inj_e[f];
        cs.inj_a[f] = (float)P.inj_a[f];
    }
    cs.inj_vmp = (float)P.inj_vmp; cs.inj_sig = (float)P.inj_sig;
    cs.inj_ux = (float)p->V_jet; cs.inj_uy = (float)p->V_y;
    cs.seed = P.seed32; cs.cap = (uint32_t)cap;
    cs.zeta = p->zeta;
    if (p->zeta == 2) {
        cs.rot_p = (float)p->rot_prob; cs.lb_ia = (float)P.lb_ia; cs.lb_ib = (float)P.lb_ib;
        cs.body_ekt = (float)P.body_ekt; cs.inj_ekt = (float)P.inj_ekt;
        for (int f = 0; f < 4; ++f) cs.face_ekt[f] = (float)P.face_ekt[f];
    }
    // Box outside which no cell lists a segment: the segments' bounding box
    // grown by the list margin (case_setup.cpp), two cells for the distance
    // between a particle and its cell's far side, and a float safety band.
    cs.bx0 = cs.by0 = 1.0f;
    cs.bx1 = cs.by1 = -1.0f;
    if (T.nseg > 0) {
        double x0 = 1e300, y0 = 1e300, x1 = -1e300, y1 = -1e300;
        for (int k = 0; k < T.nseg; ++k)
            for (int e = 0; e < 2; ++e) {
                const double sx = T.seg[6 * k + 2 * e], sy = T.seg[6 * k + 2 * e + 1];
                x0 = std::min(x0, sx); x1 = std::max(x1, sx);
                y0 = std::min(y0, sy); y1 = std::max(y1, sy);
            }
        const double L = std::max(p->L1, p->L2);
        const double g = 2.0 * std::max(P.dx, P.dy) + 1.001 * (double)T.margin + 1e-9 * L + 1e-5 * L;
        cs.bx0 = (float)(x0 - g); cs.bx1 = (float)(x1 + g);
        cs.by0 = (float)(y0 - g); cs.by1 = (float)(y1 + g);
    }
    return cs;
}

// Fields of the C cells of one case from its sample sums: acc holds six
// blocks (sum 1, sum vx, sum vy, sum vz, sum v^2, collisions), block k of
// cell c at acc[k * stride + c].  Number density on the cut-cell fluid volume.
inline void fields_from_sums(const Param &P, const CutTables &T, const double *acc, size_t stride, uint64_t ns,
                             double *n, double *ux, double *uy, double *uz, double *Tt, double *p)
{
    const double k = P.in.boltz, m = P.in.mass, Fn = P.Fn;
    for (int c = 0; c < T.ncell; ++c) {
        const double sn = acc[c];
        const double vol = T.vol[c];
        const double nn = (ns > 0 && vol > 0) ? sn / (double)ns * Fn / vol : 0.0;
        double u = 0, v = 0, w = 0, t = 0;
        if (sn > 0) {
            u = acc[stride + c] / sn; v = acc[2 * stride + c] / sn; w = acc[3 * stride + c] / sn;
            t = m / (3.0 * k) * (acc[4 * stride + c] / sn - (u * u + v * v + w * w));
        }
        if (n) n[c] = nn;
        if (ux) ux[c] = u;
        if (uy) uy[c] = v;
        if (uz) uz[c] = w;
        if (Tt) Tt[c] = t;
        if (p) p[c] = nn * k * t;
    }
}

// Rotational temperature of the C cells from the seventh sample block (sum of
// the rotational energy per unit mass): T_rot = 2 m <e> / (zeta k).
inline void rot_temperature_from_sums(const Param &P, const CutTables &T, const double *acc, size_t stride,
                                      double *Tr)
{
    const double k = P.in.boltz, m = P.in.mass, z = (double)P.in.zeta;
    for (int c = 0; c < T.ncell; ++c) {
        const double sn = acc[c];
        Tr[c] = sn > 0 ? 2.0 * m * (acc[6 * stride + c] / sn) / (z * k) : 0.0;
    }
}

}  // namespace dsmc
