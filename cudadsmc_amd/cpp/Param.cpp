#include "Param.h"

#include <cmath>
#include <cstring>

namespace dsmc_host {

double hard_sphere_density_for_mfp(double lam, double diam)
{
    return 1.0 / (std::sqrt(2.0) * M_PI * diam * diam * lam);
}

Param cylinder_mach10(int nx, int ny, double ppc)
{
    Param p;
    p.N1 = nx;
    p.N2 = ny;
    p.ppc = ppc;
    p.T = 200.0;
    p.T_wall = 500.0;
    const double a = std::sqrt(5.0 / 3.0 * p.boltz * p.T / p.mass);
    p.V_jet = 10.0 * a;
    p.L1 = 0.6;
    p.L2 = 0.3;
    p.Radius = 0.1;
    p.Center_x = 0.5 * p.L1;
    p.Center_y = 0.5 * p.L2;
    p.n_inf = hard_sphere_density_for_mfp(0.01 * 2.0 * p.Radius, p.diam);
    const double dx = p.L1 / nx;
    const double vmp = std::sqrt(2.0 * p.boltz * p.T / p.mass);
    p.tau = 0.3 * dx / (p.V_jet + 3.0 * vmp);
    p.steps = 20000;
    p.sample_start = 10000;
    return p;
}

dsmc_params to_abi(const Param &p)
{
    dsmc_params d;
    std::memset(&d, 0, sizeof(d));
    d.boltz = p.boltz;
    d.mass = p.mass;
    d.diam = p.diam;
    d.Vtl = p.Vtl;
    d.Tref = p.Tref;
    d.L1 = p.L1;
    d.L2 = p.L2;
    d.L3 = p.L3;
    d.N1 = p.N1;
    d.N2 = p.N2;
    d.n_inf = p.n_inf;
    d.T = p.T;
    d.V_jet = p.V_jet;
    d.V_y = p.V_y;
    d.T_wall = p.T_wall;
    d.wall_bc = p.wall_bc;
    for (int f = 0; f < 4; ++f) {
        d.face_bc[f] = p.face_bc[f];
        d.face_T[f] = p.face_T[f];
    }
    d.tau = p.tau;
    d.ppc = p.ppc;
    d.Fn = 0.0;
    d.seed = p.seed;
    d.capacity = p.capacity;
    d.sample_start = p.sample_start;
    d.zeta = p.zeta;
    d.rot_prob = p.rot_prob;
    return d;
}

}  // namespace dsmc_host
