// Param.h -- case parameters of the C++ host layer.
//
// Role of the reference's src/Param.h:5-43: the physical constants (boltz,
// mass, diam, Vtl), the flow field (L1, L2, L3, T, T_wall, V_jet), the body
// (Radius, Center_x, Center_y), the grid (N1, N2, N3 = 1), the time step tau
// and the output cadence.  The reference fixes them at compile time; here
// they are runtime values (one struct per case), so one executable runs any
// case, and they travel to the device through dsmc_params (include/dsmc_abi.h).
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/dsmc_abi.h"

namespace dsmc_host {

struct Param {
    // physical constants (Param.h:5-13); argon, VHS (Bird 1994, Appendix A)
    double boltz = 1.380649e-23;   // J/K
    double mass = 6.63e-26;        // kg
    double diam = 4.17e-10;        // m, VHS reference diameter
    double Vtl = 0.81;             // viscosity index omega
    double Tref = 273.0;           // K, VHS reference temperature
    int zeta = 0;                  // rotational degrees of freedom (Param.h:14): 0, or 2 for N2
    double rot_prob = 1.0;         // 1/Z_rot: fraction of collisions exchanging rotational energy
    double tau = 0.0;              // time step (s)
    // flow field (Param.h:16-22)
    double L1 = 0.6, L2 = 0.3, L3 = 1.0;
    double n_inf = 0.0;            // freestream number density (the reference gives Rho)
    double T = 200.0;              // freestream temperature (K)
    double T_wall = 500.0;         // body wall temperature (K)
    double V_jet = 0.0;            // freestream x velocity (m/s)
    double V_y = 0.0;
    int wall_bc = DSMC_BC_DIFFUSE;
    int face_bc[4] = { DSMC_BC_STREAM, DSMC_BC_STREAM, DSMC_BC_STREAM, DSMC_BC_STREAM };
    double face_T[4] = { 0.0, 0.0, 0.0, 0.0 };
    // body (Param.h:24-28): a cylinder of Radius at (Center_x, Center_y) drawn
    // as an n_points polygon (Run.cpp:45), or a wedge
    std::string geometry = "cylinder";   // "cylinder", "wedge" or "none"
    double Radius = 0.1, Center_x = 0.3, Center_y = 0.15;
    int n_points = 256;
    double wedge_apex_x = 0.15, wedge_length = 0.3, wedge_half_angle_deg = 20.0;
    // grid (Param.h:30-32)
    int N1 = 400, N2 = 200;
    // particles: per full fluid cell at n_inf (the reference gives N_Particle)
    double ppc = 30.0;
    uint64_t seed = 20251203;
    int64_t capacity = 0;
    // run control and output (Param.h:40-43; Run::solver, Run.cpp:258-304)
    int steps = 20000;
    int sample_start = 10000;      // first sampled step (< 0: never)
    int sample_every = 0;          // > 0: GPUCells::Sample after every this many steps instead
                                   // (Run.cpp:294-296 samples after steps 0, 100, 200, ...)
    int amr_every = 0;             // > 0: adaptive collision sub-cells (Cell::genAMRmesh) set from
                                   // the sampling sums after every this many steps once sampled
    int output_every = 0;          // write result_<iter> every this many steps (0: at the end only)
    std::string out_dir = "./res"; // the reference writes ./res/result_<iter>
    bool graph = true;             // replay captured steps; false: eager phase launches
    bool phases = false;           // one call per phase per step, as Run::solver does
};

// The Mach-10 cylinder (the benchmark case): T_inf = 200 K, Kn = 0.01 on the
// diameter with the hard-sphere mean free path, R = 0.1 m centred in a
// 0.6 m x 0.3 m domain, diffuse wall at 500 K, freestream on all faces,
// dt = 0.3 dx / (U + 3 vmp).
Param cylinder_mach10(int nx = 400, int ny = 200, double ppc = 30.0);

// The device-side parameter block.
dsmc_params to_abi(const Param &p);

// Hard-sphere number density whose mean free path 1/(sqrt(2) pi d^2 n) is lam.
double hard_sphere_density_for_mfp(double lam, double diam);

}  // namespace dsmc_host
