// Kernel launchers and tile constants shared by dsmc_kernels.hip and the
// C ABI implementation.
#pragma once

#include "dsmc_device.h"

namespace dsmc {

constexpr int SCAN_IPT = 8;                  // cells per thread in k_scan
constexpr int SCAN_TILE = 256 * SCAN_IPT;    // cells per k_scan block
constexpr int CELL_P = 1024;                 // particles per k_cell tile (default)
constexpr int CELL_P_MIN = 256;              // smallest tile size the launch settings accept
constexpr int CELL_CAP = 2048;               // LDS staging capacity of one tile
constexpr int CELL_BIG = 64;                 // cells with > 64 particles listed per tile
constexpr int CELL_WAVES = 4;                // waves per k_cell block
constexpr int SAMPLE_G = 8;                  // lanes per cell in the FP64 sampling sums (k_cell, k_sample)

// The step's binning: k_move records runs into bn; launch_scan_runs scans
// the packed counts; launch_scatter runs k_place and k_rank (run bases in
// runlist, start tables, tile table).
hipError_t launch_move(const DevCase &cs, const SoA &A, const SoA &B, const Bins &bn, Ctrl *ctrl, int grid,
                       hipStream_t st);
// Many instances in one launch: cases[ninst] and cbase[ninst+1] in device
// memory, per-instance counters in ictl[ninst], totals in ctrl.
hipError_t launch_move_batch(const DevCase *cases, int ninst, const int *cbase, const SoA &A, const SoA &B,
                             const Bins &bn, Ctrl *ctrl, InstCtrl *ictl, int grid, hipStream_t st);
hipError_t launch_scan_runs(const Bins &bn, int ncell, Ctrl *ctrl, hipStream_t st);
hipError_t launch_scatter(const Bins &bn, int *start, int *tile_first, int ncell, int tile_p, const Ctrl *ctrl,
                          int grid, hipStream_t st);
// Plain exclusive scan of int counts (the fill and dsmc_bin_sort).
hipError_t launch_scan(int *count, int *lstart, int *tsum, int *toff, int ncell, Ctrl *ctrl,
                       hipStream_t st);
hipError_t launch_cell(const DevCase &cs, const SoA &A, const SoA &B, const int *start, int *perm,
                       const int4 *runlist, const int *tile_first, float *sigcmax, float *mrem, int *ncoll_cell,
                       double *acc, Ctrl *ctrl, int sample_start, int tile_p, int grid, hipStream_t st);
hipError_t launch_cell_batch(const DevCase &all, const CellConst *cc, const int *cbase, int ninst,
                             InstCtrl *ictl, const SoA &A, const SoA &B, const int *start, int *perm,
                             const int4 *runlist, const int *tile_first, float *sigcmax, float *mrem, int *ncoll_cell, double *acc,
                             Ctrl *ctrl, int tile_p, int grid, hipStream_t st);
// Adaptive collision sub-cells: after launch_cell, the cells with sub-cells
// (cs.sub) collide and sample in leaves; launch_amr_update sets the grids,
// mean free paths and VTS time steps from the sampling sums.
hipError_t launch_amr_cell(const DevCase &cs, const SoA &A, const SoA &B, const int *start, float *sigcmax,
                           float *mrem, int *ncoll_cell, double *acc, Ctrl *ctrl, int sample_start, int grid,
                           hipStream_t st);
hipError_t launch_amr_update(const DevCase &cs, double Fn, double mass, double diam, double T_inf, double omega,
                             double boltz, const double *acc, const double *vol, const Ctrl *ctrl, uint32_t *sub,
                             float *lam, double *dt, hipStream_t st);
int occupancy_amr_cell(bool rot);
hipError_t launch_fill(const DevCase &cs, const SoA &A, const double *expected, const uint8_t *cut,
                       const float *ux, const float *uy, const float *sig, int *count, int *lstart,
                       int *tsum, int *toff, Ctrl *ctrl, int grid, hipStream_t st);
hipError_t launch_ctrl_reset(Ctrl *ctrl, uint32_t n, uint32_t next_pid, hipStream_t st);
// One sample of the current (cell-ordered) state into acc, counted as a
// sampled step: the sums of k_cell's sampling, in its order.
hipError_t launch_sample(const DevCase &cs, const SoA &A, const SoA &B, const int *start, const int *ncoll_cell,
                         double *acc, Ctrl *ctrl, hipStream_t st);

// One instance of a wall-distance launch: its cells start at cbase in the
// output; cell c = j + i*ny has its centre at ((i + 0.5) dx, (j + 0.5) dy);
// its outline is poly[2*poly0 .. 2*(poly0 + npoly)) (x, y pairs).
struct WallInst {
    int cbase, ncell, ny, poly0, npoly;
    double dx, dy;
};
hipError_t launch_wall_distance(const WallInst *inst, int ninst, int max_cells, const double *poly, double *out,
                                hipStream_t st);

// Resident blocks per CU of the persistent kernels (occupancy query).
int occupancy_move();
int occupancy_move_batch();
int occupancy_cell(bool rot = false);
int occupancy_cell_batch();
int occupancy_scatter();

}  // namespace dsmc
