// Host-side case construction: the roles of the reference's src/Param.h,
// src/meshes/CartesianMesh.{h,cpp} and src/object/Geom.{h,cpp}.
//
//  - Param:          derived constants of Run::initialize (Run.cpp:29-35:
//                    Vstd, Vmax, VHS_coe) plus the inflow flux per face;
//  - CartesianMesh:  uniform grid, cell id = k + j*nz + i*ny*nz with nz = 1
//                    (CartesianMesh.cpp:171-181), cell fluid volume
//                    (CartesianMesh.cpp:100-169);
//  - Geom:           closed body polygon, one wall segment per polygon edge,
//                    and per-cell lists of the segments a particle starting
//                    in that cell can reach within one cell width
//                    (Geom.cpp:7-72 builds one chord per cut cell instead).
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "../../../include/dsmc_abi.h"

namespace dsmc {

struct Param {
    dsmc_params in{};
    // derived (double, as in Run::initialize)
    double dx = 0, dy = 0, Vcell = 0, Fn = 0, ppc_eff = 0;
    double Vstd = 0, Vmax = 0, VHS_coe = 0;
    double K1 = 0, e_exp = 0, ntc_coef = 0, sigcmax0 = 0;
    double body_vmp = 0, body_sig = 0;
    double face_vmp[4]{}, face_sig[4]{};
    double inj_e[4]{}, inj_a[4]{};
    double inj_vmp = 0, inj_sig = 0;
    // rotational energy (zeta = 2): Larsen-Borgnakke exponents 1/(5/2-omega),
    // 1/(7/2-omega) and kT/m of the diffuse walls and the freestream
    double lb_ia = 0, lb_ib = 0;
    double body_ekt = 0, face_ekt[4]{}, inj_ekt = 0;
    int sample_blocks = 6;
    uint32_t seed32 = 0;

    // Validates the input and computes the derived constants.
    // Returns an empty string on success, else the reason.
    std::string derive(const dsmc_params &p);
};

struct CutTables {
    int ncell = 0, nseg = 0, ncut = 0;
    std::vector<float> seg;          // 6 per segment: ax ay bx by nx ny
    std::vector<int> cseg_start;     // ncell + 1
    std::vector<int> cseg_idx;
    std::vector<double> vol;         // fluid volume (dx*dy*L3 minus body), m^3
    std::vector<double> frac;        // fluid fraction, exactly 1.0 for uncut cells
    std::vector<float> vol_ntc;      // volume used by NTC (floored at 0.1 cell)
    std::vector<uint8_t> cut;        // 0 < frac < 1
    std::vector<float> fbox;         // 4 per cell: bounding box of the fluid part of a cut cell
    std::vector<uint64_t> fmask;     // per cut cell: bit sx + 8 sy set when sub-box (sx, sy) of the
                                     // 8x8 split of fbox touches fluid (all bits for uncut cells)
    float margin = 0;
};

// Builds the cut-cell tables.  body may be null (no immersed body).
std::string build_cut_tables(const Param &P, const dsmc_body *body, CutTables &T);
// The same from tables a host mesh layer built (dsmc_create_from_tables):
// its wall segments, fluid fractions and cut flags are taken as given (and
// checked); the per-cell reach lists and fill helpers are derived.
std::string build_cut_tables_from_host(const Param &P, const dsmc_cut_tables *h, CutTables &T);

// Cell id of cell (i, j): the reference convention with nz = 1.
inline int cell_id(int i, int j, int ny) { return j + i * ny; }

}  // namespace dsmc
