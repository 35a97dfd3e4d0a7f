// CartesianMesh.h -- the uniform grid of a case and its cut cells, built on
// the host: the role of the reference's src/meshes/CartesianMesh.{h,cpp}
// (getIndex, cutcell) together with the segment sorting of
// src/object/Geom.cpp:7-72.  Run::TransferCellsFromHostToDevice hands its
// tables to the library (dsmc_create_from_tables), as the reference's
// Run.cpp:197-223 hands ifCut and GPUSegment tables to
// GPUCells::UploadFromHost.
#pragma once

#include <cstdint>
#include <vector>

namespace dsmc_host {

class CartesianMesh {
public:
    CartesianMesh(int N1, int N2, double L1, double L2);

    // Cell id of a position: j + i*N2 with i = (int)(x/dx), j = (int)(y/dy),
    // clamped into the grid (CartesianMesh.cpp:171-181 with nz = 1).
    int getIndex(double x, double y) const;

    // Cuts every cell against the body outline (x, y pairs, counter-clockwise;
    // empty for no body): the fluid fraction and the ifCut flag of every cell
    // (CartesianMesh.cpp:100-169, here by exact polygon clipping) and one wall
    // segment per outline edge, end points and unit normal into the fluid.
    void cutcell(const std::vector<double> &outline);

    int numberCells() const { return m_N1 * m_N2; }
    double getUnidX() const { return m_dx; }
    double getUnidY() const { return m_dy; }
    const std::vector<double> &fractions() const { return m_frac; }
    const std::vector<uint8_t> &ifcut() const { return m_cut; }
    // 6 per segment: ax ay bx by nx ny
    const std::vector<double> &segments() const { return m_seg; }

private:
    int m_N1, m_N2;
    double m_L1, m_L2, m_dx, m_dy;
    std::vector<double> m_frac;
    std::vector<uint8_t> m_cut;
    std::vector<double> m_seg;
};

}  // namespace dsmc_host
