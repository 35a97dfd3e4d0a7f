// Output.h -- field output of the C++ host layer.
//
// Role of the reference's io/Output (io/Output.h, Output.cpp:12-154):
// Write2HDF5 writes the datasets U, V, W, P, T, Rho, Numchild,
// Numgrandchild and MFP ([N1][N2], index i*N2 + j), Write2VTK a structured
// grid.  No HDF5 or VTK library is needed here: Write2HDF5 writes the same
// datasets, by the same names, as a NumPy .npz archive (numpy.load reads
// it), and Write2VTK writes a VTK XML structured grid (.vts, cell data) that
// ParaView opens.  Both match cudadsmc_amd/output.py value for value.
#pragma once

#include <string>
#include <vector>

namespace dsmc_host {

class Run;

// The datasets of one output, [N1][N2] row-major.
struct FieldSet {
    int N1 = 0, N2 = 0;
    double L1 = 0, L2 = 0;
    std::vector<double> U, V, W, P, T, Rho, MFP;
    std::vector<int> Numchild, Numgrandchild;   // adaptive sub-cells (0: no grid)
    std::vector<double> Trot;                   // rotational temperature (diatomic gas only, else empty)
};

class Output {
public:
    Output() = default;
    explicit Output(Run *run) : m_run(run) {}

    // filename without extension; writes filename.npz
    void Write2HDF5(const std::string &filename);
    // filename without extension; writes filename.vts
    void Write2VTK(const std::string &filename);

    static void write_npz(const std::string &path, const FieldSet &f);
    static void write_vts(const std::string &path, const FieldSet &f);

protected:
    Run *m_run = nullptr;
};

}  // namespace dsmc_host
