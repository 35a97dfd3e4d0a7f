// Run.h -- the C++ host driver of one DSMC case over the C ABI.
//
// Keeps the role and the method names of the reference's Run
// (src/Run.h:22-63, Run.cpp): initialize(argc, argv) sets up the mesh,
// body and particles, solver() runs the time loop with output,
// finalize() releases everything; particlemove(), ressignParticle() and
// collision() are the three phases of a step (Run.cpp:159-162, 245-256),
// assignParticle() the initial fill (Run.cpp:108-157) and
// TransferCellsFromDeviceToHost() brings the sampled fields back for
// Output.  The device state lives behind a dsmc_handle (include/dsmc_abi.h):
// the host mesh (CartesianMesh: getIndex, cutcell) builds the cut-cell
// tables and TransferCellsFromHostToDevice() hands them over
// (dsmc_create_from_tables); the particle fill happens inside
// dsmc_init_freestream; Sample() is GPUCells::Sample.  Every ABI status is
// checked: a failure throws std::runtime_error with dsmc_last_error().
#pragma once

#include <memory>
#include <string>
#include <vector>

#include "../../include/dsmc_abi.h"
#include "CartesianMesh.h"
#include "Output.h"
#include "Param.h"

namespace dsmc_host {

class Run {
public:
    Run() = default;
    ~Run();
    Run(const Run &) = delete;
    Run &operator=(const Run &) = delete;

    /*============ Initialize ============*/
    // Options (all optional; the default is the Mach-10 cylinder benchmark):
    //   --case cylinder|wedge|none  --mach M --kn KN --tw RATIO --half-angle DEG
    //   --nx N1 --ny N2 --ppc P --seed S --steps K --sample-from STEP
    //   --sample-every K --amr K --output-every K --out DIR --eager --phases
    void initialize(int argc, char **argv);
    void initialize(const Param &p);
    /*============ Main loop ============*/
    void solver();
    /*============ Finalize ============*/
    void finalize();

    void assignParticle(const double &coef);
    void particlemove();
    void ressignParticle();
    void collision();
    void Sample();
    void genAMRmesh();
    void TransferCellsFromHostToDevice();
    void TransferCellsFromDeviceToHost();
    const CartesianMesh &mesh() const { return *m_mesh; }

    const Param &param() const { return m_param; }
    const std::vector<double> &body() const { return m_body; }
    const FieldSet &fields() const { return m_fields; }
    dsmc_status status() const;
    // The case as JSON: every dsmc_params field and the body outline.
    void write_case(const std::string &path) const;

protected:
    Param m_param;
    std::vector<double> m_body;
    dsmc_handle m_h = nullptr;
    std::unique_ptr<CartesianMesh> m_mesh;
    FieldSet m_fields;
    std::unique_ptr<Output> m_output;

    void write_output(int iter);

    friend class Output;
};

// Throws std::runtime_error with the library's message when rc is not DSMC_OK.
void check(int rc, const char *what);

// Parses the command line into a Param (throws std::invalid_argument).
Param parse_args(int argc, char **argv);

}  // namespace dsmc_host
