// dsmc_run -- one DSMC case from the command line (the reference's
// src/main.cpp: Run run; run.initialize(argc, argv); run.solver();
// run.finalize()).  Options: see Run.h.  Exit status 0 on success, 1 on
// any error (message on stderr), 2 on a bad command line.
//
// `dsmc_run --write-fields DIR` writes a synthetic field set through Output
// (no device needed): a check of the .npz and .vts writers.
// `dsmc_run --write-mesh FILE [case options]` writes the host mesh's cut-cell
// tables (no device needed): a check against the library's own tables.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <stdexcept>

#include "CartesianMesh.h"
#include "Geom.h"
#include "Run.h"

namespace {

int write_fields(const char *dir)
{
    dsmc_host::FieldSet f;
    f.N1 = 7;
    f.N2 = 5;
    f.L1 = 0.7;
    f.L2 = 0.25;
    const size_t C = 35;
    f.U.resize(C); f.V.resize(C); f.W.resize(C); f.P.resize(C); f.T.resize(C); f.Rho.resize(C); f.MFP.resize(C);
    f.Numchild.assign(C, 0);
    f.Numgrandchild.assign(C, 0);
    for (size_t c = 0; c < C; ++c) {
        const double x = (double)c;
        f.U[c] = 1000.0 * std::sin(x);
        f.V[c] = -3.25e-3 * x;
        f.W[c] = c == 3 ? -0.0 : 1.0 / (x + 1.0);
        f.P[c] = 1e-5 * std::exp(x / 7.0);
        f.T[c] = c == 4 ? NAN : 200.0 + x;
        f.Rho[c] = c == 5 ? 0.0 : 6.63e-26 * 1e20 * (x + 0.5);
        f.MFP[c] = c == 6 ? INFINITY : 1e-3 / (x + 1.0);
        f.Numchild[c] = (int)c - 17;
    }
    const std::string base = std::string(dir) + "/fields";
    dsmc_host::Output::write_npz(base + ".npz", f);
    dsmc_host::Output::write_vts(base + ".vts", f);
    return 0;
}

// The host mesh of a case, cut against its body, as one binary file: int32
// N1, N2, nseg, then N1*N2 float64 fluid fractions, N1*N2 uint8 cut flags
// and 6*nseg float64 segment values (no device needed).
int write_mesh(const char *path, int argc, char **argv)
{
    const dsmc_host::Param p = dsmc_host::parse_args(argc, argv);
    dsmc_host::CartesianMesh m(p.N1, p.N2, p.L1, p.L2);
    m.cutcell(dsmc_host::body_of(p));
    FILE *f = std::fopen(path, "wb");
    if (!f) throw std::runtime_error(std::string("cannot write ") + path);
    const int32_t hdr[3] = { p.N1, p.N2, (int32_t)(m.segments().size() / 6) };
    bool ok = std::fwrite(hdr, sizeof(hdr), 1, f) == 1;
    ok = ok && std::fwrite(m.fractions().data(), sizeof(double), m.fractions().size(), f) == m.fractions().size();
    ok = ok && std::fwrite(m.ifcut().data(), 1, m.ifcut().size(), f) == m.ifcut().size();
    ok = ok && std::fwrite(m.segments().data(), sizeof(double), m.segments().size(), f) == m.segments().size();
    ok = (std::fclose(f) == 0) && ok;
    if (!ok) throw std::runtime_error(std::string("write failed: ") + path);
    return 0;
}

}  // namespace

int main(int argc, char **argv)
{
    if (argc > 1 && (!std::strcmp(argv[1], "--help") || !std::strcmp(argv[1], "-h"))) {
        std::printf("usage: dsmc_run [--case cylinder|wedge|none] [--mach M] [--kn KN] [--tw RATIO]\n"
                    "                [--half-angle DEG] [--nx N1] [--ny N2] [--ppc P] [--seed S] [--steps K]\n"
                    "                [--sample-from STEP] [--sample-every K] [--output-every K] [--out DIR]\n"
                    "                [--gas argon|nitrogen] [--zrot Z] [--eager] [--phases]\n");
        return 0;
    }
    if (argc >= 3 && !std::strcmp(argv[1], "--write-mesh")) {
        try {
            return write_mesh(argv[2], argc - 2, argv + 2);
        } catch (const std::exception &e) {
            std::fprintf(stderr, "dsmc_run: %s\n", e.what());
            return 1;
        }
    }
    if (argc == 3 && !std::strcmp(argv[1], "--write-fields")) {
        try {
            return write_fields(argv[2]);
        } catch (const std::exception &e) {
            std::fprintf(stderr, "dsmc_run: %s\n", e.what());
            return 1;
        }
    }
    dsmc_host::Param p;
    try {
        p = dsmc_host::parse_args(argc, argv);
    } catch (const std::invalid_argument &e) {
        std::fprintf(stderr, "dsmc_run: %s (see --help)\n", e.what());
        return 2;
    }
    try {
        dsmc_host::Run run;
        run.initialize(p);
        run.solver();
        run.finalize();
    } catch (const std::exception &e) {
        std::fprintf(stderr, "dsmc_run: %s\n", e.what());
        return 1;
    }
    return 0;
}
