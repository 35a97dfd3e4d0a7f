#include "Run.h"

#include <sys/stat.h>

#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <stdexcept>

#include "CartesianMesh.h"
#include "Geom.h"

namespace dsmc_host {

void check(int rc, const char *what)
{
    if (rc != DSMC_OK) {
        const char *m = dsmc_last_error();
        throw std::runtime_error(std::string(what) + " failed (status " + std::to_string(rc) + "): " + (m ? m : ""));
    }
}

namespace {

double arg_d(const char *s, const char *name)
{
    char *end = nullptr;
    const double v = std::strtod(s, &end);
    if (!end || *end || !std::isfinite(v)) throw std::invalid_argument(std::string("bad value for ") + name);
    return v;
}

long arg_i(const char *s, const char *name)
{
    char *end = nullptr;
    const long v = std::strtol(s, &end, 10);
    if (!end || *end) throw std::invalid_argument(std::string("bad value for ") + name);
    return v;
}

// Supersonic argon over a cylinder (R = 0.1 m, centred) or a wedge (apex at
// x = 0.15 m, 0.3 m long or shorter so that its base stays 1 cm inside the
// domain) in a 0.6 m x 0.3 m domain: the case family the sweeps draw from
// (cudadsmc_amd/cases.py flow_case).
Param flow_case(const Param &gas, const std::string &geometry, double mach, double kn, double tw_ratio, int nx,
                int ny, double ppc, double half_angle_deg)
{
    Param p = cylinder_mach10(nx, ny, ppc);
    p.mass = gas.mass; p.diam = gas.diam; p.Vtl = gas.Vtl; p.Tref = gas.Tref;
    p.zeta = gas.zeta; p.rot_prob = gas.rot_prob;
    // speed of sound with the gas's ratio of specific heats (5/3, or 7/5 with
    // two rotational degrees of freedom)
    const double gamma = (5.0 + p.zeta) / (3.0 + p.zeta);
    const double a = std::sqrt(gamma * p.boltz * p.T / p.mass);
    p.V_jet = mach * a;
    p.T_wall = tw_ratio * p.T;
    p.geometry = geometry;
    p.wedge_half_angle_deg = half_angle_deg;
    // 0.3 m long unless its base would come closer than 1 cm to the y faces
    p.wedge_length = std::min(0.3, (0.5 * p.L2 - 0.01) / std::tan(half_angle_deg * M_PI / 180.0));
    const double size = geometry == "wedge" ? p.wedge_length : 2.0 * p.Radius;
    p.n_inf = hard_sphere_density_for_mfp(kn * size, p.diam);
    const double dx = p.L1 / nx;
    const double vmp = std::sqrt(2.0 * p.boltz * p.T / p.mass);
    p.tau = 0.3 * dx / (p.V_jet + 3.0 * vmp);
    return p;
}

void make_dir(const std::string &d)
{
    std::string cur;
    for (size_t i = 0; i <= d.size(); ++i) {
        if (i == d.size() || d[i] == '/') {
            if (!cur.empty()) mkdir(cur.c_str(), 0755);
        }
        if (i < d.size()) cur.push_back(d[i]);
    }
}

}  // namespace

Param parse_args(int argc, char **argv)
{
    std::string geometry = "cylinder";
    double mach = 10.0, kn = 0.01, tw = 2.5, half = 20.0, ppc = 30.0;
    int nx = 400, ny = 200;
    long steps = -1, sample_from = -2, every = 0, seed = -1, sample_every = 0, amr = 0;
    std::string out = "./res", gas = "argon";
    double zrot = 1.0;
    bool eager = false, phases = false;
    for (int k = 1; k < argc; ++k) {
        const std::string a = argv[k];
        auto val = [&]() -> const char * {
            if (k + 1 >= argc) throw std::invalid_argument("missing value for " + a);
            return argv[++k];
        };
        if (a == "--case") geometry = val();
        else if (a == "--mach") mach = arg_d(val(), "--mach");
        else if (a == "--kn") kn = arg_d(val(), "--kn");
        else if (a == "--tw") tw = arg_d(val(), "--tw");
        else if (a == "--half-angle") half = arg_d(val(), "--half-angle");
        else if (a == "--nx") nx = (int)arg_i(val(), "--nx");
        else if (a == "--ny") ny = (int)arg_i(val(), "--ny");
        else if (a == "--ppc") ppc = arg_d(val(), "--ppc");
        else if (a == "--seed") seed = arg_i(val(), "--seed");
        else if (a == "--steps") steps = arg_i(val(), "--steps");
        else if (a == "--sample-from") sample_from = arg_i(val(), "--sample-from");
        else if (a == "--output-every") every = arg_i(val(), "--output-every");
        else if (a == "--sample-every") sample_every = arg_i(val(), "--sample-every");
        else if (a == "--amr") amr = arg_i(val(), "--amr");
        else if (a == "--out") out = val();
        else if (a == "--eager") eager = true;
        else if (a == "--phases") phases = true;
        else if (a == "--gas") gas = val();
        else if (a == "--zrot") zrot = arg_d(val(), "--zrot");
        else throw std::invalid_argument("unknown option " + a);
    }
    if (nx < 1 || ny < 1) throw std::invalid_argument("--nx and --ny must be positive");
    if (gas != "argon" && gas != "nitrogen") throw std::invalid_argument("--gas must be argon or nitrogen");
    if (!(zrot >= 1.0)) throw std::invalid_argument("--zrot must be >= 1");
    if (geometry != "cylinder" && geometry != "wedge" && geometry != "none")
        throw std::invalid_argument("--case must be cylinder, wedge or none");
    Param base;
    if (gas == "nitrogen") {
        // N2, VHS (Bird 1994, Appendix A), two rotational degrees of freedom:
        // the reference's gas (Param.h:7,14)
        base.mass = 4.65e-26;
        base.Vtl = 0.74;
        base.zeta = 2;
        base.rot_prob = 1.0 / zrot;
    }
    Param p = flow_case(base, geometry == "none" ? "cylinder" : geometry, mach, kn, tw, nx, ny, ppc, half);
    p.geometry = geometry;
    if (steps >= 0) p.steps = (int)steps;
    if (sample_every < 0) throw std::invalid_argument("--sample-every must be >= 0");
    if (amr < 0) throw std::invalid_argument("--amr must be >= 0");
    p.amr_every = (int)amr;
    p.sample_every = (int)sample_every;
    if (sample_every > 0) p.sample_start = -1;   // the steps' own sampling off: Sample() instead
    if (sample_from >= -1) p.sample_start = (int)sample_from;
    if (seed >= 0) p.seed = (uint64_t)seed;
    p.output_every = (int)every;
    p.out_dir = out;
    p.graph = !eager;
    p.phases = phases;
    return p;
}

Run::~Run()
{
    if (m_h) dsmc_destroy(m_h);
}

void Run::initialize(int argc, char **argv)
{
    initialize(parse_args(argc, argv));
}

void Run::initialize(const Param &p)
{
    m_param = p;
    m_body = body_of(p);
    // mesh part (Run.cpp:37-62): the grid, then the cells cut against the body
    m_mesh = std::make_unique<CartesianMesh>(p.N1, p.N2, p.L1, p.L2);
    m_mesh->cutcell(m_body);
    TransferCellsFromHostToDevice();
    m_output = std::make_unique<Output>(this);
    make_dir(p.out_dir);
    write_case(p.out_dir + "/case.json");
    assignParticle(1.0);
    const dsmc_status s = status();
    std::printf("case %s %dx%d: %u particles, dt %.6g s, Fn from %.3g particles per cell\n", p.geometry.c_str(),
                p.N1, p.N2, s.n_live, p.tau, p.ppc);
}

void Run::assignParticle(const double &coef)
{
    // the library fills every fluid cell at n_inf; other fill fractions go
    // through dsmc_init_fields
    if (coef != 1.0) throw std::invalid_argument("assignParticle: only the full freestream fill (coef = 1)");
    check(dsmc_init_freestream(m_h), "dsmc_init_freestream");
}

void Run::TransferCellsFromHostToDevice()
{
    // the cut flags, fluid fractions and wall segments of the host mesh
    // (GPUCells::UploadFromHost's ifCut and GPUSegment tables, Run.cpp:197-223);
    // the handle that holds them on the device is created here
    const dsmc_params d = to_abi(m_param);
    const std::vector<double> &seg = m_mesh->segments();
    dsmc_cut_tables t{ (int)(seg.size() / 6), seg.empty() ? nullptr : seg.data(), m_mesh->fractions().data(),
                       m_mesh->ifcut().data() };
    if (m_h) {
        check(dsmc_destroy(m_h), "dsmc_destroy");
        m_h = nullptr;
    }
    check(dsmc_create_from_tables(&d, &t, &m_h), "dsmc_create_from_tables");
}

void Run::Sample() { check(dsmc_sample(m_h), "dsmc_sample"); }

// The sub-cell grids of every cell from its sampled state (Cell::VTS and
// Cell::genAMRmesh, Cell.cpp:117-279, which the reference leaves uncalled).
void Run::genAMRmesh() { check(dsmc_amr_update(m_h), "dsmc_amr_update"); }

void Run::particlemove() { check(dsmc_move(m_h), "dsmc_move"); }
void Run::ressignParticle() { check(dsmc_reassign(m_h), "dsmc_reassign"); }
void Run::collision() { check(dsmc_collide(m_h), "dsmc_collide"); }

dsmc_status Run::status() const
{
    dsmc_status s;
    check(dsmc_status_get(m_h, &s), "dsmc_status_get");
    return s;
}

void Run::TransferCellsFromDeviceToHost()
{
    const int n1 = m_param.N1, n2 = m_param.N2;
    const size_t C = (size_t)n1 * n2;
    std::vector<double> n(C), ux(C), uy(C), uz(C), T(C), p(C);
    check(dsmc_download_fields(m_h, n.data(), ux.data(), uy.data(), uz.data(), T.data(), p.data()),
          "dsmc_download_fields");
    FieldSet &f = m_fields;
    f.N1 = n1;
    f.N2 = n2;
    f.L1 = m_param.L1;
    f.L2 = m_param.L2;
    f.U = ux;
    f.V = uy;
    f.W = uz;
    f.P = p;
    f.T = T;
    f.Rho.assign(C, 0.0);
    f.MFP.assign(C, 0.0);
    f.Numchild.assign(C, 0);
    f.Numgrandchild.assign(C, 0);
    {
        // adaptive sub-cells: level-2 cells and level-3 cells in all of them
        // (Output.cpp:45-63); zeros without grids
        std::vector<uint32_t> sub(C);
        check(dsmc_download_amr(m_h, sub.data(), nullptr, nullptr), "dsmc_download_amr");
        for (size_t c = 0; c < C; ++c) {
            const int n2 = (int)(sub[c] & 0xffu) * (int)((sub[c] >> 8) & 0xffu);
            const int n3 = (int)((sub[c] >> 16) & 0xffu) * (int)((sub[c] >> 24) & 0xffu);
            f.Numchild[c] = n2;
            f.Numgrandchild[c] = n2 * n3;
        }
    }
    f.Trot.clear();
    if (m_param.zeta) {
        f.Trot.resize(C);
        check(dsmc_download_rot_temperature(m_h, f.Trot.data()), "dsmc_download_rot_temperature");
    }
    // MFP as the reference's Cell.cpp:125, m / (sqrt(2) rho pi d^2) (T/T_inf)^(omega - 1/2)
    const double d2 = m_param.diam * m_param.diam;
    for (size_t c = 0; c < C; ++c) {
        const double rho = n[c] * m_param.mass;
        f.Rho[c] = rho;
        if (rho > 0) {
            const double Tc = T[c] > 0 ? T[c] : 1.0;
            f.MFP[c] = m_param.mass / (std::sqrt(2.0) * rho * M_PI * d2) * std::pow(Tc / m_param.T, m_param.Vtl - 0.5);
        }
    }
}

void Run::write_output(int iter)
{
    TransferCellsFromDeviceToHost();
    const std::string base = m_param.out_dir + "/result_" + std::to_string(iter);
    m_output->Write2HDF5(base);
    m_output->Write2VTK(base);
}

void Run::solver()
{
    const Param &p = m_param;
    using clk = std::chrono::steady_clock;
    const auto t0 = clk::now();
    const uint64_t ps0 = status().particle_steps;
    int iter = 0;
    while (iter < p.steps) {
        int next = p.steps;
        if (p.output_every > 0) next = std::min(p.steps, (iter / p.output_every + 1) * p.output_every);
        if (p.phases) {
            // one call per phase and step, timed as Run::solver prints them
            for (; iter < next; ++iter) {
                const auto a = clk::now();
                particlemove();
                check(dsmc_sync(m_h), "dsmc_sync");
                const auto b = clk::now();
                ressignParticle();
                check(dsmc_sync(m_h), "dsmc_sync");
                const auto c = clk::now();
                collision();
                check(dsmc_sync(m_h), "dsmc_sync");
                const auto d = clk::now();
                if (iter % 100 == 0) {
                    auto ms = [](clk::time_point x, clk::time_point y) {
                        return std::chrono::duration<double, std::milli>(y - x).count();
                    };
                    std::printf("step %d: move %.3f ms, reassign %.3f ms, collision %.3f ms, total %.3f ms\n", iter,
                                ms(a, b), ms(b, c), ms(c, d), ms(a, d));
                }
            }
        } else if (p.sample_every > 0) {
            // GPUCells::Sample after every sample_every-th step, the first
            // after step 0 (Run.cpp:294-296)
            const int se = p.sample_every;
            while (iter < next) {
                const int s = (iter + se - 1) / se * se;     // next sampled step index >= iter
                const int target = s < next ? s + 1 : next;
                check(dsmc_run(m_h, target - iter, p.graph ? 1 : 0), "dsmc_run");
                iter = target;
                if ((iter - 1) % se == 0) Sample();
            }
        } else if (p.amr_every > 0) {
            // the grids renewed after every amr_every-th step, once sampled
            while (iter < next) {
                const int target = std::min(next, (iter / p.amr_every + 1) * p.amr_every);
                check(dsmc_run(m_h, target - iter, p.graph ? 1 : 0), "dsmc_run");
                iter = target;
                if (iter % p.amr_every == 0 && status().nsamples > 0) genAMRmesh();
            }
        } else {
            check(dsmc_run(m_h, next - iter, p.graph ? 1 : 0), "dsmc_run");
            iter = next;
        }
        check(dsmc_sync(m_h), "dsmc_sync");
        const dsmc_status s = status();
        const double sec = std::chrono::duration<double>(clk::now() - t0).count();
        std::printf("step %d: %u particles, %llu collisions, %.4g particle-steps/s%s\n", iter, s.n_live,
                    (unsigned long long)s.collisions, (double)(s.particle_steps - ps0) / sec,
                    s.err ? " (device error bits set)" : "");
        if (s.err) throw std::runtime_error("device error bits " + std::to_string(s.err));
        if (p.output_every > 0 || iter == p.steps) write_output(iter);
    }
}

void Run::finalize()
{
    if (m_h) {
        check(dsmc_destroy(m_h), "dsmc_destroy");
        m_h = nullptr;
    }
    std::printf("Simulation Finished\n");
}

void Run::write_case(const std::string &path) const
{
    const dsmc_params d = to_abi(m_param);
    FILE *f = std::fopen(path.c_str(), "w");
    if (!f) throw std::runtime_error("cannot write " + path);
    std::fprintf(f, "{\"params\": {\"boltz\": %.17g, \"mass\": %.17g, \"diam\": %.17g, \"Vtl\": %.17g, \"Tref\": %.17g,\n",
                 d.boltz, d.mass, d.diam, d.Vtl, d.Tref);
    std::fprintf(f, " \"L1\": %.17g, \"L2\": %.17g, \"L3\": %.17g, \"N1\": %d, \"N2\": %d, \"n_inf\": %.17g, \"T\": %.17g,\n",
                 d.L1, d.L2, d.L3, d.N1, d.N2, d.n_inf, d.T);
    std::fprintf(f, " \"V_jet\": %.17g, \"V_y\": %.17g, \"T_wall\": %.17g, \"wall_bc\": %d,\n", d.V_jet, d.V_y, d.T_wall,
                 d.wall_bc);
    std::fprintf(f, " \"face_bc\": [%d, %d, %d, %d], \"face_T\": [%.17g, %.17g, %.17g, %.17g],\n", d.face_bc[0],
                 d.face_bc[1], d.face_bc[2], d.face_bc[3], d.face_T[0], d.face_T[1], d.face_T[2], d.face_T[3]);
    std::fprintf(f, " \"tau\": %.17g, \"ppc\": %.17g, \"Fn\": %.17g, \"seed\": %llu, \"capacity\": %lld, \"sample_start\": %d,\n",
                 d.tau, d.ppc, d.Fn, (unsigned long long)d.seed, (long long)d.capacity, d.sample_start);
    std::fprintf(f, " \"zeta\": %d, \"rot_prob\": %.17g},\n", d.zeta, d.rot_prob);
    std::fprintf(f, " \"geometry\": \"%s\", \"steps\": %d, \"output_every\": %d,\n \"body\": [", m_param.geometry.c_str(),
                 m_param.steps, m_param.output_every);
    for (size_t k = 0; k + 1 < m_body.size(); k += 2)
        std::fprintf(f, "%s[%.17g, %.17g]", k ? ", " : "", m_body[k], m_body[k + 1]);
    std::fprintf(f, "]}\n");
    std::fclose(f);
}

}  // namespace dsmc_host
