// C ABI implementation (include/dsmc_abi.h): handle, device buffers, the
// step phases, hipGraph replay, per-kernel event timing and field readout.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/dsmc_abi.h"
#include "host_common.h"

using namespace dsmc;

namespace dsmc {

std::string &last_error_slot()
{
    static thread_local std::string err;
    return err;
}

}  // namespace dsmc

namespace {

#define g_err (dsmc::last_error_slot())

}  // namespace

struct dsmc_handle_s {
    int device = 0;
    hipStream_t stream = nullptr;
    Param P;
    CutTables T;
    DevCase cs{};
    int ncell = 0, ntile_scan = 0;
    int64_t cap = 0;
    SoA A{}, B{};
    Bins bn{};                   // the step's binning (k_move, k_scan_runs, k_place, k_rank)
    int *perm = nullptr, *count = nullptr, *lstart = nullptr, *tsum = nullptr, *toff = nullptr;   // + fill
    int *start = nullptr, *tile_first = nullptr, *ncoll_cell = nullptr;
    float *sigcmax = nullptr, *mrem = nullptr;
    double *acc = nullptr;
    Ctrl *ctrl = nullptr;
    float *d_seg = nullptr, *d_vol_ntc = nullptr, *d_ux = nullptr, *d_uy = nullptr, *d_sig = nullptr;
    int *d_cseg_start = nullptr, *d_cseg_idx = nullptr;
    uint8_t *d_cut = nullptr;
    float *d_fbox = nullptr;
    uint64_t *d_fmask = nullptr;
    double *d_expected = nullptr;
    // adaptive collision sub-cells (allocated on first use)
    uint32_t *d_sub = nullptr;
    float *d_lam = nullptr;
    double *d_dtv = nullptr, *d_vol = nullptr;
    int grid_amr = 0;
    int grid_move = 0, grid_scatter = 0, grid_cell = 0, grid_fill = 0;
    int tile_p = CELL_P, graph_len = 8;
    int timing = 0;
    bool cell_order = false;     // the current state is in cell order, start[] its (a step completed)
    hipGraphExec_t gexec = nullptr;
    hipGraph_t graph = nullptr;
    int graph_steps = 0;
    hipEvent_t ev[5]{};
    hipEvent_t ev_flight[2]{};   // in-flight throttle of graph replays
    std::vector<hipEvent_t> tev; // timing mode: 5 events per step of a ring
    int flight_k = 0;
    double t_ms[4]{}, t_bytes[4]{};
    int64_t t_launches = 0;
    std::vector<void *> allocs;

    template <class T>
    hipError_t alloc(T **p, size_t n)
    {
        hipError_t e = dalloc(p, n);
        if (e == hipSuccess) allocs.push_back((void *)*p);
        return e;
    }
};

namespace {

int set_device(dsmc_handle h)
{
    HIPCHK(hipSetDevice(h->device));
    return DSMC_OK;
}

int launch_step(dsmc_handle h, int phase_mask)
{
    hipStream_t st = h->stream;
    if (phase_mask & 1) HIPCHK(launch_move(h->cs, h->A, h->B, h->bn, h->ctrl, h->grid_move, st));
    if (phase_mask & 2) {
        HIPCHK(launch_scan_runs(h->bn, h->ncell, h->ctrl, st));
        HIPCHK(launch_scatter(h->bn, h->start, h->tile_first, h->ncell, h->tile_p, h->ctrl,
                              h->grid_scatter, st));
    }
    if (phase_mask & 4) {
        HIPCHK(launch_cell(h->cs, h->A, h->B, h->start, h->perm, h->bn.runlist, h->tile_first, h->sigcmax, h->mrem,
                           h->ncoll_cell, h->acc, h->ctrl, h->P.in.sample_start, h->tile_p,
                           h->grid_cell, st));
        if (h->cs.sub)
            HIPCHK(launch_amr_cell(h->cs, h->A, h->B, h->start, h->sigcmax, h->mrem, h->ncoll_cell, h->acc,
                                   h->ctrl, h->P.in.sample_start, h->grid_amr, st));
        h->cell_order = true;
    } else if (phase_mask & 1) {
        h->cell_order = false;
    }
    return DSMC_OK;
}

int read_ctrl(dsmc_handle h, Ctrl &c)
{
    HIPCHK(hipMemcpyAsync(&c, h->ctrl, sizeof(Ctrl), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    return DSMC_OK;
}

void drop_graph(dsmc_handle h)
{
    if (h->gexec) (void)hipGraphExecDestroy(h->gexec);
    if (h->graph) (void)hipGraphDestroy(h->graph);
    h->gexec = nullptr;
    h->graph = nullptr;
    h->graph_steps = 0;
}

int reset_cell_state(dsmc_handle h)
{
    std::vector<float> s0(h->ncell, (float)h->P.sigcmax0);
    HIPCHK(hipMemcpyAsync(h->sigcmax, s0.data(), sizeof(float) * h->ncell, hipMemcpyHostToDevice, h->stream));
    HIPCHK(hipMemsetAsync(h->mrem, 0, sizeof(float) * h->ncell, h->stream));
    HIPCHK(hipMemsetAsync(h->acc, 0, sizeof(double) * h->P.sample_blocks * (size_t)h->ncell, h->stream));
    HIPCHK(hipMemsetAsync(h->count, 0, sizeof(int) * h->ncell, h->stream));
    HIPCHK(hipMemsetAsync(h->bn.cnt, 0, sizeof(unsigned long long) * h->ncell, h->stream));
    HIPCHK(hipMemsetAsync(h->ncoll_cell, 0, sizeof(int) * h->ncell, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    return DSMC_OK;
}

void make_info(const Param &P, const CutTables &T, const DevCase &cs, int64_t cap, dsmc_case_info *o)
{
    std::memset(o, 0, sizeof(*o));
    o->ncell = T.ncell; o->nseg = T.nseg; o->ncut = T.ncut; o->cseg_len = (int)T.cseg_idx.size();
    o->Fn = P.Fn; o->dt = P.in.tau;
    o->L1 = cs.L1; o->L2 = cs.L2; o->dx = cs.dx; o->dy = cs.dy;
    o->ntc_coef = cs.ntc_coef; o->K1 = cs.K1; o->e_exp = cs.e_exp; o->sigcmax0 = (float)P.sigcmax0;
    o->body_vmp = cs.body_vmp; o->body_sig = cs.body_sig;
    for (int f = 0; f < 4; ++f) {
        o->face_vmp[f] = cs.face_vmp[f]; o->face_sig[f] = cs.face_sig[f];
        o->inj_e[f] = cs.inj_e[f]; o->inj_a[f] = cs.inj_a[f]; o->face_bc[f] = cs.face_bc[f];
    }
    o->inj_vmp = cs.inj_vmp; o->inj_sig = cs.inj_sig; o->inj_ux = cs.inj_ux; o->inj_uy = cs.inj_uy;
    o->margin = cs.margin; o->body_bc = cs.body_bc; o->seed = cs.seed; o->cell_volume = P.Vcell;
    o->ppc_eff = P.ppc_eff;
    o->capacity = cap;
    o->zeta = P.in.zeta;
    o->sample_blocks = P.sample_blocks;
    o->rot_p = cs.rot_p; o->lb_ia = cs.lb_ia; o->lb_ib = cs.lb_ib;
    o->body_ekt = cs.body_ekt; o->inj_ekt = cs.inj_ekt;
    for (int f = 0; f < 4; ++f) o->face_ekt[f] = cs.face_ekt[f];
}

void copy_tables(const CutTables &T, float *seg, int *cseg_start, int *cseg_idx, double *vol, float *vol_ntc,
                 uint8_t *cut, double *frac, float *fbox, uint64_t *fmask)
{
    if (fmask) std::copy(T.fmask.begin(), T.fmask.end(), fmask);
    if (seg) std::copy(T.seg.begin(), T.seg.end(), seg);
    if (cseg_start) std::copy(T.cseg_start.begin(), T.cseg_start.end(), cseg_start);
    if (cseg_idx) std::copy(T.cseg_idx.begin(), T.cseg_idx.end(), cseg_idx);
    if (vol) std::copy(T.vol.begin(), T.vol.end(), vol);
    if (vol_ntc) std::copy(T.vol_ntc.begin(), T.vol_ntc.end(), vol_ntc);
    if (cut) std::copy(T.cut.begin(), T.cut.end(), cut);
    if (frac) std::copy(T.frac.begin(), T.frac.end(), frac);
    if (fbox) std::copy(T.fbox.begin(), T.fbox.end(), fbox);
}

int fill(dsmc_handle h, const std::vector<double> &expected, const std::vector<float> &ux,
         const std::vector<float> &uy, const std::vector<float> &sig)
{
    double tot = 0;
    for (double e : expected) tot += e;
    if (tot + 6.0 * std::sqrt(tot + 1.0) + (double)h->ncell > (double)h->cap)
        return fail(DSMC_ERR_CAPACITY, "initial fill does not fit the particle capacity");
    int r = reset_cell_state(h);
    if (r) return r;
    const size_t C = h->ncell;
    HIPCHK(hipMemcpyAsync(h->d_expected, expected.data(), sizeof(double) * C, hipMemcpyHostToDevice, h->stream));
    HIPCHK(hipMemcpyAsync(h->d_ux, ux.data(), sizeof(float) * C, hipMemcpyHostToDevice, h->stream));
    HIPCHK(hipMemcpyAsync(h->d_uy, uy.data(), sizeof(float) * C, hipMemcpyHostToDevice, h->stream));
    HIPCHK(hipMemcpyAsync(h->d_sig, sig.data(), sizeof(float) * C, hipMemcpyHostToDevice, h->stream));
    HIPCHK(hipMemsetAsync(h->ctrl, 0, sizeof(Ctrl), h->stream));
    HIPCHK(launch_fill(h->cs, h->A, h->d_expected, h->d_cut, h->d_ux, h->d_uy, h->d_sig, h->count, h->lstart,
                       h->tsum, h->toff, h->ctrl, h->grid_fill, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    h->cell_order = false;
    return DSMC_OK;
}

}  // namespace

extern "C" {

const char *dsmc_last_error(void) { return g_err.c_str(); }

int dsmc_device_count(int *n)
{
    if (!n) return fail(DSMC_ERR_ARG, "null argument");
    int c = 0;
    if (hipGetDeviceCount(&c) != hipSuccess) c = 0;
    *n = c;
    return DSMC_OK;
}

int dsmc_device_cus(int *n)
{
    if (!n) return fail(DSMC_ERR_ARG, "null argument");
    int dev = 0, cu = 0;
    HIPCHK(hipGetDevice(&dev));
    HIPCHK(hipDeviceGetAttribute(&cu, hipDeviceAttributeMultiprocessorCount, dev));
    *n = cu;
    return DSMC_OK;
}

}  // extern "C"

namespace {

// Creates a handle: parameters and cut-cell tables are derived and checked
// on the host first (build fills the tables), then the device buffers are
// allocated and the tables uploaded.
template <class Build>
int create_handle(const dsmc_params *p, Build build, dsmc_handle *out)
{
    if (!p || !out) return fail(DSMC_ERR_ARG, "null argument");
    *out = nullptr;
    dsmc_handle h = new dsmc_handle_s();
    std::string msg = h->P.derive(*p);
    if (!msg.empty()) { delete h; return fail(DSMC_ERR_ARG, msg); }
    msg = build(h->P, h->T);
    if (!msg.empty()) { delete h; return fail(DSMC_ERR_GEOMETRY, msg); }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) { delete h; return fail(DSMC_ERR_NODEVICE, "no HIP device"); }
    const Param &P = h->P;
    const CutTables &T = h->T;
    h->ncell = T.ncell;
    h->ntile_scan = (T.ncell + SCAN_TILE - 1) / SCAN_TILE;
    h->cap = default_capacity(P, T);
    if (h->cap > (int64_t)0x7fffffff - 1) { delete h; return fail(DSMC_ERR_CAPACITY, "capacity above 2^31"); }

    auto bad = [&](hipError_t e, const char *what) {
        std::string m = std::string(what) + ": " + hipGetErrorString(e);
        dsmc_destroy(h);
        return fail(DSMC_ERR_HIP, m);
    };
    hipError_t e;
    if ((e = hipGetDevice(&h->device)) != hipSuccess) return bad(e, "hipGetDevice");
    if ((e = hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking)) != hipSuccess) return bad(e, "stream");
    for (int k = 0; k < 5; ++k)
        if ((e = hipEventCreate(&h->ev[k])) != hipSuccess) return bad(e, "event");
    for (int k = 0; k < 2; ++k)
        if ((e = hipEventCreateWithFlags(&h->ev_flight[k], hipEventDisableTiming)) != hipSuccess)
            return bad(e, "event");
    const size_t N = (size_t)h->cap, C = (size_t)T.ncell;
    SoA *bufs[2] = { &h->A, &h->B };
    for (SoA *s : bufs) {
        if ((e = h->alloc(&s->x, N)) || (e = h->alloc(&s->y, N)) || (e = h->alloc(&s->vx, N)) ||
            (e = h->alloc(&s->vy, N)) || (e = h->alloc(&s->vz, N)) || (e = h->alloc(&s->pid, N)) ||
            (P.in.zeta != 0 && (e = h->alloc(&s->er, N))))
            return bad(e, "particle buffers");
    }
    const int ntiles_cell = (int)((N + CELL_P_MIN - 1) / CELL_P_MIN) + 1;
    const size_t NR = (N + 63) / 64 * 64 + 64;   // run records: 64 per chunk of moved particles
    Bins &bn = h->bn;
    if ((e = h->alloc(&bn.cnt, C)) || (e = h->alloc(&bn.lstart, C)) || (e = h->alloc(&bn.tsum, h->ntile_scan)) ||
        (e = h->alloc(&bn.toff, h->ntile_scan)) || (e = h->alloc(&bn.runrec, NR)) ||
        (e = h->alloc(&bn.chunk_runs, NR / 64)) ||
        (e = h->alloc(&bn.runlist, N)))
        return bad(e, "binning buffers");
    if ((e = h->alloc(&h->perm, N)) || (e = h->alloc(&h->count, C)) ||
        (e = h->alloc(&h->lstart, C)) || (e = h->alloc(&h->tsum, h->ntile_scan)) ||
        (e = h->alloc(&h->toff, h->ntile_scan)) || (e = h->alloc(&h->start, 2 * ((size_t)C + 1))) ||
        (e = h->alloc(&h->tile_first, 4 * (size_t)ntiles_cell)) || (e = h->alloc(&h->ncoll_cell, C)) ||
        (e = h->alloc(&h->sigcmax, C)) || (e = h->alloc(&h->mrem, C)) || (e = h->alloc(&h->acc, (size_t)P.sample_blocks * C)) ||
        (e = h->alloc(&h->ctrl, CTRL_ALLOC_COUNT)) || (e = h->alloc(&h->d_seg, T.seg.size())) ||
        (e = h->alloc(&h->d_cseg_start, C + 1)) || (e = h->alloc(&h->d_cseg_idx, T.cseg_idx.size())) ||
        (e = h->alloc(&h->d_vol_ntc, C)) || (e = h->alloc(&h->d_cut, C)) || (e = h->alloc(&h->d_fbox, 4 * C)) || (e = h->alloc(&h->d_fmask, C)) ||
        (e = h->alloc(&h->d_expected, C)) || (e = h->alloc(&h->d_ux, C)) || (e = h->alloc(&h->d_uy, C)) ||
        (e = h->alloc(&h->d_sig, C)))
        return bad(e, "cell buffers");
    auto up = [&](void *d, const void *s, size_t bytes) {
        return bytes ? hipMemcpy(d, s, bytes, hipMemcpyHostToDevice) : hipSuccess;
    };
    if ((e = up(h->d_seg, T.seg.data(), sizeof(float) * T.seg.size())) ||
        (e = up(h->d_cseg_start, T.cseg_start.data(), sizeof(int) * (C + 1))) ||
        (e = up(h->d_cseg_idx, T.cseg_idx.data(), sizeof(int) * T.cseg_idx.size())) ||
        (e = up(h->d_vol_ntc, T.vol_ntc.data(), sizeof(float) * C)) ||
        (e = up(h->d_cut, T.cut.data(), C)) || (e = up(h->d_fbox, T.fbox.data(), sizeof(float) * 4 * C)) ||
        (e = up(h->d_fmask, T.fmask.data(), sizeof(uint64_t) * C)) ||
        (e = hipMemset(h->ctrl, 0, CTRL_ALLOC_BYTES)))
        return bad(e, "table upload");

    DevCase &cs = h->cs;
    cs = make_devcase(P, T, h->cap);
    cs.seg = h->d_seg; cs.cseg_start = h->d_cseg_start; cs.cseg_idx = h->d_cseg_idx; cs.vol_ntc = h->d_vol_ntc;
    cs.fbox = h->d_fbox;
    cs.fmask = h->d_fmask;

    int cu = 0;
    if ((e = hipDeviceGetAttribute(&cu, hipDeviceAttributeMultiprocessorCount, h->device)) != hipSuccess)
        return bad(e, "device attribute");
    h->grid_move = cu * occupancy_move();
    h->grid_scatter = cu * occupancy_scatter();
    h->grid_cell = cu * occupancy_cell(P.in.zeta != 0);
    h->grid_fill = cu * 4;
    int r = reset_cell_state(h);
    if (r) { std::string m = g_err; dsmc_destroy(h); return fail(r, m); }
    *out = h;
    return DSMC_OK;
}

}  // namespace

extern "C" {

int dsmc_create(const dsmc_params *p, const dsmc_body *body, dsmc_handle *out)
{
    return create_handle(p, [&](const Param &P, CutTables &T) { return build_cut_tables(P, body, T); }, out);
}

int dsmc_create_from_tables(const dsmc_params *p, const dsmc_cut_tables *t, dsmc_handle *out)
{
    if (!t) return fail(DSMC_ERR_ARG, "null host tables");
    return create_handle(p, [&](const Param &P, CutTables &T) { return build_cut_tables_from_host(P, t, T); }, out);
}

int dsmc_destroy(dsmc_handle h)
{
    if (!h) return DSMC_OK;
    (void)hipSetDevice(h->device);
    if (h->stream) (void)hipStreamSynchronize(h->stream);
    drop_graph(h);
    for (void *p : h->allocs) (void)hipFree(p);
    for (int k = 0; k < 5; ++k)
        if (h->ev[k]) (void)hipEventDestroy(h->ev[k]);
    for (int k = 0; k < 2; ++k)
        if (h->ev_flight[k]) (void)hipEventDestroy(h->ev_flight[k]);
    for (auto &e : h->tev)
        if (e) (void)hipEventDestroy(e);
    if (h->stream) (void)hipStreamDestroy(h->stream);
    delete h;
    return DSMC_OK;
}

int dsmc_case_info_get(dsmc_handle h, dsmc_case_info *o)
{
    if (!h || !o) return fail(DSMC_ERR_ARG, "null argument");
    make_info(h->P, h->T, h->cs, h->cap, o);
    return DSMC_OK;
}

int dsmc_case_tables(dsmc_handle h, float *seg, int *cseg_start, int *cseg_idx, double *vol,
                     float *vol_ntc, uint8_t *cut, float *fbox, uint64_t *fmask)
{
    if (!h) return fail(DSMC_ERR_ARG, "null handle");
    copy_tables(h->T, seg, cseg_start, cseg_idx, vol, vol_ntc, cut, nullptr, fbox, fmask);
    return DSMC_OK;
}

int dsmc_case_describe(const dsmc_params *p, const dsmc_body *body, dsmc_case_info *info, float *seg,
                       int *cseg_start, int *cseg_idx, double *vol, float *vol_ntc, uint8_t *cut,
                       double *frac, float *fbox, uint64_t *fmask)
{
    if (!p || !info) return fail(DSMC_ERR_ARG, "null argument");
    Param P;
    CutTables T;
    std::string msg = P.derive(*p);
    if (!msg.empty()) return fail(DSMC_ERR_ARG, msg);
    msg = build_cut_tables(P, body, T);
    if (!msg.empty()) return fail(DSMC_ERR_GEOMETRY, msg);
    int64_t cap = default_capacity(P, T);
    DevCase cs = make_devcase(P, T, cap);
    make_info(P, T, cs, cap, info);
    copy_tables(T, seg, cseg_start, cseg_idx, vol, vol_ntc, cut, frac, fbox, fmask);
    return DSMC_OK;
}

int dsmc_init_freestream(dsmc_handle h)
{
    if (!h) return fail(DSMC_ERR_ARG, "null handle");
    int r = set_device(h);
    if (r) return r;
    const size_t C = h->ncell;
    std::vector<double> e(C);
    for (size_t c = 0; c < C; ++c) e[c] = h->P.ppc_eff * h->T.frac[c];
    std::vector<float> ux(C, (float)h->P.in.V_jet), uy(C, (float)h->P.in.V_y), sg(C, (float)h->P.inj_sig);
    return fill(h, e, ux, uy, sg);
}

int dsmc_init_fields(dsmc_handle h, const double *n, const double *ux, const double *uy, const double *T)
{
    if (!h || !n || !ux || !uy || !T) return fail(DSMC_ERR_ARG, "null argument");
    int r = set_device(h);
    if (r) return r;
    const size_t C = h->ncell;
    std::vector<double> e(C);
    std::vector<float> fx(C), fy(C), sg(C);
    for (size_t c = 0; c < C; ++c) {
        if (!(n[c] >= 0.0) || !(T[c] > 0.0) || !std::isfinite(ux[c]) || !std::isfinite(uy[c]))
            return fail(DSMC_ERR_ARG, "init_fields: n must be >= 0, T > 0, velocities finite");
        e[c] = h->P.ppc_eff * h->T.frac[c] * (n[c] / h->P.in.n_inf);
        fx[c] = (float)ux[c];
        fy[c] = (float)uy[c];
        sg[c] = (float)std::sqrt(h->P.in.boltz * T[c] / h->P.in.mass);
    }
    return fill(h, e, fx, fy, sg);
}

int dsmc_upload_particles(dsmc_handle h, int64_t n, const float *x, const float *y, const float *vx,
                          const float *vy, const float *vz, const uint32_t *pid)
{
    if (!h || n < 0 || (n > 0 && (!x || !y || !vx || !vy || !vz))) return fail(DSMC_ERR_ARG, "bad argument");
    if (n > h->cap) return fail(DSMC_ERR_CAPACITY, "upload exceeds particle capacity");
    int r = set_device(h);
    if (r) return r;
    std::vector<uint32_t> ids;
    uint32_t next = (uint32_t)n;
    if (!pid) {
        ids.resize(n);
        for (int64_t i = 0; i < n; ++i) ids[i] = (uint32_t)i;
        pid = ids.data();
    } else {
        next = 0;
        for (int64_t i = 0; i < n; ++i) next = std::max(next, pid[i] + 1u);
    }
    const size_t b = sizeof(float) * (size_t)n;
    if (n > 0) {
        HIPCHK(hipMemcpyAsync(h->A.x, x, b, hipMemcpyHostToDevice, h->stream));
        HIPCHK(hipMemcpyAsync(h->A.y, y, b, hipMemcpyHostToDevice, h->stream));
        HIPCHK(hipMemcpyAsync(h->A.vx, vx, b, hipMemcpyHostToDevice, h->stream));
        HIPCHK(hipMemcpyAsync(h->A.vy, vy, b, hipMemcpyHostToDevice, h->stream));
        HIPCHK(hipMemcpyAsync(h->A.vz, vz, b, hipMemcpyHostToDevice, h->stream));
        HIPCHK(hipMemcpyAsync(h->A.pid, pid, sizeof(uint32_t) * (size_t)n, hipMemcpyHostToDevice, h->stream));
        // no rotational energy until dsmc_upload_rotational supplies it
        if (h->A.er) HIPCHK(hipMemsetAsync(h->A.er, 0, b, h->stream));
    }
    HIPCHK(hipMemsetAsync(h->count, 0, sizeof(int) * h->ncell, h->stream));
    HIPCHK(launch_ctrl_reset(h->ctrl, (uint32_t)n, next, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    h->cell_order = false;
    return DSMC_OK;
}

namespace {

int download_range(dsmc_handle h, bool moved, int64_t cap, float *x, float *y, float *vx, float *vy, float *vz,
                   uint32_t *pid, int64_t *n_out)
{
    int r = set_device(h);
    if (r) return r;
    Ctrl c;
    if ((r = read_ctrl(h, c))) return r;
    const uint32_t n = moved ? c.n_move : c.n_live;
    if (n_out) *n_out = n;
    if ((int64_t)n > cap) return fail(DSMC_ERR_ARG, "download buffer too small");
    const SoA &S = (c.step & 1u) ? h->B : h->A;
    const size_t b = sizeof(float) * n;
    if (n) {
        if (x) HIPCHK(hipMemcpyAsync(x, S.x, b, hipMemcpyDeviceToHost, h->stream));
        if (y) HIPCHK(hipMemcpyAsync(y, S.y, b, hipMemcpyDeviceToHost, h->stream));
        if (vx) HIPCHK(hipMemcpyAsync(vx, S.vx, b, hipMemcpyDeviceToHost, h->stream));
        if (vy) HIPCHK(hipMemcpyAsync(vy, S.vy, b, hipMemcpyDeviceToHost, h->stream));
        if (vz) HIPCHK(hipMemcpyAsync(vz, S.vz, b, hipMemcpyDeviceToHost, h->stream));
        if (pid) HIPCHK(hipMemcpyAsync(pid, S.pid, sizeof(uint32_t) * n, hipMemcpyDeviceToHost, h->stream));
    }
    HIPCHK(hipStreamSynchronize(h->stream));
    return DSMC_OK;
}

}  // namespace

int dsmc_download_particles(dsmc_handle h, int64_t cap, float *x, float *y, float *vx, float *vy, float *vz,
                            uint32_t *pid, int64_t *n_out)
{
    if (!h) return fail(DSMC_ERR_ARG, "null handle");
    return download_range(h, false, cap, x, y, vx, vy, vz, pid, n_out);
}

int dsmc_download_moved(dsmc_handle h, int64_t cap, float *x, float *y, float *vx, float *vy, float *vz,
                        uint32_t *pid, int64_t *n_out)
{
    if (!h) return fail(DSMC_ERR_ARG, "null handle");
    return download_range(h, true, cap, x, y, vx, vy, vz, pid, n_out);
}

int dsmc_move(dsmc_handle h)
{
    if (!h) return fail(DSMC_ERR_ARG, "null handle");
    int r = set_device(h);
    return r ? r : launch_step(h, 1);
}

int dsmc_reassign(dsmc_handle h)
{
    if (!h) return fail(DSMC_ERR_ARG, "null handle");
    int r = set_device(h);
    return r ? r : launch_step(h, 2);
}

int dsmc_collide(dsmc_handle h)
{
    if (!h) return fail(DSMC_ERR_ARG, "null handle");
    int r = set_device(h);
    return r ? r : launch_step(h, 4);
}

namespace {

int build_graph(dsmc_handle h)
{
    drop_graph(h);
    HIPCHK(hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
    int r = DSMC_OK;
    for (int s = 0; s < h->graph_len && r == DSMC_OK; ++s) r = launch_step(h, 7);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(h->stream, &g);
    if (r) { if (g) (void)hipGraphDestroy(g); return r; }
    if (e != hipSuccess) return fail(DSMC_ERR_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
    h->graph = g;
    HIPCHK(hipGraphInstantiate(&h->gexec, g, nullptr, nullptr, 0));
    h->graph_steps = h->graph_len;
    return DSMC_OK;
}

// Algorithmic bytes of the four timing slots (k_move; k_scan_runs; k_place +
// k_rank; k_cell) over the K steps between two reads of the counters --
// what each must read and write at least, per particle, run, candidate,
// collision and cell -- added to the timing totals.  Sums over the steps:
// live at the start (the particle-steps), injected, moved (live + injected),
// live at the end, runs.  A run record is 16 B (written by k_move, read and
// written again by k_place, read by k_rank, its base written back 4 B, read
// again by k_cell, which expands it into the tile's permutation), a packed
// cell counter 8 B (read and zeroed by the scan, which writes its 8 B tile
// scan), a cell's particle and run start 4 B each, a chunk's run count 4 B
// per 64 moved particles.  A diatomic gas
// adds its rotational energy: written for every injected particle by the
// move, gathered and written back by the cell kernel, one more sample block.
void step_bytes(dsmc_handle h, const Ctrl &before, const Ctrl &after, int K)
{
    const double C = h->ncell, k = K;
    const double zr = h->P.in.zeta ? 1.0 : 0.0;
    const double snl = (double)(after.psteps - before.psteps);
    const double sJ = (double)(after.next_pid - before.next_pid);
    const double snm = snl + sJ;
    const double snn = snl - (double)before.n_live + (double)after.n_live;
    const double S = (double)(after.nsamples - before.nsamples);      // sampled steps
    const double R = (double)(after.nruns - before.nruns);            // (cell, wave) runs binned
    h->t_bytes[0] += 40.0 * snl + (24.0 + 4.0 * zr) * sJ + (4.0 + 1.0 / 16.0) * snm + 16.0 * R + 8.0 * C * k;
    h->t_bytes[1] += 24.0 * C * k;
    h->t_bytes[2] += 52.0 * R + snm / 16.0 + 16.0 * C * k;
    h->t_bytes[3] += (48.0 + 8.0 * zr) * snn + 16.0 * R + 32.0 * (double)(after.ncand - before.ncand) +
                     24.0 * (double)(after.ncoll - before.ncoll) + 16.0 * C * k +
                     (S > 0 ? 16.0 * snn * (S / k) + (96.0 + 16.0 * zr) * C * S : 0.0);
}

}  // namespace

}  // extern "C"

namespace {

// Eager steps with HIP events between the four launches of every step (the
// batched handle's timing mode, dsmc_batch.hip): a ring of kTimeRing steps'
// events in flight, kernel durations added up as slots are reused, and
// algorithmic bytes from the counters before and after the steps.  Mode 2
// records no events and keeps only the bytes (for runs under a counter
// profiler, which times the kernels itself).
constexpr int kTimeRing = 64;

int harvest(dsmc_handle h, int slot)
{
    hipEvent_t *ev = &h->tev[5 * (size_t)slot];
    HIPCHK(hipEventSynchronize(ev[4]));
    for (int q = 0; q < 4; ++q) {
        float ms = 0.f;
        HIPCHK(hipEventElapsedTime(&ms, ev[q], ev[q + 1]));
        h->t_ms[q] += ms;
    }
    return DSMC_OK;
}

int timed_steps(dsmc_handle h, int nsteps)
{
    const bool events = h->timing == 1;
    if (events && h->tev.empty()) {
        h->tev.resize(5 * (size_t)kTimeRing, nullptr);
        for (auto &e : h->tev) HIPCHK(hipEventCreate(&e));
    }
    int r;
    Ctrl before, after;
    if ((r = read_ctrl(h, before))) return r;
    hipStream_t st = h->stream;
    for (int k = 0; k < nsteps && !events; ++k)
        if ((r = launch_step(h, 7))) return r;
    for (int k = 0; k < nsteps && events; ++k) {
        const int slot = k % kTimeRing;
        if (k >= kTimeRing && (r = harvest(h, slot))) return r;
        hipEvent_t *ev = &h->tev[5 * (size_t)slot];
        HIPCHK(hipEventRecord(ev[0], st));
        if ((r = launch_step(h, 1))) return r;
        HIPCHK(hipEventRecord(ev[1], st));
        HIPCHK(launch_scan_runs(h->bn, h->ncell, h->ctrl, st));
        HIPCHK(hipEventRecord(ev[2], st));
        HIPCHK(launch_scatter(h->bn, h->start, h->tile_first, h->ncell, h->tile_p, h->ctrl,
                              h->grid_scatter, st));
        HIPCHK(hipEventRecord(ev[3], st));
        if ((r = launch_step(h, 4))) return r;
        HIPCHK(hipEventRecord(ev[4], st));
    }
    for (int k = std::max(0, nsteps - kTimeRing); k < nsteps && events; ++k)
        if ((r = harvest(h, k % kTimeRing))) return r;
    if ((r = read_ctrl(h, after))) return r;
    step_bytes(h, before, after, nsteps);
    h->t_launches += nsteps;
    return DSMC_OK;
}

}  // namespace

extern "C" {

int dsmc_run(dsmc_handle h, int nsteps, int graph)
{
    if (!h || nsteps < 0) return fail(DSMC_ERR_ARG, "bad argument");
    int r = set_device(h);
    if (r) return r;
    if (graph && !h->timing) {
        if (!h->gexec) {
            // capture launches the kernels into the graph only; state is untouched
            if ((r = build_graph(h))) return r;
        }
        // Replays are throttled to at most 2*kFlight graphs queued ahead of the
        // device: every kFlight replays an event is recorded after waiting for
        // the one recorded kFlight replays earlier, so the queue never drains.
        constexpr int kFlight = 32;
        int full = nsteps / h->graph_steps;
        for (int k = 0; k < full; ++k) {
            HIPCHK(hipGraphLaunch(h->gexec, h->stream));
            if (++h->flight_k % kFlight == 0) {
                hipEvent_t e = h->ev_flight[(h->flight_k / kFlight) & 1];
                HIPCHK(hipEventSynchronize(e));
                HIPCHK(hipEventRecord(e, h->stream));
            }
        }
        for (int k = 0; k < nsteps % h->graph_steps; ++k)
            if ((r = launch_step(h, 7))) return r;
        return DSMC_OK;
    }
    if (!h->timing) {
        for (int k = 0; k < nsteps; ++k)
            if ((r = launch_step(h, 7))) return r;
        return DSMC_OK;
    }
    return timed_steps(h, nsteps);
}

int dsmc_sync(dsmc_handle h)
{
    if (!h) return fail(DSMC_ERR_ARG, "null handle");
    int r = set_device(h);
    if (r) return r;
    HIPCHK(hipStreamSynchronize(h->stream));
    return DSMC_OK;
}

int dsmc_status_get(dsmc_handle h, dsmc_status *o)
{
    if (!h || !o) return fail(DSMC_ERR_ARG, "null argument");
    int r = set_device(h);
    if (r) return r;
    Ctrl c;
    if ((r = read_ctrl(h, c))) return r;
    o->step = c.step; o->n_live = c.n_live; o->next_pid = c.next_pid; o->err = c.err;
    o->particle_steps = c.psteps; o->collisions = c.ncoll; o->nsamples = c.nsamples;
    o->last_collisions = c.last_coll; o->n_move = c.n_move;
    return DSMC_OK;
}

int dsmc_set_keep_perm(dsmc_handle h, int on)
{
    // the stable permutation is expanded from the ranked runs, which stay
    // until the next binning, so it is always available; the setting is
    // accepted and unused
    (void)on;
    if (!h) return fail(DSMC_ERR_ARG, "null handle");
    return DSMC_OK;
}

int dsmc_download_bins(dsmc_handle h, int64_t cap, int *cell, int *count)
{
    if (!h) return fail(DSMC_ERR_ARG, "null handle");
    int r = set_device(h);
    if (r) return r;
    Ctrl c;
    if ((r = read_ctrl(h, c))) return r;
    if (cell) {
        if ((int64_t)c.n_move > cap) return fail(DSMC_ERR_ARG, "cell buffer too small");
        // the cell of every moved particle, from the run records of its chunk
        const size_t nch = ((size_t)c.n_move + 63) / 64;
        std::vector<int> nr(nch);
        std::vector<int4> rec(nch * 64);
        if (nch) {
            HIPCHK(hipMemcpy(nr.data(), h->bn.chunk_runs, sizeof(int) * nch, hipMemcpyDeviceToHost));
            HIPCHK(hipMemcpy(rec.data(), h->bn.runrec, sizeof(int4) * nch * 64, hipMemcpyDeviceToHost));
        }
        for (uint32_t i = 0; i < c.n_move; ++i) cell[i] = -1;
        for (size_t q = 0; q < nch; ++q)
            for (int k = 0; k < nr[q]; ++k) {
                const int4 r = rec[q * 64 + k];
                const uint64_t m = ((uint64_t)(uint32_t)r.w << 32) | (uint32_t)r.z;
                for (int b = 0; b < 64; ++b)
                    if ((m >> b) & 1u) cell[q * 64 + b] = r.x;
            }
    }
    if (count) {
        // the per-cell counters of the last move (the scan clears them)
        std::vector<unsigned long long> v(h->ncell);
        HIPCHK(hipMemcpy(v.data(), h->bn.cnt, sizeof(unsigned long long) * h->ncell, hipMemcpyDeviceToHost));
        for (int k = 0; k < h->ncell; ++k) count[k] = (int)(uint32_t)v[k];
    }
    return DSMC_OK;
}

int dsmc_download_start(dsmc_handle h, int *start)
{
    if (!h || !start) return fail(DSMC_ERR_ARG, "null argument");
    int r = set_device(h);
    if (r) return r;
    HIPCHK(hipStreamSynchronize(h->stream));
    HIPCHK(hipMemcpy(start, h->start, sizeof(int) * (h->ncell + 1), hipMemcpyDeviceToHost));
    return DSMC_OK;
}

int dsmc_download_perm(dsmc_handle h, int64_t cap, int *perm)
{
    if (!h || !perm) return fail(DSMC_ERR_ARG, "null argument");
    int r = set_device(h);
    if (r) return r;
    Ctrl c;
    if ((r = read_ctrl(h, c))) return r;
    if ((int64_t)c.n_next > cap) return fail(DSMC_ERR_ARG, "perm buffer too small");
    // the permutation k_cell expands from the runs: run (id, base, lane mask)
    // holds particles 64 (id / 64) + (the mask's set bits) at positions base, ...
    std::vector<int4> runs(c.n_runs);
    if (c.n_runs) HIPCHK(hipMemcpy(runs.data(), h->bn.runlist, sizeof(int4) * c.n_runs, hipMemcpyDeviceToHost));
    for (const int4 &e : runs) {
        uint64_t m = ((uint64_t)(uint32_t)e.w << 32) | (uint32_t)e.z;
        for (int r = 0; m; ++r, m &= m - 1)
            if (e.y + r < (int64_t)c.n_next) perm[e.y + r] = (e.x & ~63) + __builtin_ctzll(m);
    }
    return DSMC_OK;
}

int dsmc_cell_state(dsmc_handle h, float *sigcmax, float *mrem, int *ncoll)
{
    if (!h) return fail(DSMC_ERR_ARG, "null handle");
    int r = set_device(h);
    if (r) return r;
    HIPCHK(hipStreamSynchronize(h->stream));
    const size_t C = h->ncell;
    if (sigcmax) HIPCHK(hipMemcpy(sigcmax, h->sigcmax, sizeof(float) * C, hipMemcpyDeviceToHost));
    if (mrem) HIPCHK(hipMemcpy(mrem, h->mrem, sizeof(float) * C, hipMemcpyDeviceToHost));
    if (ncoll) HIPCHK(hipMemcpy(ncoll, h->ncoll_cell, sizeof(int) * C, hipMemcpyDeviceToHost));
    return DSMC_OK;
}

int dsmc_set_cell_state(dsmc_handle h, const float *sigcmax, const float *mrem)
{
    if (!h) return fail(DSMC_ERR_ARG, "null handle");
    int r = set_device(h);
    if (r) return r;
    const size_t C = h->ncell;
    if (sigcmax) HIPCHK(hipMemcpy(h->sigcmax, sigcmax, sizeof(float) * C, hipMemcpyHostToDevice));
    if (mrem) HIPCHK(hipMemcpy(h->mrem, mrem, sizeof(float) * C, hipMemcpyHostToDevice));
    return DSMC_OK;
}

int dsmc_restore_counters(dsmc_handle h, const dsmc_status *s)
{
    if (!h || !s) return fail(DSMC_ERR_ARG, "null argument");
    int r = set_device(h);
    if (r) return r;
    Ctrl c;
    if ((r = read_ctrl(h, c))) return r;
    if (s->next_pid < c.next_pid) return fail(DSMC_ERR_ARG, "next_pid below the uploaded particle ids");
    // The current state lives in A on even steps and in B on odd ones; an
    // upload always writes A (and sets step 0), so an odd step moves it.
    if ((s->step & 1u) != (c.step & 1u) && c.n_live > 0) {
        const SoA &from = (c.step & 1u) ? h->B : h->A, &to = (c.step & 1u) ? h->A : h->B;
        const size_t b = sizeof(float) * (size_t)c.n_live;
        HIPCHK(hipMemcpyAsync(to.x, from.x, b, hipMemcpyDeviceToDevice, h->stream));
        HIPCHK(hipMemcpyAsync(to.y, from.y, b, hipMemcpyDeviceToDevice, h->stream));
        HIPCHK(hipMemcpyAsync(to.vx, from.vx, b, hipMemcpyDeviceToDevice, h->stream));
        HIPCHK(hipMemcpyAsync(to.vy, from.vy, b, hipMemcpyDeviceToDevice, h->stream));
        HIPCHK(hipMemcpyAsync(to.vz, from.vz, b, hipMemcpyDeviceToDevice, h->stream));
        HIPCHK(hipMemcpyAsync(to.pid, from.pid, sizeof(uint32_t) * (size_t)c.n_live, hipMemcpyDeviceToDevice,
                              h->stream));
        if (from.er) HIPCHK(hipMemcpyAsync(to.er, from.er, b, hipMemcpyDeviceToDevice, h->stream));
    }
    c.step = s->step;
    c.next_pid = s->next_pid;
    c.psteps = s->particle_steps;
    c.ncoll = s->collisions;
    c.nsamples = s->nsamples;
    c.last_coll = s->last_collisions;
    HIPCHK(hipMemcpyAsync(h->ctrl, &c, sizeof(Ctrl), hipMemcpyHostToDevice, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    return DSMC_OK;
}

int dsmc_upload_samples(dsmc_handle h, const double *acc, uint64_t nsamples)
{
    if (!h || !acc) return fail(DSMC_ERR_ARG, "null argument");
    int r = set_device(h);
    if (r) return r;
    HIPCHK(hipMemcpyAsync(h->acc, acc, sizeof(double) * h->P.sample_blocks * (size_t)h->ncell,
                          hipMemcpyHostToDevice, h->stream));
    unsigned long long ns = nsamples;
    HIPCHK(hipMemcpyAsync(&h->ctrl->nsamples, &ns, sizeof(ns), hipMemcpyHostToDevice, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    return DSMC_OK;
}

int dsmc_sample(dsmc_handle h)
{
    if (!h) return fail(DSMC_ERR_ARG, "null handle");
    if (!h->cell_order)
        return fail(DSMC_ERR_ARG, "dsmc_sample: the state is in cell order only after a completed step");
    int r = set_device(h);
    if (r) return r;
    HIPCHK(launch_sample(h->cs, h->A, h->B, h->start, h->ncoll_cell, h->acc, h->ctrl, h->stream));
    return DSMC_OK;
}

int dsmc_reset_samples(dsmc_handle h)
{
    if (!h) return fail(DSMC_ERR_ARG, "null handle");
    int r = set_device(h);
    if (r) return r;
    HIPCHK(hipMemsetAsync(h->acc, 0, sizeof(double) * h->P.sample_blocks * (size_t)h->ncell, h->stream));
    HIPCHK(hipMemsetAsync(&h->ctrl->nsamples, 0, sizeof(unsigned long long), h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    return DSMC_OK;
}

// Device buffers of the sub-cell grids, the fluid volumes (FP64) and the
// launch shape of k_amr_cell, on first use.
static int amr_alloc(dsmc_handle h)
{
    if (h->d_sub) return DSMC_OK;
    const size_t C = (size_t)h->ncell;
    hipError_t e;
    if ((e = h->alloc(&h->d_sub, C)) || (e = h->alloc(&h->d_lam, C)) || (e = h->alloc(&h->d_dtv, C)) ||
        (e = h->alloc(&h->d_vol, C)))
        return fail(DSMC_ERR_HIP, std::string("sub-cell buffers: ") + hipGetErrorString(e));
    HIPCHK(hipMemcpy(h->d_vol, h->T.vol.data(), sizeof(double) * C, hipMemcpyHostToDevice));
    HIPCHK(hipMemset(h->d_sub, 0, sizeof(uint32_t) * C));
    HIPCHK(hipMemset(h->d_lam, 0, sizeof(float) * C));
    HIPCHK(hipMemset(h->d_dtv, 0, sizeof(double) * C));
    int cu = 0;
    HIPCHK(hipDeviceGetAttribute(&cu, hipDeviceAttributeMultiprocessorCount, h->device));
    h->grid_amr = cu * occupancy_amr_cell(h->P.in.zeta != 0);
    return DSMC_OK;
}

int dsmc_set_subcells(dsmc_handle h, const uint32_t *sub)
{
    if (!h) return fail(DSMC_ERR_ARG, "null handle");
    int r = set_device(h);
    if (r) return r;
    drop_graph(h);
    if (!sub) {
        h->cs.sub = nullptr;
        return DSMC_OK;
    }
    for (int c = 0; c < h->ncell; ++c) {
        const uint32_t v = sub[c];
        if (v && (!(v & 0xffu) || !((v >> 8) & 0xffu) || !((v >> 16) & 0xffu) || !((v >> 24) & 0xffu)))
            return fail(DSMC_ERR_ARG, "dsmc_set_subcells: every level of a grid needs at least one sub-cell");
    }
    if ((r = amr_alloc(h))) return r;
    HIPCHK(hipMemcpy(h->d_sub, sub, sizeof(uint32_t) * (size_t)h->ncell, hipMemcpyHostToDevice));
    h->cs.sub = h->d_sub;
    return DSMC_OK;
}

int dsmc_amr_update(dsmc_handle h)
{
    if (!h) return fail(DSMC_ERR_ARG, "null handle");
    int r = set_device(h);
    if (r) return r;
    Ctrl c;
    if ((r = read_ctrl(h, c))) return r;
    if (c.nsamples == 0) return fail(DSMC_ERR_ARG, "dsmc_amr_update: no sampled step yet");
    if ((r = amr_alloc(h))) return r;
    const dsmc_params &p = h->P.in;
    HIPCHK(launch_amr_update(h->cs, h->P.Fn, p.mass, p.diam, p.T, p.Vtl, p.boltz, h->acc, h->d_vol, h->ctrl,
                             h->d_sub, h->d_lam, h->d_dtv, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    if (!h->cs.sub) {
        drop_graph(h);
        h->cs.sub = h->d_sub;
    }
    return DSMC_OK;
}

int dsmc_download_amr(dsmc_handle h, uint32_t *sub, float *lam, double *dt)
{
    if (!h) return fail(DSMC_ERR_ARG, "null handle");
    int r = set_device(h);
    if (r) return r;
    HIPCHK(hipStreamSynchronize(h->stream));
    const size_t C = (size_t)h->ncell;
    if (!h->d_sub) {
        if (sub) std::fill(sub, sub + C, 0u);
        if (lam) std::fill(lam, lam + C, 0.0f);
        if (dt) std::fill(dt, dt + C, 0.0);
        return DSMC_OK;
    }
    if (sub) HIPCHK(hipMemcpy(sub, h->d_sub, sizeof(uint32_t) * C, hipMemcpyDeviceToHost));
    if (lam) HIPCHK(hipMemcpy(lam, h->d_lam, sizeof(float) * C, hipMemcpyDeviceToHost));
    if (dt) HIPCHK(hipMemcpy(dt, h->d_dtv, sizeof(double) * C, hipMemcpyDeviceToHost));
    return DSMC_OK;
}

int dsmc_download_samples(dsmc_handle h, double *acc, uint64_t *nsamples)
{
    if (!h) return fail(DSMC_ERR_ARG, "null handle");
    int r = set_device(h);
    if (r) return r;
    Ctrl c;
    if ((r = read_ctrl(h, c))) return r;
    if (nsamples) *nsamples = c.nsamples;
    if (acc)
        HIPCHK(hipMemcpy(acc, h->acc, sizeof(double) * h->P.sample_blocks * (size_t)h->ncell, hipMemcpyDeviceToHost));
    return DSMC_OK;
}

int dsmc_download_fields(dsmc_handle h, double *n, double *ux, double *uy, double *uz, double *T, double *p)
{
    if (!h) return fail(DSMC_ERR_ARG, "null handle");
    const size_t C = h->ncell;
    std::vector<double> a(h->P.sample_blocks * C);
    uint64_t ns = 0;
    int r = dsmc_download_samples(h, a.data(), &ns);
    if (r) return r;
    fields_from_sums(h->P, h->T, a.data(), C, ns, n, ux, uy, uz, T, p);
    return DSMC_OK;
}

int dsmc_download_rot_temperature(dsmc_handle h, double *Trot)
{
    if (!h || !Trot) return fail(DSMC_ERR_ARG, "null argument");
    if (h->P.in.zeta == 0) return fail(DSMC_ERR_ARG, "rot_temperature: the gas is monatomic (zeta = 0)");
    const size_t C = h->ncell;
    std::vector<double> a(h->P.sample_blocks * C);
    uint64_t ns = 0;
    int r = dsmc_download_samples(h, a.data(), &ns);
    if (r) return r;
    rot_temperature_from_sums(h->P, h->T, a.data(), C, Trot);
    return DSMC_OK;
}

int dsmc_upload_rotational(dsmc_handle h, int64_t n, const float *er)
{
    if (!h || (n > 0 && !er)) return fail(DSMC_ERR_ARG, "bad argument");
    if (h->P.in.zeta == 0) return fail(DSMC_ERR_ARG, "upload_rotational: the gas is monatomic (zeta = 0)");
    int r = set_device(h);
    if (r) return r;
    Ctrl c;
    if ((r = read_ctrl(h, c))) return r;
    if (n != (int64_t)c.n_live)
        return fail(DSMC_ERR_ARG, "upload_rotational: n must equal the particles of the current state");
    for (int64_t i = 0; i < n; ++i)
        if (!(er[i] >= 0.0f) || !std::isfinite(er[i]))
            return fail(DSMC_ERR_ARG, "upload_rotational: energies must be finite and >= 0");
    const SoA &S = (c.step & 1u) ? h->B : h->A;
    if (n) HIPCHK(hipMemcpyAsync(S.er, er, sizeof(float) * (size_t)n, hipMemcpyHostToDevice, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    return DSMC_OK;
}

int dsmc_download_rotational(dsmc_handle h, int moved, int64_t cap, float *er, int64_t *n_out)
{
    if (!h) return fail(DSMC_ERR_ARG, "null handle");
    if (h->P.in.zeta == 0) return fail(DSMC_ERR_ARG, "download_rotational: the gas is monatomic (zeta = 0)");
    int r = set_device(h);
    if (r) return r;
    Ctrl c;
    if ((r = read_ctrl(h, c))) return r;
    const uint32_t n = moved ? c.n_move : c.n_live;
    if (n_out) *n_out = n;
    if ((int64_t)n > cap) return fail(DSMC_ERR_ARG, "download buffer too small");
    const SoA &S = (c.step & 1u) ? h->B : h->A;
    if (n && er) HIPCHK(hipMemcpyAsync(er, S.er, sizeof(float) * n, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    return DSMC_OK;
}

int dsmc_get_launch(dsmc_handle h, dsmc_launch *o)
{
    if (!h || !o) return fail(DSMC_ERR_ARG, "null argument");
    o->move_blocks = h->grid_move;
    o->scatter_blocks = h->grid_scatter;
    o->cell_blocks = h->grid_cell;
    o->cell_tile = h->tile_p;
    o->graph_steps = h->graph_len;
    return DSMC_OK;
}

int dsmc_set_launch(dsmc_handle h, const dsmc_launch *c)
{
    if (!h || !c) return fail(DSMC_ERR_ARG, "null argument");
    int r = check_launch(c);
    if (r) return r;
    h->grid_move = c->move_blocks;
    h->grid_scatter = c->scatter_blocks;
    h->grid_cell = c->cell_blocks;
    h->tile_p = c->cell_tile;
    h->graph_len = c->graph_steps;
    drop_graph(h);
    return DSMC_OK;
}

int dsmc_set_timing(dsmc_handle h, int on)
{
    if (!h) return fail(DSMC_ERR_ARG, "null handle");
    if (on < 0 || on > 2) return fail(DSMC_ERR_ARG, "set_timing: 0 off, 1 events and bytes, 2 bytes only");
    h->timing = on;
    for (int q = 0; q < 4; ++q) h->t_ms[q] = h->t_bytes[q] = 0.0;
    h->t_launches = 0;
    return DSMC_OK;
}

int dsmc_kernel_times(dsmc_handle h, double ms[4], double bytes[4], int64_t *launches)
{
    if (!h) return fail(DSMC_ERR_ARG, "null handle");
    for (int q = 0; q < 4; ++q) {
        double n = h->t_launches ? (double)h->t_launches : 1.0;
        if (ms) ms[q] = h->t_ms[q] / n;
        if (bytes) bytes[q] = h->t_bytes[q] / n;
    }
    if (launches) *launches = h->t_launches;
    return DSMC_OK;
}

}  // extern "C"
