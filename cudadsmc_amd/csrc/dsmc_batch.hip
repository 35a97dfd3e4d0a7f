// Batched instances (include/dsmc_abi.h, dsmc_batch_*): many independent
// DSMC cases in one set of device arrays, advanced by one launch per phase.
//
// Layout: the cells of all instances are concatenated in instance order
// (instance i owns global cells [cbase_i, cbase_i + C_i)); the particles
// are kept in global cell order by k_cell, so instance i's particles are a
// contiguous range whose start is the prefix of the per-instance live
// counts.  The scan and the scatter run unchanged over all cells; k_move
// and k_cell find each particle's / cell's instance in LDS offset tables.
// No instance sees another: every random stream is keyed on the instance's
// own particle ids, cell ids and seed, so an instance evolves exactly as it
// would in a dsmc_handle of its own.
#include <hip/hip_runtime.h>

#include <cstring>
#include <vector>

#include "../../include/dsmc_abi.h"
#include "host_common.h"

using namespace dsmc;

struct dsmc_batch_s {
    int device = 0;
    hipStream_t stream = nullptr;
    int ninst = 0;
    std::vector<Param> P;
    std::vector<CutTables> T;
    std::vector<DevCase> cs;           // per instance, device table pointers set
    std::vector<int> cbase;            // ninst + 1
    std::vector<int64_t> capi;         // per-instance capacity
    std::vector<double> poly;          // body outlines of all instances, x y pairs
    std::vector<int> poly0, npoly;     // per instance: first vertex in poly, vertex count
    int ncell = 0, ntile_scan = 0;
    int64_t cap = 0;
    SoA A{}, B{};
    Bins bn{};                          // the step's binning over all instances' cells
    int *perm = nullptr, *count = nullptr, *lstart = nullptr, *tsum = nullptr, *toff = nullptr;   // + fill
    int *start = nullptr, *tile_first = nullptr, *ncoll_cell = nullptr;
    float *sigcmax = nullptr, *mrem = nullptr;
    double *acc = nullptr;
    Ctrl *ctrl = nullptr, *fill_ctrl = nullptr;
    InstCtrl *ictl = nullptr;
    DevCase *d_cases = nullptr;
    CellConst *d_cc = nullptr;
    int *d_cbase = nullptr;
    float *d_seg = nullptr, *d_vol_ntc = nullptr, *d_fbox = nullptr;
    uint64_t *d_fmask = nullptr;
    float *d_ux = nullptr, *d_uy = nullptr, *d_sig = nullptr;
    int *d_cseg_start = nullptr, *d_cseg_idx = nullptr;
    uint8_t *d_cut = nullptr;
    double *d_expected = nullptr;
    DevCase all{};                     // what k_cell reads of "the case": all cells
    int grid_move = 0, grid_scatter = 0, grid_cell = 0, grid_fill = 0;
    int tile_p = CELL_P, graph_len = 8;
    hipGraphExec_t gexec = nullptr;
    hipGraph_t graph = nullptr;
    int graph_steps = 0;
    hipEvent_t ev[5]{};
    hipEvent_t ev_flight[2]{};
    int flight_k = 0;
    int timing = 0;
    std::vector<hipEvent_t> tev;        // timing mode: 5 events per step of a ring
    double t_ms[4]{}, t_bytes[4]{};
    int64_t t_launches = 0;
    std::vector<void *> allocs;

    template <class U>
    hipError_t alloc(U **p, size_t n)
    {
        hipError_t e = dalloc(p, n);
        if (e == hipSuccess) allocs.push_back((void *)*p);
        return e;
    }
};

namespace {

int set_device(dsmc_batch b)
{
    HIPCHK(hipSetDevice(b->device));
    return DSMC_OK;
}

int launch_step(dsmc_batch b, int mask)
{
    hipStream_t st = b->stream;
    if (mask & 1)
        HIPCHK(launch_move_batch(b->d_cases, b->ninst, b->d_cbase, b->A, b->B, b->bn, b->ctrl, b->ictl,
                                 b->grid_move, st));
    if (mask & 2) HIPCHK(launch_scan_runs(b->bn, b->ncell, b->ctrl, st));
    if (mask & 4)
        HIPCHK(launch_scatter(b->bn, b->start, b->tile_first, b->ncell, b->tile_p, b->ctrl,
                              b->grid_scatter, st));
    if (mask & 8)
        HIPCHK(launch_cell_batch(b->all, b->d_cc, b->d_cbase, b->ninst, b->ictl, b->A, b->B, b->start, b->perm,
                                 b->bn.runlist, b->tile_first, b->sigcmax, b->mrem, b->ncoll_cell, b->acc, b->ctrl, b->tile_p,
                                 b->grid_cell, st));
    return DSMC_OK;
}

int read_ctrl(dsmc_batch b, Ctrl &c, std::vector<InstCtrl> *ic = nullptr)
{
    HIPCHK(hipMemcpyAsync(&c, b->ctrl, sizeof(Ctrl), hipMemcpyDeviceToHost, b->stream));
    if (ic) {
        ic->resize(b->ninst);
        HIPCHK(hipMemcpyAsync(ic->data(), b->ictl, sizeof(InstCtrl) * b->ninst, hipMemcpyDeviceToHost, b->stream));
    }
    HIPCHK(hipStreamSynchronize(b->stream));
    return DSMC_OK;
}

void drop_graph(dsmc_batch b)
{
    if (b->gexec) (void)hipGraphExecDestroy(b->gexec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->gexec = nullptr;
    b->graph = nullptr;
    b->graph_steps = 0;
}

int build_graph(dsmc_batch b)
{
    drop_graph(b);
    HIPCHK(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
    int r = DSMC_OK;
    for (int s = 0; s < b->graph_len && r == DSMC_OK; ++s) r = launch_step(b, 15);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(b->stream, &g);
    if (r) { if (g) (void)hipGraphDestroy(g); return r; }
    if (e != hipSuccess) return fail(DSMC_ERR_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
    b->graph = g;
    HIPCHK(hipGraphInstantiate(&b->gexec, g, nullptr, nullptr, 0));
    b->graph_steps = b->graph_len;
    return DSMC_OK;
}

int reset_cell_state(dsmc_batch b)
{
    std::vector<float> s0(b->ncell);
    for (int i = 0; i < b->ninst; ++i)
        for (int c = b->cbase[i]; c < b->cbase[i + 1]; ++c) s0[c] = (float)b->P[i].sigcmax0;
    const size_t C = b->ncell;
    HIPCHK(hipMemcpyAsync(b->sigcmax, s0.data(), sizeof(float) * C, hipMemcpyHostToDevice, b->stream));
    HIPCHK(hipMemsetAsync(b->mrem, 0, sizeof(float) * C, b->stream));
    HIPCHK(hipMemsetAsync(b->acc, 0, sizeof(double) * 6 * C, b->stream));
    HIPCHK(hipMemsetAsync(b->count, 0, sizeof(int) * C, b->stream));
    HIPCHK(hipMemsetAsync(b->bn.cnt, 0, sizeof(unsigned long long) * C, b->stream));
    HIPCHK(hipMemsetAsync(b->ncoll_cell, 0, sizeof(int) * C, b->stream));
    HIPCHK(hipStreamSynchronize(b->stream));
    return DSMC_OK;
}

// Fills every instance from per-cell expected counts and drifting Maxwellians
// (the same kernels as a single case, one instance after the other, each
// writing its particles after the previous instance's).
int fill(dsmc_batch b, const std::vector<double> &e, const std::vector<float> &ux, const std::vector<float> &uy,
         const std::vector<float> &sig)
{
    for (int i = 0; i < b->ninst; ++i) {
        double tot = 0;
        for (int c = b->cbase[i]; c < b->cbase[i + 1]; ++c) tot += e[c];
        if (tot + 6.0 * std::sqrt(tot + 1.0) + (double)(b->cbase[i + 1] - b->cbase[i]) > (double)b->capi[i])
            return fail(DSMC_ERR_CAPACITY, "instance " + std::to_string(i) + ": fill exceeds its particle capacity");
    }
    drop_graph(b);
    int r = reset_cell_state(b);
    if (r) return r;
    const size_t C = b->ncell;
    hipStream_t st = b->stream;
    HIPCHK(hipMemcpyAsync(b->d_expected, e.data(), sizeof(double) * C, hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(b->d_ux, ux.data(), sizeof(float) * C, hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(b->d_uy, uy.data(), sizeof(float) * C, hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(b->d_sig, sig.data(), sizeof(float) * C, hipMemcpyHostToDevice, st));
    std::vector<InstCtrl> ic(b->ninst);
    uint32_t off = 0, err_all = 0;
    for (int i = 0; i < b->ninst; ++i) {
        const int cb = b->cbase[i];
        SoA Ai{ b->A.x + off, b->A.y + off, b->A.vx + off, b->A.vy + off, b->A.vz + off, b->A.pid + off };
        HIPCHK(hipMemsetAsync(b->fill_ctrl, 0, sizeof(Ctrl), st));
        HIPCHK(launch_fill(b->cs[i], Ai, b->d_expected + cb, b->d_cut + cb, b->d_ux + cb, b->d_uy + cb,
                           b->d_sig + cb, b->count + cb, b->lstart, b->tsum, b->toff, b->fill_ctrl, b->grid_fill, st));
        Ctrl f;
        HIPCHK(hipMemcpyAsync(&f, b->fill_ctrl, sizeof(Ctrl), hipMemcpyDeviceToHost, st));
        HIPCHK(hipStreamSynchronize(st));
        std::memset(&ic[i], 0, sizeof(InstCtrl));
        ic[i].n_live = ic[i].n_move = ic[i].next_pid = f.n_live;
        ic[i].err = f.err;
        err_all |= f.err;
        off += f.n_live;
    }
    Ctrl g;
    std::memset(&g, 0, sizeof(g));
    g.n_live = g.n_move = g.n_next = off;
    g.err = err_all;
    HIPCHK(hipMemcpyAsync(b->ictl, ic.data(), sizeof(InstCtrl) * b->ninst, hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(b->ctrl, &g, sizeof(Ctrl), hipMemcpyHostToDevice, st));
    HIPCHK(hipStreamSynchronize(st));
    return DSMC_OK;
}

bool valid_index(dsmc_batch b, int i) { return b && i >= 0 && i < b->ninst; }

}  // namespace

extern "C" {

int dsmc_batch_destroy(dsmc_batch b)
{
    if (!b) return DSMC_OK;
    (void)hipSetDevice(b->device);
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    drop_graph(b);
    for (void *p : b->allocs) (void)hipFree(p);
    for (auto &e : b->ev)
        if (e) (void)hipEventDestroy(e);
    for (auto &e : b->ev_flight)
        if (e) (void)hipEventDestroy(e);
    for (auto &e : b->tev)
        if (e) (void)hipEventDestroy(e);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
    return DSMC_OK;
}

namespace {
int setup_device(dsmc_batch b, dsmc_batch *out);
}

int dsmc_batch_create(const dsmc_params *params, const dsmc_body *bodies, int n, dsmc_batch *out)
{
    if (!params || !out || n < 1) return fail(DSMC_ERR_ARG, "batch needs n >= 1 cases");
    if (n > MAX_INST) return fail(DSMC_ERR_ARG, "batch holds at most " + std::to_string(MAX_INST) + " instances");
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return fail(DSMC_ERR_NODEVICE, "no HIP device");
    dsmc_batch b = new dsmc_batch_s();
    b->ninst = n;
    b->P.resize(n);
    b->T.resize(n);
    for (int i = 0; i < n; ++i) {
        std::string msg = b->P[i].derive(params[i]);
        if (!msg.empty()) { delete b; return fail(DSMC_ERR_ARG, "instance " + std::to_string(i) + ": " + msg); }
        if (params[i].zeta != 0) {
            delete b;
            return fail(DSMC_ERR_ARG, "instance " + std::to_string(i) +
                                          ": batched instances are monatomic (zeta = 0); run a diatomic gas "
                                          "with a dsmc_handle");
        }
        const dsmc_body *bd = (bodies && bodies[i].n_points > 0) ? &bodies[i] : nullptr;
        msg = build_cut_tables(b->P[i], bd, b->T[i]);
        if (!msg.empty()) { delete b; return fail(DSMC_ERR_GEOMETRY, "instance " + std::to_string(i) + ": " + msg); }
        b->poly0.push_back((int)(b->poly.size() / 2));
        b->npoly.push_back(bd ? bd->n_points : 0);
        if (bd) b->poly.insert(b->poly.end(), bd->xy, bd->xy + 2 * (size_t)bd->n_points);
    }
    return setup_device(b, out);
}

namespace {

// Device arrays, uploaded tables and launch sizes of a batch whose host
// side (P, T, body outlines) is set; takes ownership of b (freed on failure).
int setup_device(dsmc_batch b, dsmc_batch *out)
{
    const int n = b->ninst;
    b->cs.resize(n);
    b->capi.resize(n);
    b->cbase.assign(n + 1, 0);
    int64_t C = 0, segs = 0, idx = 0;
    for (int i = 0; i < n; ++i) {
        b->capi[i] = default_capacity(b->P[i], b->T[i]);
        b->cap += b->capi[i];
        C += b->T[i].ncell;
        segs += (int64_t)b->T[i].seg.size();
        idx += (int64_t)b->T[i].cseg_idx.size();
        b->cbase[i + 1] = (int)std::min<int64_t>(C, 0x7fffffff);
    }
    if (C >= (1LL << 30) || b->cap > (int64_t)0x7fffffff - 1) {
        delete b;
        return fail(DSMC_ERR_CAPACITY, "batch too large: cells or particle capacity above 2^31");
    }
    b->ncell = (int)C;
    b->ntile_scan = (b->ncell + SCAN_TILE - 1) / SCAN_TILE;
    auto bad = [&](hipError_t e, const char *what) {
        std::string m = std::string(what) + ": " + hipGetErrorString(e);
        dsmc_batch_destroy(b);
        return fail(DSMC_ERR_HIP, m);
    };
    hipError_t e;
    if ((e = hipGetDevice(&b->device)) != hipSuccess) return bad(e, "hipGetDevice");
    if ((e = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking)) != hipSuccess) return bad(e, "stream");
    for (auto &ev : b->ev)
        if ((e = hipEventCreate(&ev)) != hipSuccess) return bad(e, "event");
    for (auto &ev : b->ev_flight)
        if ((e = hipEventCreateWithFlags(&ev, hipEventDisableTiming)) != hipSuccess) return bad(e, "event");
    const size_t N = (size_t)b->cap, Cs = (size_t)C;
    for (SoA *s : { &b->A, &b->B })
        if ((e = b->alloc(&s->x, N)) || (e = b->alloc(&s->y, N)) || (e = b->alloc(&s->vx, N)) ||
            (e = b->alloc(&s->vy, N)) || (e = b->alloc(&s->vz, N)) || (e = b->alloc(&s->pid, N)))
            return bad(e, "particle buffers");
    const int ntiles_cell = (int)((N + CELL_P_MIN - 1) / CELL_P_MIN) + 1;
    const size_t NR = (N + 63) / 64 * 64 + 64;   // run records: 64 per chunk of moved particles
    Bins &bn = b->bn;
    if ((e = b->alloc(&bn.cnt, Cs)) || (e = b->alloc(&bn.lstart, Cs)) || (e = b->alloc(&bn.tsum, b->ntile_scan)) ||
        (e = b->alloc(&bn.toff, b->ntile_scan)) || (e = b->alloc(&bn.runrec, NR)) ||
        (e = b->alloc(&bn.chunk_runs, NR / 64)) ||
        (e = b->alloc(&bn.runlist, N)))
        return bad(e, "binning buffers");
    if ((e = b->alloc(&b->perm, N)) || (e = b->alloc(&b->count, Cs)) ||
        (e = b->alloc(&b->lstart, Cs)) || (e = b->alloc(&b->tsum, b->ntile_scan)) ||
        (e = b->alloc(&b->toff, b->ntile_scan)) || (e = b->alloc(&b->start, 2 * ((size_t)Cs + 1))) ||
        (e = b->alloc(&b->tile_first, 4 * (size_t)ntiles_cell)) || (e = b->alloc(&b->ncoll_cell, Cs)) ||
        (e = b->alloc(&b->sigcmax, Cs)) || (e = b->alloc(&b->mrem, Cs)) || (e = b->alloc(&b->acc, 6 * Cs)) ||
        (e = b->alloc(&b->ctrl, CTRL_ALLOC_COUNT)) || (e = b->alloc(&b->fill_ctrl, 1)) || (e = b->alloc(&b->ictl, n)) ||
        (e = b->alloc(&b->d_cases, n)) || (e = b->alloc(&b->d_cc, n)) || (e = b->alloc(&b->d_cbase, n + 1)) ||
        (e = b->alloc(&b->d_seg, segs)) || (e = b->alloc(&b->d_cseg_start, Cs + n)) ||
        (e = b->alloc(&b->d_cseg_idx, idx)) || (e = b->alloc(&b->d_vol_ntc, Cs)) || (e = b->alloc(&b->d_cut, Cs)) ||
        (e = b->alloc(&b->d_fbox, 4 * Cs)) || (e = b->alloc(&b->d_fmask, Cs)) || (e = b->alloc(&b->d_expected, Cs)) || (e = b->alloc(&b->d_ux, Cs)) ||
        (e = b->alloc(&b->d_uy, Cs)) || (e = b->alloc(&b->d_sig, Cs)))
        return bad(e, "cell buffers");

    // concatenated tables; every instance's DevCase points at its slices
    std::vector<CellConst> cc(n);
    int64_t so = 0, io = 0;
    for (int i = 0; i < n; ++i) {
        const CutTables &Ti = b->T[i];
        const int cb = b->cbase[i];
        const size_t Ci = Ti.ncell;
        auto up = [&](void *d, const void *s, size_t bytes) {
            return bytes ? hipMemcpy(d, s, bytes, hipMemcpyHostToDevice) : hipSuccess;
        };
        if ((e = up(b->d_seg + so, Ti.seg.data(), sizeof(float) * Ti.seg.size())) ||
            (e = up(b->d_cseg_start + cb + i, Ti.cseg_start.data(), sizeof(int) * (Ci + 1))) ||
            (e = up(b->d_cseg_idx + io, Ti.cseg_idx.data(), sizeof(int) * Ti.cseg_idx.size())) ||
            (e = up(b->d_vol_ntc + cb, Ti.vol_ntc.data(), sizeof(float) * Ci)) ||
            (e = up(b->d_cut + cb, Ti.cut.data(), Ci)) ||
            (e = up(b->d_fbox + 4 * (size_t)cb, Ti.fbox.data(), sizeof(float) * 4 * Ci)) ||
            (e = up(b->d_fmask + cb, Ti.fmask.data(), sizeof(uint64_t) * Ci)))
            return bad(e, "table upload");
        DevCase &c = b->cs[i];
        c = make_devcase(b->P[i], Ti, b->capi[i]);
        c.seg = b->d_seg + so;
        c.cseg_start = b->d_cseg_start + cb + i;
        c.cseg_idx = b->d_cseg_idx + io;
        c.vol_ntc = b->d_vol_ntc + cb;
        c.fbox = b->d_fbox + 4 * (size_t)cb;
        c.fmask = b->d_fmask + cb;
        cc[i] = CellConst{ c.ntc_coef, c.K1, c.e_exp, c.seed, b->P[i].in.sample_start, 0 };
        so += (int64_t)Ti.seg.size();
        io += (int64_t)Ti.cseg_idx.size();
    }
    b->all = DevCase{};
    b->all.ncell = b->ncell;
    b->all.vol_ntc = b->d_vol_ntc;
    if ((e = hipMemcpy(b->d_cases, b->cs.data(), sizeof(DevCase) * n, hipMemcpyHostToDevice)) ||
        (e = hipMemcpy(b->d_cc, cc.data(), sizeof(CellConst) * n, hipMemcpyHostToDevice)) ||
        (e = hipMemcpy(b->d_cbase, b->cbase.data(), sizeof(int) * (n + 1), hipMemcpyHostToDevice)) ||
        (e = hipMemset(b->ctrl, 0, CTRL_ALLOC_BYTES)) || (e = hipMemset(b->ictl, 0, sizeof(InstCtrl) * n)))
        return bad(e, "table upload");

    int cu = 0;
    if ((e = hipDeviceGetAttribute(&cu, hipDeviceAttributeMultiprocessorCount, b->device)) != hipSuccess)
        return bad(e, "device attribute");
    b->grid_move = cu * occupancy_move_batch();
    b->grid_scatter = cu * occupancy_scatter();
    b->grid_cell = cu * occupancy_cell_batch();
    b->grid_fill = cu * 4;
    int r = reset_cell_state(b);
    if (r) { std::string m = last_error_slot(); dsmc_batch_destroy(b); return fail(r, m); }
    *out = b;
    return DSMC_OK;
}

}  // namespace

int dsmc_batch_case_info(dsmc_batch b, int i, dsmc_case_info *o)
{
    if (!valid_index(b, i) || !o) return fail(DSMC_ERR_ARG, "bad instance index or null argument");
    const Param &P = b->P[i];
    const CutTables &T = b->T[i];
    const DevCase &cs = b->cs[i];
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
    o->capacity = b->capi[i];
    return DSMC_OK;
}

int dsmc_batch_init_freestream(dsmc_batch b)
{
    if (!b) return fail(DSMC_ERR_ARG, "null batch");
    int r = set_device(b);
    if (r) return r;
    const size_t C = b->ncell;
    std::vector<double> e(C);
    std::vector<float> ux(C), uy(C), sg(C);
    for (int i = 0; i < b->ninst; ++i)
        for (int c = b->cbase[i]; c < b->cbase[i + 1]; ++c) {
            const int lc = c - b->cbase[i];
            e[c] = b->P[i].ppc_eff * b->T[i].frac[lc];
            ux[c] = (float)b->P[i].in.V_jet;
            uy[c] = (float)b->P[i].in.V_y;
            sg[c] = (float)b->P[i].inj_sig;
        }
    return fill(b, e, ux, uy, sg);
}

int dsmc_batch_init_fields(dsmc_batch b, const double *n, const double *ux, const double *uy, const double *T)
{
    if (!b || !n || !ux || !uy || !T) return fail(DSMC_ERR_ARG, "null argument");
    int r = set_device(b);
    if (r) return r;
    const size_t C = b->ncell;
    std::vector<double> e(C);
    std::vector<float> fx(C), fy(C), sg(C);
    for (int i = 0; i < b->ninst; ++i) {
        const Param &P = b->P[i];
        for (int c = b->cbase[i]; c < b->cbase[i + 1]; ++c) {
            if (!(n[c] >= 0.0) || !(T[c] > 0.0) || !std::isfinite(ux[c]) || !std::isfinite(uy[c]))
                return fail(DSMC_ERR_ARG, "init_fields: n must be >= 0, T > 0, velocities finite");
            e[c] = P.ppc_eff * b->T[i].frac[c - b->cbase[i]] * (n[c] / P.in.n_inf);
            fx[c] = (float)ux[c];
            fy[c] = (float)uy[c];
            sg[c] = (float)std::sqrt(P.in.boltz * T[c] / P.in.mass);
        }
    }
    return fill(b, e, fx, fy, sg);
}

}  // extern "C"

namespace {

// Eager steps with HIP events between the four launches of every step, on
// the batch's stream.  The events of the last kTimeRing steps are in flight
// at once: before a step reuses its ring slot, the slot's previous step is
// waited for and its kernel durations are added up, so the queue is
// throttled without being drained.  Algorithmic bytes follow from the
// device counters read once before and once after the steps (sums over the
// steps of live, moved and new particle counts, candidates, collisions and
// sampled steps), so no step waits for the host.  Mode 2 records no events
// and keeps only the bytes (for runs under a counter profiler, which times
// the kernels itself).
constexpr int kTimeRing = 64;

int harvest(dsmc_batch b, int slot)
{
    hipEvent_t *ev = &b->tev[5 * (size_t)slot];
    HIPCHK(hipEventSynchronize(ev[4]));
    for (int q = 0; q < 4; ++q) {
        float ms = 0.f;
        HIPCHK(hipEventElapsedTime(&ms, ev[q], ev[q + 1]));
        b->t_ms[q] += ms;
    }
    return DSMC_OK;
}

int timed_steps(dsmc_batch b, int nsteps)
{
    const bool events = b->timing == 1;
    if (events && b->tev.empty()) {
        b->tev.resize(5 * (size_t)kTimeRing, nullptr);
        for (auto &e : b->tev) HIPCHK(hipEventCreate(&e));
    }
    int r;
    Ctrl before, after;
    std::vector<InstCtrl> ib, ia;
    if ((r = read_ctrl(b, before, &ib))) return r;
    hipStream_t st = b->stream;
    for (int k = 0; k < nsteps && !events; ++k)
        if ((r = launch_step(b, 15))) return r;
    for (int k = 0; k < nsteps && events; ++k) {
        const int slot = k % kTimeRing;
        if (k >= kTimeRing && (r = harvest(b, slot))) return r;
        hipEvent_t *ev = &b->tev[5 * (size_t)slot];
        HIPCHK(hipEventRecord(ev[0], st));
        for (int q = 0; q < 4; ++q) {
            if ((r = launch_step(b, 1 << q))) return r;
            HIPCHK(hipEventRecord(ev[q + 1], st));
        }
    }
    for (int k = std::max(0, nsteps - kTimeRing); k < nsteps && events; ++k)
        if ((r = harvest(b, k % kTimeRing))) return r;
    if ((r = read_ctrl(b, after, &ia))) return r;
    // sums over the steps: live at the start (the step's particle-steps),
    // injected, moved (live + injected), live at the end, (cell, wave) runs;
    // the per-unit figures are those of the single-instance handle
    const double K = nsteps, C = b->ncell;
    const double R = (double)(after.nruns - before.nruns);
    const double snl = (double)(after.psteps - before.psteps);
    const double sJ = (double)(after.next_pid - before.next_pid);
    const double snm = snl + sJ;
    const double snn = snl - (double)before.n_live + (double)after.n_live;
    double sampled = 0.0;
    for (int i = 0; i < b->ninst; ++i) {
        const double S = (double)(ia[i].nsamples - ib[i].nsamples);
        if (S <= 0) continue;
        const double nn_i = (double)(ia[i].psteps - ib[i].psteps) - ib[i].n_live + ia[i].n_live;
        sampled += 16.0 * nn_i * (S / K) + 96.0 * (double)(b->cbase[i + 1] - b->cbase[i]) * S;
    }
    b->t_bytes[0] += 40.0 * snl + 24.0 * sJ + (4.0 + 1.0 / 16.0) * snm + 16.0 * R + 8.0 * C * K;
    b->t_bytes[1] += 24.0 * C * K;
    b->t_bytes[2] += 52.0 * R + snm / 16.0 + 16.0 * C * K;
    b->t_bytes[3] += 48.0 * snn + 16.0 * R + 32.0 * (double)(after.ncand - before.ncand) +
                     24.0 * (double)(after.ncoll - before.ncoll) + 16.0 * C * K + sampled;
    b->t_launches += nsteps;
    return DSMC_OK;
}

}  // namespace

extern "C" {

int dsmc_batch_run(dsmc_batch b, int nsteps, int graph)
{
    if (!b || nsteps < 0) return fail(DSMC_ERR_ARG, "bad argument");
    int r = set_device(b);
    if (r) return r;
    if (!b->timing) {
        int done = 0;
        if (graph) {
            if (!b->gexec && (r = build_graph(b))) return r;
            constexpr int kFlight = 32;   // at most 2*kFlight replays queued ahead
            for (; done + b->graph_steps <= nsteps; done += b->graph_steps) {
                HIPCHK(hipGraphLaunch(b->gexec, b->stream));
                if (++b->flight_k % kFlight == 0) {
                    hipEvent_t ev = b->ev_flight[(b->flight_k / kFlight) & 1];
                    HIPCHK(hipEventSynchronize(ev));
                    HIPCHK(hipEventRecord(ev, b->stream));
                }
            }
        }
        for (; done < nsteps; ++done)
            if ((r = launch_step(b, 15))) return r;
        return DSMC_OK;
    }
    return timed_steps(b, nsteps);
}

int dsmc_batch_sync(dsmc_batch b)
{
    if (!b) return fail(DSMC_ERR_ARG, "null batch");
    int r = set_device(b);
    if (r) return r;
    HIPCHK(hipStreamSynchronize(b->stream));
    return DSMC_OK;
}

int dsmc_batch_status(dsmc_batch b, int i, dsmc_status *o)
{
    if (!b || !o || i < -1 || i >= b->ninst) return fail(DSMC_ERR_ARG, "bad instance index or null argument");
    int r = set_device(b);
    if (r) return r;
    Ctrl c;
    std::vector<InstCtrl> ic;
    if ((r = read_ctrl(b, c, &ic))) return r;
    std::memset(o, 0, sizeof(*o));
    o->step = c.step;
    if (i < 0) {
        o->n_live = c.n_live; o->next_pid = 0; o->err = c.err;
        o->particle_steps = c.psteps; o->collisions = c.ncoll; o->last_collisions = c.last_coll;
        o->n_move = c.n_move;
        for (const InstCtrl &x : ic) { o->err |= x.err; o->nsamples = std::max<uint64_t>(o->nsamples, x.nsamples); }
        return DSMC_OK;
    }
    const InstCtrl &x = ic[i];
    o->n_live = x.n_live; o->next_pid = x.next_pid; o->err = x.err;
    o->particle_steps = x.psteps; o->collisions = x.ncoll; o->nsamples = x.nsamples;
    o->last_collisions = x.last_coll; o->n_move = x.n_move;
    return DSMC_OK;
}

int dsmc_batch_download_particles(dsmc_batch b, int i, int64_t cap, float *x, float *y, float *vx, float *vy,
                                  float *vz, uint32_t *pid, int64_t *n_out)
{
    if (!valid_index(b, i)) return fail(DSMC_ERR_ARG, "bad instance index");
    int r = set_device(b);
    if (r) return r;
    Ctrl c;
    std::vector<InstCtrl> ic;
    if ((r = read_ctrl(b, c, &ic))) return r;
    size_t off = 0;
    for (int k = 0; k < i; ++k) off += ic[k].n_live;
    const int64_t n = ic[i].n_live;
    if (n_out) *n_out = n;
    if (n > cap) return fail(DSMC_ERR_CAPACITY, "output arrays too small");
    const SoA &S = (c.step & 1u) ? b->B : b->A;
    const size_t bytes = sizeof(float) * (size_t)n;
    if (n > 0) {
        if (x) HIPCHK(hipMemcpyAsync(x, S.x + off, bytes, hipMemcpyDeviceToHost, b->stream));
        if (y) HIPCHK(hipMemcpyAsync(y, S.y + off, bytes, hipMemcpyDeviceToHost, b->stream));
        if (vx) HIPCHK(hipMemcpyAsync(vx, S.vx + off, bytes, hipMemcpyDeviceToHost, b->stream));
        if (vy) HIPCHK(hipMemcpyAsync(vy, S.vy + off, bytes, hipMemcpyDeviceToHost, b->stream));
        if (vz) HIPCHK(hipMemcpyAsync(vz, S.vz + off, bytes, hipMemcpyDeviceToHost, b->stream));
        if (pid) HIPCHK(hipMemcpyAsync(pid, S.pid + off, sizeof(uint32_t) * n, hipMemcpyDeviceToHost, b->stream));
    }
    HIPCHK(hipStreamSynchronize(b->stream));
    return DSMC_OK;
}

int dsmc_batch_cell_state(dsmc_batch b, int i, float *sigcmax, float *mrem, int *ncoll)
{
    if (!valid_index(b, i)) return fail(DSMC_ERR_ARG, "bad instance index");
    int r = set_device(b);
    if (r) return r;
    const int cb = b->cbase[i];
    const size_t Ci = b->cbase[i + 1] - cb;
    HIPCHK(hipStreamSynchronize(b->stream));
    if (sigcmax) HIPCHK(hipMemcpy(sigcmax, b->sigcmax + cb, sizeof(float) * Ci, hipMemcpyDeviceToHost));
    if (mrem) HIPCHK(hipMemcpy(mrem, b->mrem + cb, sizeof(float) * Ci, hipMemcpyDeviceToHost));
    if (ncoll) HIPCHK(hipMemcpy(ncoll, b->ncoll_cell + cb, sizeof(int) * Ci, hipMemcpyDeviceToHost));
    return DSMC_OK;
}

int dsmc_batch_reset_samples(dsmc_batch b)
{
    if (!b) return fail(DSMC_ERR_ARG, "null batch");
    int r = set_device(b);
    if (r) return r;
    Ctrl c;
    std::vector<InstCtrl> ic;
    if ((r = read_ctrl(b, c, &ic))) return r;
    for (InstCtrl &x : ic) x.nsamples = 0;
    HIPCHK(hipMemsetAsync(b->acc, 0, sizeof(double) * 6 * (size_t)b->ncell, b->stream));
    HIPCHK(hipMemcpyAsync(b->ictl, ic.data(), sizeof(InstCtrl) * b->ninst, hipMemcpyHostToDevice, b->stream));
    HIPCHK(hipStreamSynchronize(b->stream));
    return DSMC_OK;
}

int dsmc_batch_download_samples(dsmc_batch b, int i, double *acc, uint64_t *nsamples)
{
    if (!valid_index(b, i)) return fail(DSMC_ERR_ARG, "bad instance index");
    int r = set_device(b);
    if (r) return r;
    Ctrl c;
    std::vector<InstCtrl> ic;
    if ((r = read_ctrl(b, c, &ic))) return r;
    if (nsamples) *nsamples = ic[i].nsamples;
    if (acc) {
        const int cb = b->cbase[i];
        const size_t Ci = b->cbase[i + 1] - cb, C = b->ncell;
        for (int k = 0; k < 6; ++k)
            HIPCHK(hipMemcpy(acc + k * Ci, b->acc + k * C + cb, sizeof(double) * Ci, hipMemcpyDeviceToHost));
    }
    return DSMC_OK;
}

int dsmc_batch_download_counts(dsmc_batch b, double *counts, uint64_t *nsamples)
{
    if (!b || !counts) return fail(DSMC_ERR_ARG, "null argument");
    int r = set_device(b);
    if (r) return r;
    std::vector<InstCtrl> ic(b->ninst);
    HIPCHK(hipMemcpyAsync(counts, b->acc, sizeof(double) * (size_t)b->ncell, hipMemcpyDeviceToHost, b->stream));
    if (nsamples)
        HIPCHK(hipMemcpyAsync(ic.data(), b->ictl, sizeof(InstCtrl) * b->ninst, hipMemcpyDeviceToHost, b->stream));
    HIPCHK(hipStreamSynchronize(b->stream));
    if (nsamples)
        for (int i = 0; i < b->ninst; ++i) nsamples[i] = ic[i].nsamples;
    return DSMC_OK;
}

int dsmc_batch_download_moments(dsmc_batch b, double *sums, uint64_t *nsamples)
{
    if (!b || !sums) return fail(DSMC_ERR_ARG, "null argument");
    int r = set_device(b);
    if (r) return r;
    std::vector<InstCtrl> ic(b->ninst);
    // the first five accumulator blocks lie back to back over all cells
    HIPCHK(hipMemcpyAsync(sums, b->acc, sizeof(double) * 5 * (size_t)b->ncell, hipMemcpyDeviceToHost, b->stream));
    if (nsamples)
        HIPCHK(hipMemcpyAsync(ic.data(), b->ictl, sizeof(InstCtrl) * b->ninst, hipMemcpyDeviceToHost, b->stream));
    HIPCHK(hipStreamSynchronize(b->stream));
    if (nsamples)
        for (int i = 0; i < b->ninst; ++i) nsamples[i] = ic[i].nsamples;
    return DSMC_OK;
}

int dsmc_batch_wall_distance(dsmc_batch b, double *d)
{
    if (!b || !d) return fail(DSMC_ERR_ARG, "null argument");
    int r = set_device(b);
    if (r) return r;
    std::vector<WallInst> wi(b->ninst);
    int maxc = 1;
    for (int i = 0; i < b->ninst; ++i) {
        const dsmc_params &p = b->P[i].in;
        wi[i] = WallInst{ b->cbase[i], b->T[i].ncell, p.N2, b->poly0[i], b->npoly[i], p.L1 / p.N1, p.L2 / p.N2 };
        maxc = std::max(maxc, b->T[i].ncell);
    }
    WallInst *dw = nullptr;
    double *dp = nullptr, *dd = nullptr;
    hipError_t e = dalloc(&dw, wi.size());
    if (e == hipSuccess) e = dalloc(&dp, b->poly.size());
    if (e == hipSuccess) e = dalloc(&dd, (size_t)b->ncell);
    if (e == hipSuccess)
        e = hipMemcpyAsync(dw, wi.data(), sizeof(WallInst) * wi.size(), hipMemcpyHostToDevice, b->stream);
    if (e == hipSuccess && !b->poly.empty())
        e = hipMemcpyAsync(dp, b->poly.data(), sizeof(double) * b->poly.size(), hipMemcpyHostToDevice, b->stream);
    if (e == hipSuccess) e = launch_wall_distance(dw, b->ninst, maxc, dp, dd, b->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(d, dd, sizeof(double) * (size_t)b->ncell, hipMemcpyDeviceToHost, b->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(b->stream);
    (void)hipFree(dw);
    (void)hipFree(dp);
    (void)hipFree(dd);
    if (e != hipSuccess) return fail(DSMC_ERR_HIP, std::string("wall distance: ") + hipGetErrorString(e));
    return DSMC_OK;
}

// Keeps the instances keep[0..nkeep) (ascending) and drops the rest: a new
// set of device arrays sized for the kept instances (their host tables are
// copied, not rebuilt) receives their particle
// ranges, per-cell state (sigcmax, NTC remainders, collision counts, sample
// sums) and counters, then takes the handle's place.  Random streams are
// keyed on step, particle id and local cell id, none of which change, so a
// kept instance continues bit-identically.
int dsmc_batch_compact(dsmc_batch b, const int *keep, int nkeep)
{
    if (!b || !keep || nkeep < 1 || nkeep > b->ninst) return fail(DSMC_ERR_ARG, "compact: 1 <= nkeep <= n");
    for (int k = 0; k < nkeep; ++k)
        if (keep[k] < 0 || keep[k] >= b->ninst || (k > 0 && keep[k] <= keep[k - 1]))
            return fail(DSMC_ERR_ARG, "compact: keep must list instance indices in ascending order");
    int r = set_device(b);
    if (r) return r;
    Ctrl c;
    std::vector<InstCtrl> ic;
    if ((r = read_ctrl(b, c, &ic))) return r;
    if (nkeep == b->ninst) return DSMC_OK;
    dsmc_batch nb = new dsmc_batch_s();
    nb->ninst = nkeep;
    for (int k = 0; k < nkeep; ++k) {
        const int i = keep[k];
        nb->P.push_back(b->P[i]);
        nb->T.push_back(b->T[i]);
        nb->poly0.push_back((int)(nb->poly.size() / 2));
        nb->npoly.push_back(b->npoly[i]);
        const double *v = b->poly.data() + 2 * (size_t)b->poly0[i];
        nb->poly.insert(nb->poly.end(), v, v + 2 * (size_t)b->npoly[i]);
    }
    if ((r = setup_device(nb, &nb))) return r;
    nb->grid_move = b->grid_move;
    nb->grid_scatter = b->grid_scatter;
    nb->grid_cell = b->grid_cell;
    nb->tile_p = b->tile_p;
    nb->graph_len = b->graph_len;
    nb->timing = b->timing;

    const SoA &So = (c.step & 1u) ? b->B : b->A;
    const SoA &Sn = (c.step & 1u) ? nb->B : nb->A;
    std::vector<size_t> off(b->ninst + 1, 0);
    for (int i = 0; i < b->ninst; ++i) off[i + 1] = off[i] + ic[i].n_live;
    const size_t Co = b->ncell, Cn = nb->ncell;
    hipStream_t st = nb->stream;
    Ctrl g = c;
    g.psteps = g.ncoll = g.ncand = 0ull;
    g.last_coll = g.last_cand = 0u;
    std::vector<InstCtrl> nic(nkeep);
    size_t dst = 0;
    hipError_t e = hipSuccess;
    auto cp = [&](void *to, const void *from, size_t bytes) {
        if (e == hipSuccess && bytes) e = hipMemcpyAsync(to, from, bytes, hipMemcpyDeviceToDevice, st);
    };
    for (int k = 0; k < nkeep; ++k) {
        const int i = keep[k];
        const size_t n = ic[i].n_live, s = off[i];
        cp(Sn.x + dst, So.x + s, sizeof(float) * n);
        cp(Sn.y + dst, So.y + s, sizeof(float) * n);
        cp(Sn.vx + dst, So.vx + s, sizeof(float) * n);
        cp(Sn.vy + dst, So.vy + s, sizeof(float) * n);
        cp(Sn.vz + dst, So.vz + s, sizeof(float) * n);
        cp(Sn.pid + dst, So.pid + s, sizeof(uint32_t) * n);
        dst += n;
        const size_t co = b->cbase[i], cn = nb->cbase[k], Ci = b->T[i].ncell;
        cp(nb->sigcmax + cn, b->sigcmax + co, sizeof(float) * Ci);
        cp(nb->mrem + cn, b->mrem + co, sizeof(float) * Ci);
        cp(nb->ncoll_cell + cn, b->ncoll_cell + co, sizeof(int) * Ci);
        for (int q = 0; q < 6; ++q) cp(nb->acc + q * Cn + cn, b->acc + q * Co + co, sizeof(double) * Ci);
        nic[k] = ic[i];
        g.psteps += ic[i].psteps;
        g.ncoll += ic[i].ncoll;
        g.ncand += ic[i].ncand;
        g.last_coll += ic[i].last_coll;
        g.last_cand += ic[i].last_cand;
    }
    g.n_live = g.n_move = g.n_next = (uint32_t)dst;
    if (e == hipSuccess) e = hipMemcpyAsync(nb->ictl, nic.data(), sizeof(InstCtrl) * nkeep, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipMemcpyAsync(nb->ctrl, &g, sizeof(Ctrl), hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) {
        dsmc_batch_destroy(nb);
        return fail(DSMC_ERR_HIP, std::string("compact: ") + hipGetErrorString(e));
    }
    std::swap(*b, *nb);
    dsmc_batch_destroy(nb);
    return DSMC_OK;
}

int dsmc_batch_download_fields(dsmc_batch b, int i, double *n, double *ux, double *uy, double *uz, double *T,
                               double *p)
{
    if (!valid_index(b, i)) return fail(DSMC_ERR_ARG, "bad instance index");
    const size_t Ci = b->cbase[i + 1] - b->cbase[i];
    std::vector<double> a(6 * Ci);
    uint64_t ns = 0;
    int r = dsmc_batch_download_samples(b, i, a.data(), &ns);
    if (r) return r;
    fields_from_sums(b->P[i], b->T[i], a.data(), Ci, ns, n, ux, uy, uz, T, p);
    return DSMC_OK;
}

int dsmc_batch_get_launch(dsmc_batch b, dsmc_launch *o)
{
    if (!b || !o) return fail(DSMC_ERR_ARG, "null argument");
    o->move_blocks = b->grid_move;
    o->scatter_blocks = b->grid_scatter;
    o->cell_blocks = b->grid_cell;
    o->cell_tile = b->tile_p;
    o->graph_steps = b->graph_len;
    return DSMC_OK;
}

int dsmc_batch_set_launch(dsmc_batch b, const dsmc_launch *c)
{
    if (!b || !c) return fail(DSMC_ERR_ARG, "null argument");
    int r = check_launch(c);
    if (r) return r;
    b->grid_move = c->move_blocks;
    b->grid_scatter = c->scatter_blocks;
    b->grid_cell = c->cell_blocks;
    b->tile_p = c->cell_tile;
    b->graph_len = c->graph_steps;
    drop_graph(b);
    return DSMC_OK;
}

int dsmc_batch_set_timing(dsmc_batch b, int on)
{
    if (!b) return fail(DSMC_ERR_ARG, "null batch");
    if (on < 0 || on > 2) return fail(DSMC_ERR_ARG, "set_timing: 0 off, 1 events and bytes, 2 bytes only");
    b->timing = on;
    for (int q = 0; q < 4; ++q) b->t_ms[q] = b->t_bytes[q] = 0.0;
    b->t_launches = 0;
    return DSMC_OK;
}

int dsmc_batch_kernel_times(dsmc_batch b, double ms[4], double bytes[4], int64_t *launches)
{
    if (!b) return fail(DSMC_ERR_ARG, "null batch");
    for (int q = 0; q < 4; ++q) {
        if (ms) ms[q] = b->t_launches ? b->t_ms[q] / (double)b->t_launches : 0.0;
        if (bytes) bytes[q] = b->t_launches ? b->t_bytes[q] / (double)b->t_launches : 0.0;
    }
    if (launches) *launches = b->t_launches;
    return DSMC_OK;
}

}  // extern "C"
