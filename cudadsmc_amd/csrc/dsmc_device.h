// Device building blocks of the DSMC hot path on gfx950.
//
// Arithmetic contract (shared with the CPU oracle, which is why results are
// bit-identical): FP32 state, every + - * / sqrt correctly rounded and
// never fused (compile with -ffp-contract=off), log / exp / sin-cos(2 pi u)
// from the routines below (IEEE basic operations only), FP64 only for the
// sampling sums and the inflow remainder.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace dsmc {

enum { BC_STREAM = 0, BC_SPECULAR = 1, BC_DIFFUSE = 2 };
enum { SUB_MOVE = 1, SUB_INJECT = 2, SUB_COLLIDE = 3, SUB_INIT = 4 };
enum { ERR_BOUNCE = 1, ERR_INFLOW = 2, ERR_CAPACITY = 4, ERR_FILL = 8 };
constexpr int MAX_BOUNCE = 64;
constexpr int INFLOW_TRIES = 256;
constexpr int FILL_TRIES = 1024;
constexpr float SEG_TOL = 1e-5f;   // relative overlap at body segment ends
constexpr float SEG_BAND = 1e-4f;  // wall band (fraction of a cell) caught as an immediate hit

// Structure-of-arrays particle state (24 B per particle; 28 B with the
// rotational energy per unit mass er of a diatomic gas, else er is null).
struct SoA {
    float *x, *y, *vx, *vy, *vz;
    uint32_t *pid;
    float *er;
};

// Per-instance constants, passed by value to every kernel.
struct DevCase {
    int nx, ny, ncell, nseg;
    float L1, L2, dx, dy, dt;
    float ntc_coef, K1, e_exp, margin;
    float bx0, by0, bx1, by1;      // outside this box every cell's segment list is empty
    int body_bc;
    float body_vmp, body_sig;
    int face_bc[4];
    float face_vmp[4], face_sig[4];
    double inj_e[4];
    float inj_a[4];
    float inj_vmp, inj_sig, inj_ux, inj_uy;
    uint32_t seed, cap;
    int zeta;                      // rotational degrees of freedom: 0, or 2 (Larsen-Borgnakke)
    float rot_p, lb_ia, lb_ib;     // exchange fraction, 1/(5/2-omega), 1/(7/2-omega)
    float body_ekt, face_ekt[4], inj_ekt;   // kT/m of diffuse walls and the freestream
    const float *seg;
    const int *cseg_start, *cseg_idx;
    const float *vol_ntc;
    const float *fbox;      // 4 per cell: fluid-part bounding box of cut cells
    const uint64_t *fmask;  // per cell: 8x8 sub-boxes of fbox that touch fluid
    const uint32_t *sub;    // per cell: adaptive collision sub-cells (packed grid), null: none
};

// Adaptive collision sub-cells (Cell::genAMRmesh, Cell.cpp:168-279): a cell's
// grid is packed n2x | n2y << 8 | n3x << 16 | n3y << 24 (level 2, and level 3
// in every level-2 cell; 0: no grid), its leaves an nsx x nsy uniform grid,
// nsx = n2x n3x.  A cell with more than AMR_CAP particles or AMR_LEAVES
// leaves collides without sub-cells.
constexpr int AMR_MAX = 16;
constexpr int AMR_CAP = 2048;
constexpr int AMR_LEAVES = 256;

__host__ __device__ inline int amr_leaves(uint32_t s, int &nsx, int &nsy)
{
    if (!s) { nsx = 1; nsy = 1; return 1; }
    nsx = (int)(s & 0xffu) * (int)((s >> 16) & 0xffu);
    nsy = (int)((s >> 8) & 0xffu) * (int)((s >> 24) & 0xffu);
    return nsx * nsy;
}

// Whether cell c (k particles) collides in sub-cells.
__device__ __forceinline__ bool amr_cell(const DevCase &cs, int c, int k)
{
    if (!cs.sub || k < 2 || k > AMR_CAP) return false;
    int nsx, nsy;
    const int nl = amr_leaves(cs.sub[c], nsx, nsy);
    return nl > 1 && nl <= AMR_LEAVES;
}

// Device-resident step control: everything a step needs that changes
// between steps lives here, so a step has no host round trip and can be
// captured in a hipGraph.
struct Ctrl {
    uint32_t step, n_live, n_move, n_next;
    uint32_t next_pid, err, scan_done, cell_done;
    uint32_t step_coll, last_coll, step_cand, last_cand;
    uint32_t tile_next;            // k_cell work queue
    uint32_t n_runs;               // (cell, k_move wave) runs of the last binning
    unsigned long long psteps, ncoll, nsamples, ncand;
    unsigned long long nruns;      // runs binned over all steps (algorithmic byte counts)
};

// k_cell counts the end of a step in EXIT_GROUPS groups of blocks, each
// group on its own 4 KB line after the Ctrl (sums of collisions and
// candidates, blocks done), so that no single line takes an atomic from
// every block.  A Ctrl allocation is CTRL_ALLOC_BYTES, zeroed at creation;
// the last block of a group clears its line.
constexpr int EXIT_GROUPS = 16;
constexpr size_t EXIT_OFFSET = 4096, EXIT_STRIDE = 4096;
constexpr size_t CTRL_ALLOC_BYTES = EXIT_OFFSET + EXIT_GROUPS * EXIT_STRIDE;
constexpr size_t CTRL_ALLOC_COUNT = (CTRL_ALLOC_BYTES + sizeof(Ctrl) - 1) / sizeof(Ctrl);
static_assert(sizeof(Ctrl) <= EXIT_OFFSET, "the exit groups follow the Ctrl");
__host__ __device__ inline unsigned long long *exit_group(Ctrl *c, int g)
{
    return reinterpret_cast<unsigned long long *>(reinterpret_cast<char *>(c) + EXIT_OFFSET + (size_t)g * EXIT_STRIDE);
}

// Binning of one step.  k_move groups the lanes of each wave (64 consecutive
// particles, a "chunk") by cell: each group is a run, counted with one 64-bit
// atomic (runs << 32 | particles) per cell and recorded, compacted per chunk,
// as (cell, its run slot in the cell, lane mask).  One scan covers both
// counts; k_place lists every cell's runs contiguously, k_rank ranks a
// cell's runs by chunk and gives each its destination base, and k_cell
// expands the runs of its tile into the permutation (every particle at its
// run's base + its rank in the run): the stable (ascending index) order.
struct Bins {
    unsigned long long *cnt;       // per cell, zeroed again by the scan
    unsigned long long *lstart;    // per cell: exclusive scan of cnt within its scan tile
    unsigned long long *tsum, *toff;   // per scan tile: total, and exclusive scan of the totals
    int4 *runrec;                  // chunk q's runs at [64q, 64q + chunk_runs[q]): cell, slot, mask lo, hi
    int *chunk_runs;               // runs per chunk
    int4 *runlist;                 // runs in cell order: run id (64q + k), destination base (k_rank), mask lo, hi
};

// Batched runs: every instance keeps its own counters (the global Ctrl holds
// the totals that drive the shared scan, scatter and tile queue).
constexpr int MAX_INST = 256;
struct InstCtrl {
    uint32_t n_live, n_move, next_pid, err;
    uint32_t step_coll, step_cand, last_coll, last_cand;
    unsigned long long psteps, ncoll, nsamples, ncand;
};

// What k_cell needs of an instance, per cell.
struct CellConst {
    float ntc_coef, K1, e_exp;
    uint32_t seed;
    int sample_start;   // < 0: never
    int pad;
};

// ---------------------------------------------------------------------------
// Philox4x32-10 (Salmon et al., SC'11), counter-based: key = (id, seed),
// counter = (draw, step, substream, extra).
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint4 philox(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3,
                                        uint32_t k0, uint32_t k1)
{
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        if (r) { k0 += 0x9E3779B9u; k1 += 0xBB67AE85u; }
        const uint32_t lo0 = 0xD2511F53u * c0, hi0 = __umulhi(0xD2511F53u, c0);
        const uint32_t lo1 = 0xCD9E8D57u * c2, hi1 = __umulhi(0xCD9E8D57u, c2);
        const uint32_t n0 = hi1 ^ c1 ^ k0, n2 = hi0 ^ c3 ^ k1;
        c0 = n0; c1 = lo1; c2 = n2; c3 = lo0;
    }
    return make_uint4(c0, c1, c2, c3);
}

__device__ __forceinline__ float u01(uint32_t w) { return (float)(w >> 8) * 0x1p-24f; }
__device__ __forceinline__ float u01o(uint32_t w) { return (float)((w >> 8) + 1u) * 0x1p-24f; }
__device__ __forceinline__ int mulhi_idx(uint32_t w, int n) { return (int)__umulhi(w, (uint32_t)n); }

// ---------------------------------------------------------------------------
// log, exp, sin/cos(2 pi u) from IEEE basic operations.
// ---------------------------------------------------------------------------
#define DM_LN2_HI 0x1.62e400p-1f
#define DM_LN2_LO 0x1.7f7d1cp-20f
#define DM_LOG2E  0x1.715476p+0f
#define DM_PIO2   0x1.921fb6p+0f

// x = 2^e m, m in [sqrt(1/2), sqrt(2)); log m = 2 atanh((m-1)/(m+1)).
__device__ __forceinline__ float dm_logf(float x)
{
    uint32_t ix = __float_as_uint(x);
    int e = (int)((ix >> 23) & 0xffu) - 127;
    float m = __uint_as_float((ix & 0x007fffffu) | 0x3f800000u);
    if (m > 0x1.6a09e6p+0f) { m = m * 0.5f; e = e + 1; }
    float f = m - 1.0f;
    float s = f / (2.0f + f);
    float s2 = s * s;
    float p = s2 * (0x1.555556p-2f + s2 * (0x1.99999ap-3f + s2 * (0x1.24924ap-3f + s2 * 0x1.c71c72p-4f)));
    float two_s = 2.0f * s;
    float logm = two_s + two_s * p;
    float ef = (float)e;
    return ef * DM_LN2_HI + (ef * DM_LN2_LO + logm);
}

// x = k ln2 + r (Cody-Waite), e^r by Taylor to r^7.
__device__ __forceinline__ float dm_expf(float x)
{
    if (x < -87.0f) return 0.0f;
    if (x > 88.0f) x = 88.0f;
    float kf = floorf(x * DM_LOG2E + 0.5f);
    float r = (x - kf * DM_LN2_HI) - kf * DM_LN2_LO;
    float p = 1.0f + r * (1.0f + r * (0.5f + r * (0x1.555556p-3f + r * (0x1.555556p-5f +
              r * (0x1.111112p-7f + r * (0x1.6c16c2p-10f + r * 0x1.a01a02p-13f))))));
    int k = (int)kf;
    if (k > 127) return p * __uint_as_float((uint32_t)(127 + 127) << 23) * 2.0f;
    if (k < -126) return p * __uint_as_float((uint32_t)(k + 126 + 127) << 23) * 0x1p-126f;
    return p * __uint_as_float((uint32_t)(k + 127) << 23);
}

// sin, cos of 2 pi u, u in [0,1): exact quadrant reduction, Taylor on |t|<=pi/4.
__device__ __forceinline__ void dm_sincos2pi(float u, float &sn, float &cs)
{
    float t = u * 4.0f;
    float q = floorf(t + 0.5f);
    float f = t - q;
    float th = f * DM_PIO2;
    float t2 = th * th;
    float sp = th + th * t2 * (-0x1.555556p-3f + t2 * (0x1.111112p-7f + t2 * (-0x1.a01a02p-13f + t2 * 0x1.71de3ap-19f)));
    float cp = 1.0f + t2 * (-0.5f + t2 * (0x1.555556p-5f + t2 * (-0x1.6c16c2p-10f + t2 * (0x1.a01a02p-16f + t2 * -0x1.27e4fcp-22f))));
    int qi = ((int)q) & 3;
    if (qi == 0)      { sn = sp;  cs = cp; }
    else if (qi == 1) { sn = cp;  cs = -sp; }
    else if (qi == 2) { sn = -sp; cs = -cp; }
    else              { sn = -cp; cs = sp; }
}

// Cell index (cudaDSMC GPUCell.cu:184-189: i = (int)(x/dx), cell = j + i*ny),
// clamped into the grid.
__device__ __forceinline__ int cell_of(const DevCase &cs, float x, float y)
{
    int i = (int)(x / cs.dx);
    int j = (int)(y / cs.dy);
    i = i < 0 ? 0 : (i > cs.nx - 1 ? cs.nx - 1 : i);
    j = j < 0 ? 0 : (j > cs.ny - 1 ? cs.ny - 1 : j);
    return j + i * cs.ny;
}

// Leaf of a particle at (x, y) in cell c with an nsx x nsy leaf grid: the
// same FP32 division as cell_of, then the fraction of the cell scaled.
__device__ __forceinline__ int amr_leaf(const DevCase &cs, int c, float x, float y, int nsx, int nsy)
{
    const int i = c / cs.ny, j = c % cs.ny;
    const float fx = x / cs.dx - (float)i, fy = y / cs.dy - (float)j;
    int lx = (int)(fx * (float)nsx), ly = (int)(fy * (float)nsy);
    lx = lx < 0 ? 0 : (lx > nsx - 1 ? nsx - 1 : lx);
    ly = ly < 0 ? 0 : (ly > nsy - 1 ? nsy - 1 : ly);
    return ly + lx * nsy;
}

// Rotational energy per unit mass of a molecule with two rotational degrees
// of freedom in equilibrium at kT/m = ekt: -ln(U) ekt, exponential with mean
// ekt (cudaDSMC Random::RotationalenergySample, Random.cpp:48-51).
__device__ __forceinline__ float rot_energy(uint32_t w, float ekt) { return -dm_logf(u01o(w)) * ekt; }

// Diffuse re-emission from a wall with unit normal (nx, ny) into the fluid:
// flux-weighted normal speed, independent Gaussian tangential and z parts;
// with rot, the rotational energy accommodated to the wall (kTw/m = ekt).
__device__ __forceinline__ void diffuse_emit(float nx, float ny, float vmp, float sig, float ekt, bool rot,
                                             uint32_t pid, uint32_t seed, uint32_t step, uint32_t b,
                                             float &vx, float &vy, float &vz, float &er)
{
    uint4 w = philox(b, step, SUB_MOVE, 0u, pid, seed);
    float vn = sqrtf(-dm_logf(u01o(w.x))) * vmp;
    float r = sqrtf(-2.0f * dm_logf(u01o(w.y))) * sig;
    float s, c;
    dm_sincos2pi(u01(w.z), s, c);
    float vt = r * c;
    vx = vn * nx + vt * (-ny);
    vy = vn * ny + vt * nx;
    vz = r * s;
    if (rot) er = rot_energy(w.w, ekt);
}

// Flight of duration t with wall events (faces and body segments), earliest
// event first.  Returns true if the particle left through a stream face.
// pid is only read (from *pidp) when a diffuse wall needs random numbers.
// vchg reports a changed velocity; with zeta = 2 a diffuse wall re-draws er
// as well, and echg reports that (er is only written then).
__device__ __forceinline__ bool move_one(const DevCase &cs, float &x, float &y, float &vx, float &vy,
                                         float &vz, float &er, const uint32_t *pidp, uint32_t pid_known,
                                         bool have_pid, uint32_t step, float t, uint32_t &err,
                                         bool &vchg, bool &echg)
{
    const bool rot = cs.zeta != 0;
    bool dead = false;
    for (int b = 0;; ++b) {
        if (b == MAX_BOUNCE) { err |= ERR_BOUNCE; break; }
        float best = t;
        int kind = -1;
        // A face's time s = d / |v| (d the distance to it) is computed only
        // when it can beat best: for best > 0, d > |v| best (1 + 1e-4)
        // implies fl(d / |v|) >= best, so skipping changes no outcome.
        if (vx < 0.0f) {
            const float d = 0.0f - x;
            if (!(best > 0.0f && -d > -vx * best * 1.0001f)) {
                float s = d / vx;
                if (s < best) { best = s; kind = 0; }
            }
        } else if (vx > 0.0f) {
            const float d = cs.L1 - x;
            if (!(best > 0.0f && d > vx * best * 1.0001f)) {
                float s = d / vx;
                if (s < best) { best = s; kind = 1; }
            }
        }
        if (vy < 0.0f) {
            const float d = 0.0f - y;
            if (!(best > 0.0f && -d > -vy * best * 1.0001f)) {
                float s = d / vy;
                if (s < best) { best = s; kind = 2; }
            }
        } else if (vy > 0.0f) {
            const float d = cs.L2 - y;
            if (!(best > 0.0f && d > vy * best * 1.0001f)) {
                float s = d / vy;
                if (s < best) { best = s; kind = 3; }
            }
        }
        if (cs.nseg > 0) {
            // reach = sqrt(v^2) t against the margin; the square decides
            // unless it lies within 0.1 % of margin^2
            const float v2 = vx * vx + vy * vy;
            const float r2 = v2 * t * t, m2 = cs.margin * cs.margin;
            const bool all = r2 > 1.001f * m2 ? true : (r2 < 0.999f * m2 ? false : sqrtf(v2) * t > cs.margin);
            int cnt, c0 = 0;
            if (all) {
                cnt = cs.nseg;
            } else if (fminf(fmaxf(x, 0.0f), cs.L1) < cs.bx0 || fminf(fmaxf(x, 0.0f), cs.L1) > cs.bx1 ||
                       fminf(fmaxf(y, 0.0f), cs.L2) < cs.by0 || fminf(fmaxf(y, 0.0f), cs.L2) > cs.by1) {
                cnt = 0;   // the clamped position's cell lists no segment
            } else {
                int c = cell_of(cs, x, y);
                c0 = cs.cseg_start[c];
                cnt = cs.cseg_start[c + 1] - c0;
            }
            for (int q = 0; q < cnt; ++q) {
                int k = all ? q : cs.cseg_idx[c0 + q];
                const float *S = cs.seg + 6 * k;
                float snx = S[4], sny = S[5];
                float vn = vx * snx + vy * sny;
                if (!(vn < 0.0f)) continue;
                float ax = S[0], ay = S[1];
                float ex = S[2] - ax, ey = S[3] - ay;
                float den = vx * ey - vy * ex;
                if (den == 0.0f) continue;
                float wx = ax - x, wy = ay - y;
                float s = (wx * ey - wy * ex) / den;
                if (!(s < best)) continue;
                if (!(s > 0.0f)) {
                    // on or just behind the wall line (a previous step's hit
                    // point rounded to the solid side) and moving in: hit now,
                    // at the start of a flight only.  After a reflection the
                    // particle leaves the face it hit, which for a convex body
                    // is leaving the body; a band hit on the neighbouring face
                    // of an acute vertex would trap it there.
                    if (b > 0) continue;
                    float d = -(wx * snx + wy * sny);
                    if (d < -(cs.margin * SEG_BAND)) continue;
                    s = 0.0f;
                }
                // small overlap at the segment ends: a ray through a shared
                // vertex must not slip between two rounded tests
                float u = (wx * vy - wy * vx) / den;
                if (u < -SEG_TOL || u > 1.0f + SEG_TOL) continue;
                // past the segment's own end (in the overlap): a hit only for
                // a flight that also enters the solid side of the segment
                // sharing that vertex (one closed polygon, segment k from
                // vertex k to k+1).  The overlaps of the two faces of an acute
                // vertex would otherwise bound a thin region outside the body
                // in which a re-emitted particle bounces without end.
                if (u < 0.0f || u > 1.0f) {
                    const int nb = u > 1.0f ? (k + 1 == cs.nseg ? 0 : k + 1) : (k == 0 ? cs.nseg - 1 : k - 1);
                    const float *Q = cs.seg + 6 * nb;
                    if (!(vx * Q[4] + vy * Q[5] < 0.0f)) continue;
                }
                best = s;
                kind = 4 + k;
            }
        }
        if (kind < 0) {
            x = x + vx * t;
            y = y + vy * t;
            break;
        }
        x = x + vx * best;
        y = y + vy * best;
        t = t - best;
        vchg = true;
        if (kind < 4) {
            if (kind == 0) x = 0.0f;
            else if (kind == 1) x = cs.L1;
            else if (kind == 2) y = 0.0f;
            else y = cs.L2;
            int bc = cs.face_bc[kind];
            if (bc == BC_STREAM) { dead = true; break; }
            if (bc == BC_SPECULAR) {
                if (kind < 2) vx = -vx; else vy = -vy;
            } else {
                float nx = kind == 0 ? 1.0f : (kind == 1 ? -1.0f : 0.0f);
                float ny = kind == 2 ? 1.0f : (kind == 3 ? -1.0f : 0.0f);
                uint32_t pid = have_pid ? pid_known : *pidp;
                diffuse_emit(nx, ny, cs.face_vmp[kind], cs.face_sig[kind], cs.face_ekt[kind], rot, pid, cs.seed,
                             step, (uint32_t)b, vx, vy, vz, er);
                echg = rot;
            }
        } else {
            const float *S = cs.seg + 6 * (kind - 4);
            float snx = S[4], sny = S[5];
            if (cs.body_bc == BC_SPECULAR) {
                float vn = vx * snx + vy * sny;
                vx = vx - 2.0f * vn * snx;
                vy = vy - 2.0f * vn * sny;
            } else {
                uint32_t pid = have_pid ? pid_known : *pidp;
                diffuse_emit(snx, sny, cs.body_vmp, cs.body_sig, cs.body_ekt, rot, pid, cs.seed, step,
                             (uint32_t)b, vx, vy, vz, er);
                echg = rot;
            }
        }
    }
    return dead;
}

// Inflow particles created at `step` on each stream face: exact remainder
// carry floor((s+1) e) - floor(s e) of the expected flux e per step.
__device__ __forceinline__ void inject_counts(const DevCase &cs, uint32_t step, int J[4])
{
#pragma unroll
    for (int f = 0; f < 4; ++f) {
        double e = cs.inj_e[f];
        J[f] = (cs.face_bc[f] == BC_STREAM && e > 0.0)
                   ? (int)(floor(((double)step + 1.0) * e) - floor((double)step * e))
                   : 0;
    }
}

// One inflow particle on face f: position on the face, drifting-Maxwellian
// flux-weighted velocity, then flight for a random fraction of dt.
__device__ __forceinline__ bool inject_one(const DevCase &cs, int f, uint32_t pid, uint32_t step,
                                           float &x, float &y, float &vx, float &vy, float &vz, float &er,
                                           uint32_t &err)
{
    uint4 w = philox(0u, step, SUB_INJECT, 0u, pid, cs.seed);
    float upos = u01(w.x);
    float ufrac = u01(w.y);
    float r = sqrtf(-2.0f * dm_logf(u01o(w.z)));
    float sn, cn;
    dm_sincos2pi(u01(w.w), sn, cn);
    float g0 = r * cn, g1 = r * sn;

    float a = cs.inj_a[f];
    float lo, hi;
    if (a >= -3.5f) {
        lo = a - 3.5f;
        if (lo < 0.0f) lo = 0.0f;
        hi = (a > 0.0f ? a : 0.0f) + 3.5f;
    } else {
        lo = 0.0f;
        hi = -4.0f / a;
    }
    float ws = 0.5f * (a + sqrtf(a * a + 2.0f));
    float dws = ws - a;
    float ref = dws * dws;
    float wn = lo;
    bool ok = false;
    for (int tr = 0; tr < INFLOW_TRIES && !ok; tr += 2) {
        uint4 q = philox(1u + (uint32_t)(tr / 2), step, SUB_INJECT, 0u, pid, cs.seed);
        float cand = lo + (hi - lo) * u01(q.x);
        float d = cand - a;
        float acc = (cand / ws) * dm_expf(ref - d * d);
        wn = cand;
        if (u01(q.y) < acc) { ok = true; break; }
        cand = lo + (hi - lo) * u01(q.z);
        d = cand - a;
        acc = (cand / ws) * dm_expf(ref - d * d);
        wn = cand;
        if (u01(q.w) < acc) { ok = true; break; }
    }
    if (!ok) err |= ERR_INFLOW;
    float vn = wn * cs.inj_vmp;
    vz = cs.inj_sig * g1;
    // faces 0, 1 (x = 0, L1): normal along x; faces 2, 3 (y = 0, L2): along y;
    // odd faces point the normal speed inward as -vn (selects, not branches)
    const bool xf = f < 2;
    const float vin = (f & 1) ? -vn : vn;
    const float vtan = (xf ? cs.inj_uy : cs.inj_ux) + cs.inj_sig * g0;
    const float wall = (f & 1) ? (xf ? cs.L1 : cs.L2) : 0.0f;
    x = xf ? wall : upos * cs.L1;
    y = xf ? upos * cs.L2 : wall;
    vx = xf ? vin : vtan;
    vy = xf ? vtan : vin;
    // rotational energy of the freestream (GPUParticle.cu:409 draws it and
    // leaves it unstored, D13)
    er = 0.0f;
    if (cs.zeta != 0) er = rot_energy(philox(0u, step, SUB_INJECT, 1u, pid, cs.seed).x, cs.inj_ekt);
    bool vchg = false, echg = false;
    return move_one(cs, x, y, vx, vy, vz, er, nullptr, pid, true, step, ufrac * cs.dt, err, vchg, echg);
}

// Cross-lane moves within a row of 16 lanes by DPP (a VALU operand
// modifier, no LDS round trip).  row_shr:k (0x110 + k) gives lane i the
// value of lane i - k of its row; row_bcast:15 / :31 (0x142 / 0x143, rows
// chosen by the row mask) give a row lane 15 / 31.  A lane whose source is
// outside the row, or whose row is masked, gets 0.
template <int CTRL, int ROWS = 0xf>
__device__ __forceinline__ int dpp_i(int x)
{
    return __builtin_amdgcn_update_dpp(0, x, CTRL, ROWS, 0xf, true);
}
template <int CTRL, int ROWS = 0xf>
__device__ __forceinline__ float dpp_f(float x)
{
    return __int_as_float(dpp_i<CTRL, ROWS>(__float_as_int(x)));
}

// Inclusive prefix sums over the wave (0 from outside a row adds nothing).
__device__ __forceinline__ int wave_prefix_sum(int x)
{
    x += dpp_i<0x111>(x);
    x += dpp_i<0x112>(x);
    x += dpp_i<0x114>(x);
    x += dpp_i<0x118>(x);
    x += dpp_i<0x142, 0xa>(x);
    x += dpp_i<0x143, 0xc>(x);
    return x;
}
template <int CTRL, int ROWS = 0xf>
__device__ __forceinline__ unsigned long long dpp_u64(unsigned long long x)
{
    const unsigned lo = (unsigned)dpp_i<CTRL, ROWS>((int)(unsigned)x);
    const unsigned hi = (unsigned)dpp_i<CTRL, ROWS>((int)(unsigned)(x >> 32));
    return ((unsigned long long)hi << 32) | lo;
}
__device__ __forceinline__ unsigned long long wave_prefix_sum_u64(unsigned long long x)
{
    x += dpp_u64<0x111>(x);
    x += dpp_u64<0x112>(x);
    x += dpp_u64<0x114>(x);
    x += dpp_u64<0x118>(x);
    x += dpp_u64<0x142, 0xa>(x);
    x += dpp_u64<0x143, 0xc>(x);
    return x;
}

// A wave-uniform lane's value (readlane: no LDS round trip).
__device__ __forceinline__ int lane_value(int x, int l) { return __builtin_amdgcn_readlane(x, l); }

}  // namespace dsmc
