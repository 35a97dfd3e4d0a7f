// k_cell: the per-cell half of a DSMC step on gfx950.
//
// One block (4 waves of 64) takes a tile of consecutive cells whose particles
// occupy [p0, p0+n) of the cell-sorted order; tiles of about CELL_P particles
// are handed out by a device-wide queue (n <= CELL_CAP: staged in LDS;
// larger tiles stage in the destination arrays in HBM).  Per tile:
//   0. the tile's k_move runs, ranked within their cells by k_rank (each
//      run's destination base), expanded into the permutation: every cell's
//      particles in ascending index order (the order of the serial binning,
//      sortParticles, GPUParticle.cu:365-373);
//   1. gather of the state into cell order -- the outflow compaction of
//      GPUParticles::DeleteInvalid (cudaDSMC GPUParticle.cu:76-140);
//   2. NTC/VHS collisions (collisionInCells, GPUCell.cu:200-333, with the
//      host semantics of Cell::collision, Cell.cpp:335-409): candidate counts
//      one lane per cell; cells with up to 64 candidates run in groups of 8
//      lanes, heavier cells by a whole wave, both in rounds that end at the
//      first candidate sharing a particle with an earlier accepted one,
//      whose results equal the serial candidate loop bit for bit;
//   3. FP64 sampling sums, 8 lanes per cell joined by a fixed butterfly
//      (samplingInCells, GPUCell.cu:336-402, as time-averaged accumulators);
//   4. velocities and cell state written back coalesced.
// The last block to finish advances the device step control.
#include <type_traits>

#include "dsmc_kernels.h"

namespace dsmc {

namespace {

constexpr int CELL_MAXC = 128;       // cells whose metadata is staged at once
constexpr int CELL_THREADS = 64 * CELL_WAVES;   // k_cell's block
static_assert(CELL_MAXC + 1 <= CELL_THREADS, "one thread per staged cell start");
constexpr int NTC_G = 8;             // lanes per cell in the grouped NTC rounds
constexpr int NTC_GROUP_MAX = 64;    // above this many candidates a whole wave takes the cell
constexpr int GATHER_UNROLL = 4;     // particles per thread with loads in flight

__device__ __forceinline__ int lane_id() { return (int)__lane_id(); }

// Inclusive prefix max over the wave of values >= 0 (a 0 from outside a
// row leaves a max unchanged).
__device__ __forceinline__ float wave_prefix_max(float x)
{
    x = fmaxf(x, dpp_f<0x111>(x));
    x = fmaxf(x, dpp_f<0x112>(x));
    x = fmaxf(x, dpp_f<0x114>(x));
    x = fmaxf(x, dpp_f<0x118>(x));
    x = fmaxf(x, dpp_f<0x142, 0xa>(x));
    x = fmaxf(x, dpp_f<0x143, 0xc>(x));
    return x;
}

// The same within aligned segments of G lanes (G = 8 or 16: within a row),
// sub = lane % G.
template <int G>
__device__ __forceinline__ float seg_prefix_max(float x, int sub)
{
    static_assert(G == 8 || G == 16, "segments within a row of 16 lanes");
    float t = dpp_f<0x111>(x);
    if (sub >= 1) x = fmaxf(x, t);
    t = dpp_f<0x112>(x);
    if (sub >= 2) x = fmaxf(x, t);
    t = dpp_f<0x114>(x);
    if (sub >= 4) x = fmaxf(x, t);
    if constexpr (G == 16) {
        t = dpp_f<0x118>(x);
        if (sub >= 8) x = fmaxf(x, t);
    }
    return x;
}

// Whether an earlier lane of the segment (lane - k, 1 <= k <= sub) holds a1
// or a2 equal to i1 or i2.
template <int K, int G>
__device__ __forceinline__ bool seg_earlier_hit(int a1, int a2, int i1, int i2, int sub)
{
    if constexpr (K >= G) {
        return false;
    } else {
        const int a = dpp_i<0x110 + K>(a1), b = dpp_i<0x110 + K>(a2);
        const bool hit = (K <= sub) & ((a == i1) | (a == i2) | (b == i1) | (b == i2));
        return hit || seg_earlier_hit<K + 1, G>(a1, a2, i1, i2, sub);
    }
}

// The value of lane ^ D (D = 1, 2 or 4: within a segment of 8 lanes), by
// DPP with every lane active: quad_perm [1,0,3,2] (0xb1) and [2,3,0,1]
// (0x4e) for 1 and 2; for 4, row_half_mirror (0x141: lane i of a segment
// takes lane 7 - i) and then quad_perm [3,2,1,0] (0x1b: i takes i ^ 3),
// since 7 - (i ^ 3) = i ^ 4.  Exactly what __shfl_xor(x, D, 8) returns.
// (A DPP move must not sit in a branch: the lanes it reads from would be
// inactive there and read as 0.)
template <int D>
__device__ __forceinline__ int xor_lane_i(int x)
{
    static_assert(D == 1 || D == 2 || D == 4, "a partner within 8 lanes");
    if constexpr (D == 1) return dpp_i<0xb1>(x);
    else if constexpr (D == 2) return dpp_i<0x4e>(x);
    else return dpp_i<0x1b>(dpp_i<0x141>(x));
}
template <int D>
__device__ __forceinline__ double xor_lane_d(double x)
{
    const unsigned long long b = (unsigned long long)__double_as_longlong(x);
    const unsigned lo = (unsigned)xor_lane_i<D>((int)(unsigned)b);
    const unsigned hi = (unsigned)xor_lane_i<D>((int)(unsigned)(b >> 32));
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// 0, computed from v by an instruction the compiler cannot see through: a
// use of v that makes the wave wait for v (an atomic's old value)
__device__ __forceinline__ uint32_t opaque_zero(uint32_t v)
{
    uint32_t z;
    asm volatile("v_mul_lo_u32 %0, %1, 0" : "=v"(z) : "v"(v));
    return z;
}

// Called by one thread of every block at the block's end: adds (coll,
// cand) into the block's exit group (blocks g, g + ng, ... with ng =
// min(grid, EXIT_GROUPS), each group on its own line) and counts the block
// done there.  dep carries old values of earlier atomics that must be
// performed first.  In the last block of a group it returns true with the
// group's sums in tot (coll | cand << 32) and the line cleared for the next
// launch.
__device__ bool exit_group_add(Ctrl *ctrl, uint32_t coll, uint32_t cand, uint32_t dep, unsigned long long &tot)
{
    const uint32_t ng = min(gridDim.x, (uint32_t)EXIT_GROUPS);
    const uint32_t g = blockIdx.x % ng;
    const uint32_t gsize = (gridDim.x - g + ng - 1u) / ng;
    unsigned long long *eg = exit_group(ctrl, (int)g);
    uint32_t *const gdone = reinterpret_cast<uint32_t *>(eg + 1);
    const unsigned long long o = atomicAdd(eg, (unsigned long long)coll | ((unsigned long long)cand << 32));
    dep |= (uint32_t)o | (uint32_t)(o >> 32);
    if (atomicAdd(gdone, 1u + opaque_zero(dep)) != gsize - 1u) return false;
    tot = atomicExch(eg, 0ull);
    atomicExch(gdone, 0u);
    return true;
}

struct CellShared {
    int st[CELL_MAXC + 1];           // cell starts relative to the tile
    float sm[CELL_MAXC], mr[CELL_MAXC], vol[CELL_MAXC];
    int nc[CELL_MAXC], M[CELL_MAXC];
    int big[CELL_MAXC], heavy[CELL_MAXC];
    int inst[CELL_MAXC];             // batched runs: instance of each staged cell
    int nbig, nheavy;
    uint32_t coll, cand;
};

template <bool GLOBAL>
__device__ __forceinline__ void wave_sync()
{
    if (GLOBAL) __threadfence_block();
    else __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

// Collision constants of one cell: VHS sigma*c_r = K1 (c_r^2)^e_exp, the
// Philox key (instance-local cell id, instance seed) and, for a diatomic gas,
// the Larsen-Borgnakke exchange fraction and exponents.
struct CK {
    float K1, e_exp;
    uint32_t seed;
    int lc;
    float rot_p, lb_ia, lb_ib;
};

// Number of NTC candidates of a cell with remainder carry:
// M = (int)(np(np-1) Fn dt/2 (sigma c)max / V + r),  r' = that - M.
__device__ __forceinline__ int ntc_count(float ntc_coef, int np, float vol, float smax, float &mrem)
{
    float e = (float)((long long)np * (long long)(np - 1)) * ntc_coef;
    e = e * smax;
    e = e / vol;
    const float tot = e + mrem;
    const int M = (int)tot;
    mrem = tot - (float)M;
    return M;
}

// Post-collision velocities: isotropic (VHS) scattering of the relative
// velocity about the centre of mass (Particle.cpp:84-100, GPUCell.cu:278-310).
// The scattering direction and |g| depend only on the candidate's random
// numbers and velocities, not on whether it is accepted, so the rounds
// compute them before the acceptance test (it then overlaps the
// cross-section chain); scatter_pair applies them to an accepted pair.
struct Scatter {
    float cx, cy, cz;   // unit direction of the scattered relative velocity
    float g;            // its magnitude (|g|, or the Larsen-Borgnakke g')
};

__device__ __forceinline__ Scatter scatter_dir(const CK &k, uint32_t step, int m, uint32_t w3, float g2)
{
    const float cosr = 2.0f * u01(w3) - 1.0f;
    const float sinr = sqrtf(1.0f - cosr * cosr);
    const uint4 w2 = philox((uint32_t)m, step, SUB_COLLIDE, 1u, (uint32_t)k.lc, k.seed);
    float sp, cp;
    dm_sincos2pi(u01(w2.x), sp, cp);
    return Scatter{ cosr, sinr * cp, sinr * sp, sqrtf(g2) };
}

__device__ __forceinline__ void scatter_pair(const Scatter &r, float &v1x, float &v1y, float &v1z, float &v2x,
                                             float &v2y, float &v2z)
{
    const float rx = r.g * r.cx, ry = r.g * r.cy, rz = r.g * r.cz;
    const float mx = 0.5f * (v1x + v2x), my = 0.5f * (v1y + v2y), mz = 0.5f * (v1z + v2z);
    v1x = mx + 0.5f * rx; v2x = mx - 0.5f * rx;
    v1y = my + 0.5f * ry; v2y = my - 0.5f * ry;
    v1z = mz + 0.5f * rz; v2z = mz - 0.5f * rz;
}

// Larsen-Borgnakke exchange of an accepted pair of a diatomic gas (zeta = 2;
// the scheme of GPUCell.cu:256-275,316-328, written there and commented out):
// with probability rot_p the collision energy per unit mass
// E = g^2/4 + e1 + e2 (reduced mass m/2) is split into translation, fraction
// x ~ Beta(5/2 - omega, 2) drawn as U1^(1/(5/2-omega)) U2^(1/(7/2-omega)),
// and rotation, shared uniformly between the two; |g'| = 2 sqrt(x E).
__device__ __forceinline__ void lb_exchange(const CK &k, uint32_t step, int m, float g2, float &e1, float &e2,
                                            Scatter &sd)
{
    const uint4 w = philox((uint32_t)m, step, SUB_COLLIDE, 2u, (uint32_t)k.lc, k.seed);
    if (!(u01(w.x) < k.rot_p)) return;
    const float ec = 0.25f * g2 + (e1 + e2);
    const float x = dm_expf(k.lb_ia * dm_logf(u01o(w.y)) + k.lb_ib * dm_logf(u01o(w.z)));
    const float et = x * ec;
    const float erot = ec - et;
    const float a = erot * u01(w.w);
    e1 = a;
    e2 = erot - a;
    sd.g = 2.0f * sqrtf(et);
}

__device__ __forceinline__ float vhs_sigc(const CK &k, float g2)
{
    return g2 > 0.0f ? k.K1 * dm_expf(k.e_exp * dm_logf(g2)) : 0.0f;
}

// The serial candidate loop of many cells at once: a wave takes 64/G cells,
// G lanes per cell, and runs rounds of up to G candidates in each cell (the
// same rule as ntc_rounds, within G-lane segments).  list[]
// holds the chunk-local cell numbers; results go to sh.sm / sh.nc.
template <int G, bool GLOBAL, bool ROT, class KF>
__device__ void ntc_groups(const KF &kf, uint32_t step, int c0, const int *list, int nlist, CellShared &sh,
                           float *Vx, float *Vy, float *Vz, float *Er, int wave, int nwaves)
{
    const int lane = lane_id(), g = lane / G, sub = lane % G;
    const unsigned long long gmask = (G == 64 ? ~0ull : ((1ull << G) - 1ull)) << (g * G);
    constexpr int CPW = 64 / G;
    for (int base = wave * CPW; base < nlist; base += nwaves * CPW) {
        const int li = base + g;
        const bool has = li < nlist;
        const int i = has ? list[li] : 0;
        const int o = has ? sh.st[i] : 0;
        const int np = has ? sh.st[i + 1] - o : 2;
        const int M = has ? sh.M[i] : 0;
        const CK k = kf.key(i, c0 + i);
        float *vx = Vx + o, *vy = Vy + o, *vz = Vz + o;
        float *er = ROT ? Er + o : nullptr;
        float smax = has ? sh.sm[i] : 0.0f;
        int nc = 0, mb = 0;
        while (__ballot(mb < M)) {
            const int nact = mb < M ? min(G, M - mb) : 0;
            const bool act = sub < nact;
            const int m = mb + sub;
            int i1 = -1, i2 = -2;
            uint4 w = make_uint4(0u, 0u, 0u, 0u);
            if (act) {
                w = philox((uint32_t)m, step, SUB_COLLIDE, 0u, (uint32_t)k.lc, k.seed);
                i1 = mulhi_idx(w.x, np);
                i2 = mulhi_idx(w.y, np - 1);
                if (i2 >= i1) i2 = i2 + 1;
            }
            float sc = 0.0f, g2 = 0.0f;
            float v1x = 0.f, v1y = 0.f, v1z = 0.f, v2x = 0.f, v2y = 0.f, v2z = 0.f;
            Scatter sd{ 0.f, 0.f, 0.f, 0.f };
            if (act) {
                v1x = vx[i1]; v1y = vy[i1]; v1z = vz[i1];
                v2x = vx[i2]; v2y = vy[i2]; v2z = vz[i2];
                const float gx = v1x - v2x, gy = v1y - v2y, gz = v1z - v2z;
                g2 = gx * gx + gy * gy + gz * gz;
                sc = vhs_sigc(k, g2);
                sd = scatter_dir(k, step, m, w.w, g2);
            }
            const float pm = seg_prefix_max<G>(sc, sub);
            const float sml = fmaxf(smax, pm);
            const bool acc = act && (u01(w.z) < sc / sml);
            // the round ends before the first candidate sharing a particle
            // with an earlier accepted one (a rejected candidate changes no
            // velocity, so its neighbours in the round read current values)
            const int a1 = acc ? i1 : -3, a2 = acc ? i2 : -4;
            const bool conf = seg_earlier_hit<1, G>(a1, a2, i1, i2, sub);
            const unsigned long long cm = __ballot(act && conf) & gmask;
            const int f = cm ? (int)__ffsll((long long)cm) - 1 - g * G : nact;
            const bool accd = acc && sub < f;
            if (accd) {
                if constexpr (ROT) {
                    float e1 = er[i1], e2 = er[i2];
                    lb_exchange(k, step, m, g2, e1, e2, sd);
                    er[i1] = e1; er[i2] = e2;
                }
                scatter_pair(sd, v1x, v1y, v1z, v2x, v2y, v2z);
                vx[i1] = v1x; vy[i1] = v1y; vz[i1] = v1z;
                vx[i2] = v2x; vy[i2] = v2y; vz[i2] = v2z;
            }
            nc += __popcll(__ballot(accd) & gmask);
            const float last = __shfl(pm, f > 0 ? f - 1 : 0, G);
            if (f > 0) smax = fmaxf(smax, last);
            mb += f;
            wave_sync<GLOBAL>();
        }
        if (has && sub == 0) { sh.sm[i] = smax; sh.nc[i] = nc; }
    }
}

// The same loop by a whole wave: each round takes the longest prefix of the
// next <= 64 candidates sharing no particle with an earlier accepted
// candidate of the round; (sigma c)max becomes a prefix max.  Within that
// prefix every candidate reads the velocities the serial loop would, so the
// results are identical.
// GLOBAL: velocities in HBM; LMARK: conflict marks in LDS (else found by
// shuffles).
template <bool GLOBAL, bool ROT, bool LMARK = !GLOBAL>
__device__ int ntc_rounds(const CK &k, uint32_t step, int np, int M, float *Vx, float *Vy,
                          float *Vz, float *Er, float &smax, int *mark)
{
    const int lane = lane_id();
    int nc = 0;
    // a whole-wave cell is the longest serial chain of its tile (on a
    // body's stagnation side, of the launch): its wave issues first on its
    // SIMD (the instruction priority only; no result depends on it)
    __builtin_amdgcn_s_setprio(3);
    if (LMARK) {
        // mark[q] = lowest accepted lane of the round that uses particle q
        // (64 = none)
        for (int q = lane; q < np; q += 64) mark[q] = 64;
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        __builtin_amdgcn_wave_barrier();
    }
    for (int mb = 0; mb < M;) {
        const int m = mb + lane;
        const int nact = min(64, M - mb);
        const bool act = lane < nact;
        int i1 = -1, i2 = -2;
        uint4 w = make_uint4(0u, 0u, 0u, 0u);
        if (act) {
            w = philox((uint32_t)m, step, SUB_COLLIDE, 0u, (uint32_t)k.lc, k.seed);
            i1 = mulhi_idx(w.x, np);
            i2 = mulhi_idx(w.y, np - 1);
            if (i2 >= i1) i2 = i2 + 1;
        }
        float sc = 0.0f, g2 = 0.0f;
        float v1x = 0.f, v1y = 0.f, v1z = 0.f, v2x = 0.f, v2y = 0.f, v2z = 0.f;
        Scatter sd{ 0.f, 0.f, 0.f, 0.f };
        if (act) {
            v1x = Vx[i1]; v1y = Vy[i1]; v1z = Vz[i1];
            v2x = Vx[i2]; v2y = Vy[i2]; v2z = Vz[i2];
            float gx = v1x - v2x, gy = v1y - v2y, gz = v1z - v2z;
            g2 = gx * gx + gy * gy + gz * gz;
            sc = vhs_sigc(k, g2);
            sd = scatter_dir(k, step, m, w.w, g2);
        }
        const float pm = wave_prefix_max(sc);
        const float sml = fmaxf(smax, pm);
        const bool acc = act && (u01(w.z) < sc / sml);
        bool conf = false;
        if (LMARK) {
            if (acc) { atomicMin(&mark[i1], lane); atomicMin(&mark[i2], lane); }
            __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
            __builtin_amdgcn_wave_barrier();
            if (act) conf = mark[i1] < lane || mark[i2] < lane;
            __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
            __builtin_amdgcn_wave_barrier();
            if (acc) { mark[i1] = 64; mark[i2] = 64; }
        } else {
            const int a1 = acc ? i1 : -3, a2 = acc ? i2 : -4;
            for (int j = 0; j < nact; ++j) {
                const int a = __builtin_amdgcn_readlane(a1, j), b = __builtin_amdgcn_readlane(a2, j);
                if (j < lane && (a == i1 || a == i2 || b == i1 || b == i2)) conf = true;
            }
        }
        const unsigned long long cm = __ballot(act && conf);
        const int f = cm ? (int)__ffsll((long long)cm) - 1 : nact;
        const bool accd = acc && lane < f;
        if (accd) {
            if constexpr (ROT) {
                float e1 = Er[i1], e2 = Er[i2];
                lb_exchange(k, step, m, g2, e1, e2, sd);
                Er[i1] = e1; Er[i2] = e2;
            }
            scatter_pair(sd, v1x, v1y, v1z, v2x, v2y, v2z);
            Vx[i1] = v1x; Vy[i1] = v1y; Vz[i1] = v1z;
            Vx[i2] = v2x; Vy[i2] = v2y; Vz[i2] = v2z;
        }
        nc += __popcll(__ballot(accd));
        smax = fmaxf(smax, __int_as_float(__builtin_amdgcn_readlane(__float_as_int(pm), f - 1)));
        mb += f;
        if (GLOBAL) __threadfence_block();
        else __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        __builtin_amdgcn_wave_barrier();
    }
    __builtin_amdgcn_s_setprio(0);
    return nc;
}

struct CellArgs {
    DevCase cs;              // batched runs: ncell = all cells, vol_ntc = all cells
    SoA A, B;
    const int *start;        // [0, C]: particle starts; [C+1, 2C+1]: run starts
    int *perm;               // a single cell above the LDS capacity: its permutation
    const int4 *runlist;     // per run, in cell order: id (64 chunk + k), destination base, lane mask
    const int *tile_first;   // int4 (first cell, its start, its run start, 0) per tile
    float *sigcmax, *mrem;
    int *ncoll_cell;
    double *acc;
    Ctrl *ctrl;
    int sample_start;
    // batched runs only
    const CellConst *cc;     // ninst
    const int *cbase;        // ninst + 1: first global cell of each instance
    int ninst;
    InstCtrl *ictl;
    int tile_p;              // particles per tile (the scatter's tile boundaries)
};

// Largest i with tab[i] <= c (tab[0] = 0, nondecreasing).
__device__ __forceinline__ int find_cell_inst(const int *tab, int ninst, int c)
{
    int lo = 0, hi = ninst - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (tab[mid] <= c) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// Per-cell constants of one case: uniform across the launch.
struct KSingle {
    const DevCase &cs;
    bool sampling;
    __device__ void prepare(CellShared &, int, int) const {}
    __device__ CK key(int, int c) const { return CK{ cs.K1, cs.e_exp, cs.seed, c, cs.rot_p, cs.lb_ia, cs.lb_ib }; }
    __device__ float ntc_coef(int) const { return cs.ntc_coef; }
    __device__ bool sample(int) const { return sampling; }
    __device__ void add(int, uint32_t, uint32_t) const {}
    // cells with adaptive sub-cells collide and sample in k_amr_cell
    __device__ bool amr(int c, int k) const { return amr_cell(cs, c, k); }
};

// Batched instances: an LDS copy of the instance table (first cell of every
// instance), the constants of the instances that the current chunk of cells
// spans (at most CC_WIN of them in LDS, the rest read from the global table)
// and their sums of collisions and candidates, added to the instance
// counters when the chunk is done.  Staging only those keeps the block's LDS
// small enough for four blocks per CU.
constexpr int CC_WIN = 8;

struct BatchShared {
    int cbase[MAX_INST + 1];
    CellConst cc[CC_WIN];            // instances w0 .. w0 + wn - 1
    uint32_t icoll[CC_WIN], icand[CC_WIN];
    int w0, wn;
    int last;
};

struct KBatch {
    BatchShared &b;
    const int *inst;     // instance of each staged cell (CellShared::inst)
    int ninst;
    uint32_t step;
    const CellConst *gcc;    // every instance's constants (global)
    InstCtrl *ictl;
    // Adds the window's sums to the instance counters.  Called by wave 0
    // only, after a block barrier that ended every add.  Returns the OR of
    // the atomics' old values: a caller that uses it has waited for every
    // add of the wave to be performed.
    __device__ uint32_t flush() const
    {
        const int t = (int)threadIdx.x;
        uint32_t dep = 0u;
        if (t < b.wn) {
            if (b.icoll[t]) dep |= atomicAdd(&ictl[b.w0 + t].step_coll, b.icoll[t]);
            if (b.icand[t]) dep |= atomicAdd(&ictl[b.w0 + t].step_cand, b.icand[t]);
        }
        return dep;
    }
    // The previous chunk ended in a block barrier.  Wave 0 flushes the old
    // window and then stages the new one: one wave runs in order, so its
    // reads of the old window precede its writes of the new; the caller's
    // barrier publishes it.
    __device__ void prepare(CellShared &sh, int c0, int nc) const
    {
        for (int i = threadIdx.x; i < nc; i += blockDim.x) sh.inst[i] = find_cell_inst(b.cbase, ninst, c0 + i);
        if (threadIdx.x < 64) {
            flush();
            const int w0 = find_cell_inst(b.cbase, ninst, c0);
            const int wn = min(find_cell_inst(b.cbase, ninst, c0 + nc - 1) - w0 + 1, CC_WIN);
            const int t = (int)threadIdx.x;
            if (t < wn) {
                const CellConst q = gcc[w0 + t];
                b.cc[t] = q;
                b.icoll[t] = 0u;
                b.icand[t] = 0u;
            }
            if (t == 0) { b.w0 = w0; b.wn = wn; }
        }
    }
    __device__ CellConst cc(int u) const
    {
        const int k = u - b.w0;
        return (k >= 0 && k < b.wn) ? b.cc[k] : gcc[u];
    }
    __device__ CK key(int i, int c) const
    {
        const int u = inst[i];
        const CellConst q = cc(u);
        return CK{ q.K1, q.e_exp, q.seed, c - b.cbase[u], 0.f, 0.f, 0.f };   // monatomic
    }
    __device__ float ntc_coef(int i) const { return cc(inst[i]).ntc_coef; }
    __device__ bool sample(int i) const
    {
        const int s0 = cc(inst[i]).sample_start;
        return s0 >= 0 && step >= (uint32_t)s0;
    }
    __device__ bool amr(int, int) const { return false; }   // batched instances: no sub-cells
    __device__ void add(int i, uint32_t coll, uint32_t cand) const
    {
        const int u = inst[i], k = u - b.w0;
        if (k >= 0 && k < b.wn) {
            if (coll) atomicAdd(&b.icoll[k], coll);
            if (cand) atomicAdd(&b.icand[k], cand);
        } else {
            if (coll) atomicAdd(&ictl[u].step_coll, coll);
            if (cand) atomicAdd(&ictl[u].step_cand, cand);
        }
    }
};

// One tile: cells [ca, cb), particles [p0, p0+n), runs [ra, rb).  LDS
// selects where the permutation, velocities and (ROT: diatomic gas)
// rotational energies are staged (block LDS, or HBM in place).
template <bool LDS, bool ROT, class KF>
__device__ void cell_tile(const CellArgs &a, const KF &kf, const SoA &cur, const SoA &nxt, uint32_t step,
                          int ca, int cb, int p0, int n, int ra, int rb, int *Sp, float *Vx, float *Vy, float *Vz,
                          float *Er, CellShared &sh, uint32_t &tcoll, uint32_t &tcand, int *lmarks, int lmark_cap,
                          bool claim, int carry, int *ahead)
{
    const int tid = threadIdx.x, bs = blockDim.x, wave = tid >> 6, lane = lane_id();
    const int nwaves = bs >> 6;
    const DevCase &cs = a.cs;
    // 0. the permutation of the tile, from its runs: run (id, base, lane
    // mask) holds the particles 64 (id / 64) + (the mask's set bits), in
    // ascending order from position base.  Each run is expanded by P threads,
    // one 64/P-bit piece of its mask each (P = 4, 2 or 1, as many as one pass
    // of the block's threads allows), which bounds the serial writes per
    // thread.  A tile of one chunk of cells loads its runs, cell starts and
    // collision state in one round trip (the metadata stays staged until the
    // collisions); larger tiles stage each chunk's metadata at the
    // collisions.
    const bool one = cb - ca <= CELL_MAXC;
    // the block's next tile claimed ahead (its queue position goes to
    // *ahead; carry = one claimed before and not yet used): the atomic is
    // issued before this phase's loads, so their wait covers its return
    int qa = carry;
    if (claim && tid == 0) qa = (int)atomicAdd(&a.ctrl->tile_next, 1u) + (int)gridDim.x;
    {
        const int nc = cb - ca, nr = rb - ra;
        const int lp = nr <= CELL_THREADS / 4 ? 2 : nr <= CELL_THREADS / 2 ? 1 : 0;   // log2 P
        const int pbits = 64 >> lp;
        for (int i0 = 0; i0 < (nr << lp) || i0 == 0; i0 += CELL_THREADS) {
            const int item = i0 + tid;
            const bool has = item < (nr << lp);
            int4 e = make_int4(0, 0, 0, 0);
            if (has) e = a.runlist[ra + (item >> lp)];
            int st0 = 0;
            float sm0 = 0.0f, mr0 = 0.0f, vol0 = 0.0f;
            if (one && i0 == 0) {
                if (tid <= nc) st0 = a.start[ca + tid];
                if (tid < nc) {
                    sm0 = a.sigcmax[ca + tid];
                    mr0 = a.mrem[ca + tid];
                    vol0 = cs.vol_ntc[ca + tid];
                }
                kf.prepare(sh, ca, nc);
            }
            if (has) {
                const int q0 = e.x & ~63;
                const unsigned long long m = ((unsigned long long)(uint32_t)e.w << 32) | (uint32_t)e.z;
                const int lo = (item & ((1 << lp) - 1)) * pbits;          // this thread's piece: bits [lo, lo + pbits)
                const unsigned long long below = (1ull << lo) - 1ull;     // lo < 64
                unsigned long long mine = pbits == 64 ? m : m & (((1ull << pbits) - 1ull) << lo);
                const int b = e.y - p0 + __popcll(m & below);
                // a run always lies inside its tile; the check keeps a
                // damaged run table from writing outside the staging
                if (b < 0 || b + __popcll(mine) > n) mine = 0ull;
                int *const o = Sp + b;
                for (int r = 0; mine; ++r, mine &= mine - 1ull) o[r] = q0 + (int)__builtin_ctzll(mine);
            }
            if (one && i0 == 0) {
                if (tid <= nc) sh.st[tid] = st0 - p0;
                if (tid < nc) {
                    sh.sm[tid] = sm0;
                    sh.mr[tid] = mr0;
                    sh.vol[tid] = vol0;
                }
            }
        }
    }
    if (tid == 0) *ahead = qa;
    if (!LDS) __threadfence();
    __syncthreads();

    // 1. gather into cell order, GATHER_UNROLL particles per thread in flight
    {
        const float *__restrict__ cx = cur.x, *__restrict__ cy = cur.y;
        const float *__restrict__ cvx = cur.vx, *__restrict__ cvy = cur.vy, *__restrict__ cvz = cur.vz;
        const float *__restrict__ cer = cur.er;
        const uint32_t *__restrict__ cp = cur.pid;
        float *__restrict__ ox = nxt.x + p0, *__restrict__ oy = nxt.y + p0;
        uint32_t *__restrict__ op = nxt.pid + p0;
        for (int k0 = tid; k0 < n; k0 += GATHER_UNROLL * bs) {
            int q[GATHER_UNROLL];
            float gx[GATHER_UNROLL], gy[GATHER_UNROLL], gvx[GATHER_UNROLL], gvy[GATHER_UNROLL],
                gvz[GATHER_UNROLL], ge[GATHER_UNROLL];
            uint32_t gp[GATHER_UNROLL];
#pragma unroll
            for (int j = 0; j < GATHER_UNROLL; ++j) {
                const int k = k0 + j * bs;
                q[j] = k < n ? Sp[k] : -1;
            }
#pragma unroll
            for (int j = 0; j < GATHER_UNROLL; ++j)
                if (q[j] >= 0) {
                    gx[j] = cx[q[j]]; gy[j] = cy[q[j]]; gp[j] = cp[q[j]];
                    gvx[j] = cvx[q[j]]; gvy[j] = cvy[q[j]]; gvz[j] = cvz[q[j]];
                    if constexpr (ROT) ge[j] = cer[q[j]];
                }
#pragma unroll
            for (int j = 0; j < GATHER_UNROLL; ++j)
                if (q[j] >= 0) {
                    const int k = k0 + j * bs;
                    ox[k] = gx[j]; oy[k] = gy[j]; op[k] = gp[j];
                    Vx[k] = gvx[j]; Vy[k] = gvy[j]; Vz[k] = gvz[j];
                    if constexpr (ROT) Er[k] = ge[j];
                }
        }
    }
    if (!LDS) __threadfence();
    __syncthreads();

    // 2-3. collisions and sampling, chunk by chunk of cells
    for (int c0 = ca; c0 < cb; c0 += CELL_MAXC) {
        const int nc = min(cb - c0, CELL_MAXC);
        if (!one) {
            for (int i = tid; i <= nc; i += bs) sh.st[i] = a.start[c0 + i] - p0;
            for (int i = tid; i < nc; i += bs) {
                sh.sm[i] = a.sigcmax[c0 + i];
                sh.mr[i] = a.mrem[c0 + i];
                sh.vol[i] = cs.vol_ntc[c0 + i];
            }
            kf.prepare(sh, c0, nc);
        }
        if (tid == 0) { sh.nheavy = 0; sh.nbig = 0; }
        __syncthreads();
        // candidate counts; cells go to the grouped rounds or, with many
        // candidates, to whole-wave rounds
        for (int i = tid; i < nc; i += bs) {
            const int o = sh.st[i], k = sh.st[i + 1] - o;
            int M = 0;
            const float vol = sh.vol[i];
            if (k >= 2 && vol > 0.0f && !kf.amr(c0 + i, k)) {
                float mr = sh.mr[i];
                M = ntc_count(kf.ntc_coef(i), k, vol, sh.sm[i], mr);
                sh.mr[i] = mr;
                if (M > NTC_GROUP_MAX) sh.heavy[atomicAdd(&sh.nheavy, 1)] = i;
                else if (M > 0) sh.big[atomicAdd(&sh.nbig, 1)] = i;
            }
            sh.nc[i] = 0;
            sh.M[i] = M;
            tcand += (uint32_t)M;
            kf.add(i, 0u, (uint32_t)M);
        }
        __syncthreads();
        ntc_groups<NTC_G, !LDS, ROT>(kf, step, c0, sh.big, sh.nbig, sh, Vx, Vy, Vz, Er, wave, nwaves);
        if (!LDS) __threadfence();
        __syncthreads();
        const int nh = sh.nheavy;
        for (int h = wave; h < nh; h += nwaves) {
            const int i = sh.heavy[h];
            const int o = sh.st[i], k = sh.st[i + 1] - o;
            float smax = sh.sm[i];
            // conflict marks: the cell's permutation segment, free after the
            // gather (LDS tiles); for a single cell staged in HBM, the block's
            // idle LDS staging area when the cell fits, else shuffles
            int ncoll;
            float *eo = ROT ? Er + o : nullptr;
            if (LDS)
                ncoll = ntc_rounds<false, ROT>(kf.key(i, c0 + i), step, k, sh.M[i], Vx + o, Vy + o, Vz + o, eo, smax,
                                               Sp + o);
            else if (lmarks && k <= lmark_cap)
                ncoll = ntc_rounds<true, ROT, true>(kf.key(i, c0 + i), step, k, sh.M[i], Vx + o, Vy + o, Vz + o, eo,
                                                    smax, lmarks);
            else
                ncoll = ntc_rounds<true, ROT, false>(kf.key(i, c0 + i), step, k, sh.M[i], Vx + o, Vy + o, Vz + o,
                                                     eo, smax, Sp + o);
            if (lane == 0) { sh.sm[i] = smax; sh.nc[i] = ncoll; }
        }
        if (!LDS) __threadfence();
        __syncthreads();
        for (int i = tid; i < nc; i += bs) {
            const int c = c0 + i;
            const int o = sh.st[i], k = sh.st[i + 1] - o;
            a.sigcmax[c] = sh.sm[i];
            a.mrem[c] = sh.mr[i];
            a.ncoll_cell[c] = sh.nc[i];
            tcoll += (uint32_t)sh.nc[i];
            kf.add(i, (uint32_t)sh.nc[i], 0u);
        }
        // FP64 sampling sums: SAMPLE_G lanes per cell, each over every
        // SAMPLE_G-th particle, then a fixed butterfly -- an order that
        // depends on the cell's particle count only (not on tiles or
        // launch settings), so every run of a case sums alike
        {
            const int g = lane / SAMPLE_G, sub = lane % SAMPLE_G;
            constexpr int CPW = 64 / SAMPLE_G;
            const size_t C = (size_t)cs.ncell;
            for (int base = wave * CPW; base < nc; base += nwaves * CPW) {
                const int i = base + g;
                const int o = i < nc ? sh.st[i] : 0;
                const int k = i < nc ? sh.st[i + 1] - o : 0;
                const bool on = k > 0 && kf.sample(i) && !kf.amr(c0 + i, k);
                double sx = 0.0, sy = 0.0, sz = 0.0, s2 = 0.0, se = 0.0;
                // unrolled so that several particles' loads are in flight
                // before the adds, which keep their order
                if (on)
#pragma unroll 4
                    for (int q = sub; q < k; q += SAMPLE_G) {
                        const double u = (double)Vx[o + q], v = (double)Vy[o + q], w = (double)Vz[o + q];
                        sx += u; sy += v; sz += w;
                        s2 += u * u + v * v + w * w;
                        if constexpr (ROT) se += (double)Er[o + q];
                    }
                // the butterfly over 4, 2, 1 by DPP (each lane adds its
                // partner's value, as a shuffle would give it)
                static_assert(SAMPLE_G == 8, "the butterfly spans 8 lanes");
                auto fold = [&](auto dc) {
                    constexpr int D = decltype(dc)::value;
                    sx += xor_lane_d<D>(sx);
                    sy += xor_lane_d<D>(sy);
                    sz += xor_lane_d<D>(sz);
                    s2 += xor_lane_d<D>(s2);
                    if constexpr (ROT) se += xor_lane_d<D>(se);
                };
                fold(std::integral_constant<int, 4>{});
                fold(std::integral_constant<int, 2>{});
                fold(std::integral_constant<int, 1>{});
                // one lane per cell adds into the accumulators: the FP64
                // adds are done at the memory side (the same IEEE sum, one
                // writer per cell per step), so nothing waits for a load
                if (on && sub == 0) {
                    const int c = c0 + i;
                    unsafeAtomicAdd(&a.acc[c], (double)k);
                    unsafeAtomicAdd(&a.acc[C + c], sx);
                    unsafeAtomicAdd(&a.acc[2 * C + c], sy);
                    unsafeAtomicAdd(&a.acc[3 * C + c], sz);
                    unsafeAtomicAdd(&a.acc[4 * C + c], s2);
                    unsafeAtomicAdd(&a.acc[5 * C + c], (double)sh.nc[i]);
                    if constexpr (ROT) unsafeAtomicAdd(&a.acc[6 * C + c], se);
                }
            }
        }
        __syncthreads();
    }

    // 4. velocities back to HBM
    if (LDS)
        for (int k = tid; k < n; k += bs) {
            nxt.vx[p0 + k] = Vx[k];
            nxt.vy[p0 + k] = Vy[k];
            nxt.vz[p0 + k] = Vz[k];
            if constexpr (ROT) nxt.er[p0 + k] = Er[k];
        }
    __syncthreads();
}

struct NoBatch {};

template <bool BATCH, bool ROT>
__global__ __launch_bounds__(256, ROT ? 3 : 4) void k_cell(CellArgs a)
{
    // LDS staging of a tile: permutation, velocities and (ROT) rotational
    // energies; for a single cell staged in HBM it holds the collision
    // conflict marks instead
    constexpr int NSTAGE = ROT ? 5 : 4;
    __shared__ union { int i[NSTAGE * CELL_CAP]; float f[NSTAGE * CELL_CAP]; } stage;
    int *const s_perm = stage.i;
    float *const s_vx = stage.f + CELL_CAP, *const s_vy = stage.f + 2 * CELL_CAP, *const s_vz = stage.f + 3 * CELL_CAP;
    float *const s_er = ROT ? stage.f + 4 * CELL_CAP : nullptr;
    __shared__ CellShared sh;
    __shared__ typename std::conditional<BATCH, BatchShared, NoBatch>::type bsh;
    Ctrl *ctrl = a.ctrl;
    const int tid = threadIdx.x;
    const uint32_t step = ctrl->step;
    const SoA cur = (step & 1u) ? a.B : a.A;
    const SoA nxt = (step & 1u) ? a.A : a.B;
    const int n_next = (int)ctrl->n_next;
    const int ntiles = (n_next + a.tile_p - 1) / a.tile_p;
    const bool sampling = a.sample_start >= 0 && step >= (uint32_t)a.sample_start;
    __shared__ int s_tile;
    __shared__ int s_ahead[2];    // a tile claimed ahead, by round parity
    if (tid == 0) { sh.coll = 0u; sh.cand = 0u; }
    if constexpr (BATCH) {
        for (int t = tid; t <= a.ninst; t += blockDim.x) bsh.cbase[t] = a.cbase[t];
        if (tid == 0) { bsh.w0 = 0; bsh.wn = 0; }
    }
    uint32_t tcoll = 0, tcand = 0;
    const int4 *tiles = reinterpret_cast<const int4 *>(a.tile_first);
    int rest_c = 0, rest_cb = 0;      // the part of a cut tile still to do
    int pending = -1;                 // a queue position claimed ahead and not yet used
    for (int round = 0;; ++round) {
        int ca, cb, p0, n, ra, rb;
        int qcur = -1;
        const int *const rstart = a.start + a.cs.ncell + 1;
        if (rest_c < rest_cb) {
            ca = rest_c; cb = rest_cb;
            p0 = a.start[ca]; n = a.start[cb] - p0;
            ra = rstart[ca]; rb = rstart[cb];
        } else {
            // dynamic tile queue: dense tiles cost more, blocks take the
            // next free one (each block's first tile is its own index, no
            // atomic), from the last cell backwards -- with the inflow at
            // x = 0 the cheap, uniform freestream tiles are then the ones
            // left to fill the end of the launch
            int q = (int)blockIdx.x;
            if (round > 0) {
                if (pending >= 0) {
                    q = pending;
                    pending = -1;
                } else {
                    if (tid == 0) s_tile = (int)atomicAdd(&ctrl->tile_next, 1u) + (int)gridDim.x;
                    __syncthreads();
                    q = s_tile;
                    __syncthreads();
                }
            }
            if (q >= ntiles) break;
            qcur = q;
            const int t = ntiles - 1 - q;
            // tile t: cells [ca, cb), particles [p0, p0 + n), runs [ra, rb),
            // from the table of (first cell, its start, its run start)
            const int4 tf = tiles[t];
            const int4 tn = (t + 1 < ntiles) ? tiles[t + 1] : make_int4(a.cs.ncell, n_next, (int)ctrl->n_runs, 0);
            ca = tf.x; cb = tn.x;
            p0 = tf.y; n = tn.y - p0;
            ra = tf.z; rb = tn.z;
        }
        rest_c = rest_cb = 0;
        if (ca >= cb) continue;
        if (n > CELL_CAP) {
            // beyond the LDS capacity: this round takes the longest run of
            // whole cells that fits (at least one cell; a single cell above
            // CELL_CAP stages in HBM), the rest comes next round
            const int m = min(cb - ca, CELL_MAXC);
            for (int i = tid; i <= m; i += blockDim.x) sh.st[i] = a.start[ca + i];
            __syncthreads();
            const int fit = __syncthreads_count(tid >= 1 && tid <= m && sh.st[tid] - p0 <= CELL_CAP);
            const int c2 = ca + (fit > 0 ? fit : 1);
            n = sh.st[c2 - ca] - p0;
            rb = rstart[c2];
            __syncthreads();
            rest_c = c2; rest_cb = cb;
            cb = c2;
        }
        // the next tile is claimed during this one while the queue is far
        // from its end (a block holding a claimed tile there would leave the
        // others idle at the end of the launch)
        const bool claim = pending < 0 && qcur >= 0 && qcur + 2 * (int)gridDim.x < ntiles;
        int *const ahead = &s_ahead[round & 1];
        auto run = [&](const auto &kf) {
            if (n <= CELL_CAP)
                cell_tile<true, ROT>(a, kf, cur, nxt, step, ca, cb, p0, n, ra, rb, s_perm, s_vx, s_vy, s_vz, s_er, sh,
                                     tcoll, tcand, nullptr, 0, claim, pending, ahead);
            else    // a single cell: its collision rounds alone use the idle staging LDS for marks
                cell_tile<false, ROT>(a, kf, cur, nxt, step, ca, cb, p0, n, ra, rb, a.perm + p0, nxt.vx + p0,
                                      nxt.vy + p0, nxt.vz + p0, ROT ? nxt.er + p0 : nullptr, sh, tcoll, tcand, stage.i,
                                      4 * CELL_CAP, claim, pending, ahead);
        };
        if constexpr (BATCH) run(KBatch{ bsh, sh.inst, a.ninst, step, a.cc, a.ictl });
        else run(KSingle{ a.cs, sampling });
        pending = *ahead;
    }
    if (tcoll) atomicAdd(&sh.coll, tcoll);
    if (tcand) atomicAdd(&sh.cand, tcand);
    // End of the step.  Each block adds its sums into its group's line and
    // counts itself done there; the last block of a group moves the group's
    // sums into the step counters and counts the group done in cell_done.
    // Every add is a returning atomic whose old value feeds the next
    // increment (through a zero the compiler cannot fold): an atomic's value
    // returns once it is performed device-wide, so each counter is complete
    // when the increment after it is seen, with no device-scope release
    // (which writes back the XCD's L2, once per exiting block).  The groups
    // keep the ~1,000 exiting blocks off one line: there their atomics
    // queued for tens of microseconds and held up the tiles still running.
    uint32_t dep = 0u;
    if constexpr (BATCH) {
        __syncthreads();
        if (tid < 64) dep = KBatch{ bsh, sh.inst, a.ninst, step, a.cc, a.ictl }.flush();
    }
    __syncthreads();
    bool last = false;
    if (tid == 0) {
        unsigned long long t;
        if (exit_group_add(ctrl, sh.coll, sh.cand, dep, t)) {
            uint32_t dep2 = (uint32_t)t | (uint32_t)(t >> 32);
            if ((uint32_t)t) dep2 |= atomicAdd(&ctrl->step_coll, (uint32_t)t);
            if ((uint32_t)(t >> 32)) dep2 |= atomicAdd(&ctrl->step_cand, (uint32_t)(t >> 32));
            last = atomicAdd(&ctrl->cell_done, 1u + opaque_zero(dep2)) == min(gridDim.x, (uint32_t)EXIT_GROUPS) - 1u;
        }
        if constexpr (BATCH) bsh.last = last ? 1 : 0;
    }
    if constexpr (BATCH) {
        // the last block closes every instance's step
        __syncthreads();
        if (!bsh.last) return;
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        for (int t = tid; t < a.ninst; t += blockDim.x) {
            InstCtrl &ic = a.ictl[t];
            const uint32_t nl = ic.n_live;
            const uint32_t nn = (uint32_t)(a.start[bsh.cbase[t + 1]] - a.start[bsh.cbase[t]]);
            const uint32_t sc = __hip_atomic_load(&ic.step_coll, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            const uint32_t sm = __hip_atomic_load(&ic.step_cand, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            const int s0 = a.cc[t].sample_start;
            ic.psteps += nl;
            ic.next_pid += ic.n_move - nl;
            ic.n_live = nn;
            ic.last_coll = sc;
            ic.ncoll += sc;
            ic.last_cand = sm;
            ic.ncand += sm;
            ic.step_coll = 0u;
            ic.step_cand = 0u;
            if (s0 >= 0 && step >= (uint32_t)s0) ic.nsamples += 1ull;
        }
        if (tid != 0) return;
    } else {
        if (!last) return;
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    }
    const uint32_t n_live = ctrl->n_live;
    const uint32_t sc = __hip_atomic_load(&ctrl->step_coll, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const uint32_t sm = __hip_atomic_load(&ctrl->step_cand, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    ctrl->psteps += n_live;
    ctrl->next_pid += ctrl->n_move - n_live;
    ctrl->n_live = (uint32_t)n_next;
    ctrl->last_coll = sc;
    ctrl->ncoll += sc;
    ctrl->last_cand = sm;
    ctrl->ncand += sm;
    ctrl->step_coll = 0u;
    ctrl->step_cand = 0u;
    if (sampling) ctrl->nsamples += 1ull;
    ctrl->step = step + 1u;
    ctrl->tile_next = 0u;
    ctrl->cell_done = 0u;
}

// ---------------------------------------------------------------------------
// Adaptive collision sub-cells (Cell::genAMRmesh, Cell.cpp:168-279;
// Cell::sortParticle2children, Cell.cpp:280-295; Cell::collision over the
// leaves, Cell.cpp:335-352).  After k_cell (which leaves these cells alone
// but for the gather into cell order), one block per cell with sub-cells:
//   1. leaf of every particle; a stable order by leaf (wave 0 walks the
//      cell in chunks of 64, one LDS cursor per leaf);
//   2. the candidate numbers leaf by leaf with the remainder carried in leaf
//      order (thread 0), then one thread per leaf runs its candidates with
//      its own running (sigma c)max from the cell's -- the candidates are
//      numbered on through the leaves, one Philox stream per cell and step;
//   3. the cell's state written back in leaf order, FP64 sampling sums.
// The arithmetic is k_cell's (ntc_count, vhs_sigc, scatter_dir,
// scatter_pair, lb_exchange).
// ---------------------------------------------------------------------------
struct AmrArgs {
    DevCase cs;
    SoA A, B;
    const int *start;
    float *sigcmax, *mrem;
    int *ncoll_cell;
    double *acc;
    Ctrl *ctrl;
    int sample_start;
};

constexpr int AMR_THREADS = 256;
constexpr int AMR_WCAP = 256;        // a cell of at most this many particles is taken by one wave

// LDS of one wave's cell, and of the block's cell: leaf order, leaf of each
// particle, velocities (and rotational energies) in leaf order, and per
// leaf its start, cursor, candidate count and first candidate number.
template <int CAP, int NV>
struct AmrLds {
    int idx[CAP];
    int leaf[CAP];
    float v[NV][CAP];
    int ls[AMR_LEAVES + 1], pos[AMR_LEAVES], M[AMR_LEAVES], m0[AMR_LEAVES];
};
template <int NV>
union AmrShared {
    AmrLds<AMR_WCAP, NV> w[AMR_THREADS / 64];
    AmrLds<AMR_CAP, NV> b;
};

__device__ __forceinline__ void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

// One cell with sub-cells, np <= AMR_WCAP particles, by one wave: the steps
// of the block form below, with the leaves dealt to the wave's lanes and
// wave barriers only.  The arithmetic and every order (leaf order, candidate
// numbering, the lane-strided FP64 sampling sums and their butterfly) are
// the block form's, so both give the same bits.
template <bool ROT>
__device__ void amr_cell_wave(const AmrArgs &a, const SoA &S, uint32_t step, bool sampling, int c,
                              AmrLds<AMR_WCAP, ROT ? 4 : 3> &L, uint32_t &wcoll, uint32_t &wcand)
{
    const DevCase &cs = a.cs;
    const int lane = lane_id();
    const unsigned long long lt = (1ull << lane) - 1ull;
    const int o = a.start[c], np = a.start[c + 1] - o;
    int nsx, nsy;
    const int nl = amr_leaves(cs.sub[c], nsx, nsy);
    for (int l = lane; l < nl; l += 64) L.pos[l] = 0;
    wave_sync();
    for (int k = lane; k < np; k += 64) {
        const int lf = amr_leaf(cs, c, S.x[o + k], S.y[o + k], nsx, nsy);
        L.leaf[k] = lf;
        atomicAdd(&L.pos[lf], 1);
    }
    wave_sync();
    // leaf starts, candidate numbers with the remainder carried in leaf order
    if (lane == 0) {
        const float vol = cs.vol_ntc[c];
        const float vleaf = vol / (float)nl;
        const float smax0 = a.sigcmax[c];
        float mr = a.mrem[c];
        int st = 0, m0 = 0;
        for (int l = 0; l < nl; ++l) {
            const int npl = L.pos[l];
            L.ls[l] = st;
            L.pos[l] = st;
            st += npl;
            int M = 0;
            if (vol > 0.0f && npl >= 2) M = ntc_count(cs.ntc_coef, npl, vleaf, smax0, mr);
            L.M[l] = M;
            L.m0[l] = m0;
            m0 += M;
        }
        L.ls[nl] = st;
        a.mrem[c] = mr;
        wcand += (uint32_t)m0;
    }
    wave_sync();
    // stable order by leaf, chunk after chunk in index order
    for (int k0 = 0; k0 < np; k0 += 64) {
        const int k = k0 + lane;
        const int lf = k < np ? L.leaf[k] : -1;
        unsigned long long todo = __ballot(lf >= 0);
        while (todo) {
            const int l0 = (int)__ffsll((long long)todo) - 1;
            const int ll = __shfl(lf, l0, 64);
            const unsigned long long m = __ballot(lf == ll);
            int base = 0;
            if (lane == l0) base = L.pos[ll];
            base = __shfl(base, l0, 64);
            if (lf == ll) L.idx[base + __popcll(m & lt)] = k;
            if (lane == l0) L.pos[ll] = base + __popcll(m);
            wave_sync();
            todo &= ~m;
        }
    }
    for (int p = lane; p < np; p += 64) {
        const int k = o + L.idx[p];
        L.v[0][p] = S.vx[k];
        L.v[1][p] = S.vy[k];
        L.v[2][p] = S.vz[k];
        if constexpr (ROT) L.v[3][p] = S.er[k];
    }
    wave_sync();
    // one lane per leaf
    const CK key{ cs.K1, cs.e_exp, cs.seed, c, cs.rot_p, cs.lb_ia, cs.lb_ib };
    const float smax0 = a.sigcmax[c];
    float best = smax0;
    uint32_t nc = 0;
    for (int l = lane; l < nl; l += 64) {
        const int M = L.M[l];
        if (M == 0) continue;
        const int b = L.ls[l], npl = L.ls[l + 1] - b, m0 = L.m0[l];
        float smax = smax0;
        for (int m = m0; m < m0 + M; ++m) {
            const uint4 w = philox((uint32_t)m, step, SUB_COLLIDE, 0u, (uint32_t)c, cs.seed);
            const int i1 = mulhi_idx(w.x, npl);
            int i2 = mulhi_idx(w.y, npl - 1);
            if (i2 >= i1) i2 = i2 + 1;
            const int p1 = b + i1, p2 = b + i2;
            float v1x = L.v[0][p1], v1y = L.v[1][p1], v1z = L.v[2][p1];
            float v2x = L.v[0][p2], v2y = L.v[1][p2], v2z = L.v[2][p2];
            const float gx = v1x - v2x, gy = v1y - v2y, gz = v1z - v2z;
            const float g2 = gx * gx + gy * gy + gz * gz;
            const float sc = vhs_sigc(key, g2);
            if (sc > smax) smax = sc;
            if (u01(w.z) < sc / smax) {
                Scatter sd = scatter_dir(key, step, m, w.w, g2);
                if constexpr (ROT) lb_exchange(key, step, m, g2, L.v[3][p1], L.v[3][p2], sd);
                scatter_pair(sd, v1x, v1y, v1z, v2x, v2y, v2z);
                L.v[0][p1] = v1x; L.v[1][p1] = v1y; L.v[2][p1] = v1z;
                L.v[0][p2] = v2x; L.v[1][p2] = v2y; L.v[2][p2] = v2z;
                ++nc;
            }
        }
        best = fmaxf(best, smax);
    }
#pragma unroll
    for (int d = 32; d > 0; d >>= 1) {
        nc += __shfl_xor(nc, d, 64);
        best = fmaxf(best, __shfl_xor(best, d, 64));
    }
    wave_sync();
    // the cell in leaf order: positions and ids through registers
    constexpr int PR = AMR_WCAP / 64;
    float rx[PR], ry[PR];
    uint32_t rp[PR];
#pragma unroll
    for (int j = 0; j < PR; ++j) {
        const int p = lane + j * 64;
        if (p < np) {
            const int k = o + L.idx[p];
            rx[j] = S.x[k]; ry[j] = S.y[k]; rp[j] = S.pid[k];
        }
    }
    wave_sync();
#pragma unroll
    for (int j = 0; j < PR; ++j) {
        const int p = lane + j * 64;
        if (p < np) {
            S.x[o + p] = rx[j]; S.y[o + p] = ry[j]; S.pid[o + p] = rp[j];
            S.vx[o + p] = L.v[0][p]; S.vy[o + p] = L.v[1][p]; S.vz[o + p] = L.v[2][p];
            if constexpr (ROT) S.er[o + p] = L.v[3][p];
        }
    }
    if (lane == 0) {
        a.sigcmax[c] = best;
        a.ncoll_cell[c] = (int)nc;
        wcoll += nc;
    }
    if (sampling) {
        double sx = 0.0, sy = 0.0, sz = 0.0, s2 = 0.0, se = 0.0;
        for (int p = lane; p < np; p += 64) {
            const double u = (double)L.v[0][p], v = (double)L.v[1][p], w = (double)L.v[2][p];
            sx += u; sy += v; sz += w;
            s2 += u * u + v * v + w * w;
            if constexpr (ROT) se += (double)L.v[3][p];
        }
#pragma unroll
        for (int d = 32; d > 0; d >>= 1) {
            sx += __shfl_xor(sx, d, 64);
            sy += __shfl_xor(sy, d, 64);
            sz += __shfl_xor(sz, d, 64);
            s2 += __shfl_xor(s2, d, 64);
            if constexpr (ROT) se += __shfl_xor(se, d, 64);
        }
        if (lane == 0) {
            const size_t C = (size_t)cs.ncell;
            unsafeAtomicAdd(&a.acc[c], (double)np);
            unsafeAtomicAdd(&a.acc[C + c], sx);
            unsafeAtomicAdd(&a.acc[2 * C + c], sy);
            unsafeAtomicAdd(&a.acc[3 * C + c], sz);
            unsafeAtomicAdd(&a.acc[4 * C + c], s2);
            unsafeAtomicAdd(&a.acc[5 * C + c], (double)nc);
            if constexpr (ROT) unsafeAtomicAdd(&a.acc[6 * C + c], se);
        }
    }
    wave_sync();
}

template <bool ROT>
__global__ __launch_bounds__(AMR_THREADS) void k_amr_cell(AmrArgs a)
{
    __shared__ AmrShared<ROT ? 4 : 3> lds;
    int *const s_idx = lds.b.idx;           // leaf order -> particle in the cell
    int *const s_leaf = lds.b.leaf;         // leaf of each particle (cell order)
    float (*const s_v)[AMR_CAP] = lds.b.v;
    int *const s_ls = lds.b.ls, *const s_pos = lds.b.pos, *const s_M = lds.b.M, *const s_m0 = lds.b.m0;
    __shared__ int s_list[AMR_THREADS], s_nlist;      // cells for the block
    __shared__ int s_wlist[AMR_THREADS], s_nwlist;    // cells for a wave each
    __shared__ uint32_t s_nc, s_best;
    const DevCase &cs = a.cs;
    const int tid = threadIdx.x, lane = tid & 63;
    const uint32_t step = a.ctrl->step - 1u;          // k_cell has advanced the step
    const SoA S = ((step + 1u) & 1u) ? a.B : a.A;     // the state k_cell wrote
    const bool sampling = a.sample_start >= 0 && step >= (uint32_t)a.sample_start;
    const size_t C = (size_t)cs.ncell;
    uint32_t bcoll = 0, bcand = 0;
    // block b takes the cells b, b + G, b + 2G, ... (G blocks): cells with
    // sub-cells cluster in space, so a cyclic deal spreads them over every
    // block; each pass checks AMR_THREADS of the block's cells and lists
    // those with sub-cells
    const int G = (int)gridDim.x;
    for (int j0 = 0; (long long)j0 * G + (int)blockIdx.x < cs.ncell; j0 += AMR_THREADS) {
        if (tid == 0) { s_nlist = 0; s_nwlist = 0; }
        __syncthreads();
        {
            const long long cl = (long long)(j0 + tid) * G + (int)blockIdx.x;
            const int c = cl < cs.ncell ? (int)cl : -1;
            const int k = c >= 0 ? a.start[c + 1] - a.start[c] : 0;
            if (c >= 0 && amr_cell(cs, c, k)) {
                if (k <= AMR_WCAP) s_wlist[atomicAdd(&s_nwlist, 1)] = c;
                else s_list[atomicAdd(&s_nlist, 1)] = c;
            }
        }
        __syncthreads();
        // the small cells, a wave each
        {
            const int wave = tid >> 6, nwl = s_nwlist;
            for (int q = wave; q < nwl; q += AMR_THREADS / 64)
                amr_cell_wave<ROT>(a, S, step, sampling, s_wlist[q], lds.w[wave], bcoll, bcand);
        }
        __syncthreads();
        const int nlist = s_nlist;
        for (int q = 0; q < nlist; ++q) {
            const int c = s_list[q];
            const int o = a.start[c], np = a.start[c + 1] - o;
            int nsx, nsy;
            const int nl = amr_leaves(cs.sub[c], nsx, nsy);
            // 1. leaves and their counts
            for (int l = tid; l < nl; l += AMR_THREADS) s_pos[l] = 0;
            if (tid == 0) { s_nc = 0u; s_best = __float_as_uint(a.sigcmax[c]); }
            __syncthreads();
            for (int k = tid; k < np; k += AMR_THREADS) {
                const int lf = amr_leaf(cs, c, S.x[o + k], S.y[o + k], nsx, nsy);
                s_leaf[k] = lf;
                atomicAdd(&s_pos[lf], 1);
            }
            __syncthreads();
            // 2a. leaf starts, candidate numbers with the carried remainder
            if (tid == 0) {
                const float vol = cs.vol_ntc[c];
                const float vleaf = vol / (float)nl;
                const float smax0 = a.sigcmax[c];
                float mr = a.mrem[c];
                int st = 0, m0 = 0;
                for (int l = 0; l < nl; ++l) {
                    const int npl = s_pos[l];
                    s_ls[l] = st;
                    s_pos[l] = st;
                    st += npl;
                    int M = 0;
                    if (vol > 0.0f && npl >= 2) M = ntc_count(cs.ntc_coef, npl, vleaf, smax0, mr);
                    s_M[l] = M;
                    s_m0[l] = m0;
                    m0 += M;
                }
                s_ls[nl] = st;
                a.mrem[c] = mr;
                bcand += (uint32_t)m0;
            }
            __syncthreads();
            // 1b. stable order by leaf: wave 0, chunk after chunk in index order
            if (tid < 64) {
                const unsigned long long lt = (1ull << lane) - 1ull;
                for (int k0 = 0; k0 < np; k0 += 64) {
                    const int k = k0 + lane;
                    const int lf = k < np ? s_leaf[k] : -1;
                    unsigned long long todo = __ballot(lf >= 0);
                    while (todo) {
                        const int l0 = (int)__ffsll((long long)todo) - 1;
                        const int ll = __shfl(lf, l0, 64);
                        const unsigned long long m = __ballot(lf == ll);
                        int base = 0;
                        if (lane == l0) base = s_pos[ll];
                        base = __shfl(base, l0, 64);
                        if (lf == ll) s_idx[base + __popcll(m & lt)] = k;
                        if (lane == l0) s_pos[ll] = base + __popcll(m);
                        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
                        __builtin_amdgcn_wave_barrier();
                        todo &= ~m;
                    }
                }
            }
            __syncthreads();
            // velocities in leaf order
            for (int p = tid; p < np; p += AMR_THREADS) {
                const int k = o + s_idx[p];
                s_v[0][p] = S.vx[k];
                s_v[1][p] = S.vy[k];
                s_v[2][p] = S.vz[k];
                if constexpr (ROT) s_v[3][p] = S.er[k];
            }
            __syncthreads();
            // 2b. one thread per leaf
            const CK key{ cs.K1, cs.e_exp, cs.seed, c, cs.rot_p, cs.lb_ia, cs.lb_ib };
            for (int l = tid; l < nl; l += AMR_THREADS) {
                const int M = s_M[l];
                if (M == 0) continue;
                const int b = s_ls[l], npl = s_ls[l + 1] - b, m0 = s_m0[l];
                float smax = a.sigcmax[c];              // the cell's value at the step's start
                uint32_t nc = 0;
                for (int m = m0; m < m0 + M; ++m) {
                    const uint4 w = philox((uint32_t)m, step, SUB_COLLIDE, 0u, (uint32_t)c, cs.seed);
                    const int i1 = mulhi_idx(w.x, npl);
                    int i2 = mulhi_idx(w.y, npl - 1);
                    if (i2 >= i1) i2 = i2 + 1;
                    const int p1 = b + i1, p2 = b + i2;
                    float v1x = s_v[0][p1], v1y = s_v[1][p1], v1z = s_v[2][p1];
                    float v2x = s_v[0][p2], v2y = s_v[1][p2], v2z = s_v[2][p2];
                    const float gx = v1x - v2x, gy = v1y - v2y, gz = v1z - v2z;
                    const float g2 = gx * gx + gy * gy + gz * gz;
                    const float sc = vhs_sigc(key, g2);
                    if (sc > smax) smax = sc;
                    if (u01(w.z) < sc / smax) {
                        Scatter sd = scatter_dir(key, step, m, w.w, g2);
                        if constexpr (ROT) lb_exchange(key, step, m, g2, s_v[3][p1], s_v[3][p2], sd);
                        scatter_pair(sd, v1x, v1y, v1z, v2x, v2y, v2z);
                        s_v[0][p1] = v1x; s_v[1][p1] = v1y; s_v[2][p1] = v1z;
                        s_v[0][p2] = v2x; s_v[1][p2] = v2y; s_v[2][p2] = v2z;
                        ++nc;
                    }
                }
                atomicAdd(&s_nc, nc);
                atomicMax(&s_best, __float_as_uint(smax));   // non-negative floats order as their bits
            }
            __syncthreads();
            // 3. the cell in leaf order: positions and ids through registers
            constexpr int PR = AMR_CAP / AMR_THREADS;
            float rx[PR], ry[PR];
            uint32_t rp[PR];
#pragma unroll
            for (int j = 0; j < PR; ++j) {
                const int p = tid + j * AMR_THREADS;
                if (p < np) {
                    const int k = o + s_idx[p];
                    rx[j] = S.x[k]; ry[j] = S.y[k]; rp[j] = S.pid[k];
                }
            }
            __syncthreads();
#pragma unroll
            for (int j = 0; j < PR; ++j) {
                const int p = tid + j * AMR_THREADS;
                if (p < np) {
                    S.x[o + p] = rx[j]; S.y[o + p] = ry[j]; S.pid[o + p] = rp[j];
                    S.vx[o + p] = s_v[0][p]; S.vy[o + p] = s_v[1][p]; S.vz[o + p] = s_v[2][p];
                    if constexpr (ROT) S.er[o + p] = s_v[3][p];
                }
            }
            const uint32_t ncell_coll = s_nc;
            if (tid == 0) {
                a.sigcmax[c] = __uint_as_float(s_best);
                a.ncoll_cell[c] = (int)ncell_coll;
                bcoll += ncell_coll;
            }
            // FP64 sampling sums of the cell (wave 0)
            if (sampling && tid < 64) {
                double sx = 0.0, sy = 0.0, sz = 0.0, s2 = 0.0, se = 0.0;
                for (int p = lane; p < np; p += 64) {
                    const double u = (double)s_v[0][p], v = (double)s_v[1][p], w = (double)s_v[2][p];
                    sx += u; sy += v; sz += w;
                    s2 += u * u + v * v + w * w;
                    if constexpr (ROT) se += (double)s_v[3][p];
                }
#pragma unroll
                for (int d = 32; d > 0; d >>= 1) {
                    sx += __shfl_xor(sx, d, 64);
                    sy += __shfl_xor(sy, d, 64);
                    sz += __shfl_xor(sz, d, 64);
                    s2 += __shfl_xor(s2, d, 64);
                    if constexpr (ROT) se += __shfl_xor(se, d, 64);
                }
                if (lane == 0) {
                    unsafeAtomicAdd(&a.acc[c], (double)np);
                    unsafeAtomicAdd(&a.acc[C + c], sx);
                    unsafeAtomicAdd(&a.acc[2 * C + c], sy);
                    unsafeAtomicAdd(&a.acc[3 * C + c], sz);
                    unsafeAtomicAdd(&a.acc[4 * C + c], s2);
                    unsafeAtomicAdd(&a.acc[5 * C + c], (double)ncell_coll);
                    if constexpr (ROT) unsafeAtomicAdd(&a.acc[6 * C + c], se);
                }
            }
            __syncthreads();
        }
    }
    // the block's sums (lane 0 of each wave; the block's cells: thread 0)
    // go through the exit groups that k_cell left cleared: the last block
    // of a group adds the group's sums to the totals, so the Ctrl line takes
    // a few atomics per group instead of four from every wave
    __shared__ uint32_t s_wc[AMR_THREADS / 64], s_wd[AMR_THREADS / 64];
    if (lane == 0) { s_wc[tid >> 6] = bcoll; s_wd[tid >> 6] = bcand; }
    __syncthreads();
    if (tid == 0) {
        uint32_t coll = 0, cand = 0;
        for (int w = 0; w < AMR_THREADS / 64; ++w) { coll += s_wc[w]; cand += s_wd[w]; }
        unsigned long long t;
        if (exit_group_add(a.ctrl, coll, cand, 0u, t)) {
            const uint32_t gc = (uint32_t)t, gd = (uint32_t)(t >> 32);
            if (gc) { atomicAdd(&a.ctrl->ncoll, (unsigned long long)gc); atomicAdd(&a.ctrl->last_coll, gc); }
            if (gd) { atomicAdd(&a.ctrl->ncand, (unsigned long long)gd); atomicAdd(&a.ctrl->last_cand, gd); }
        }
    }
}

// Sub-cell grids, mean free paths and VTS time steps from the time-averaged
// sums (Cell::VTS, Cell.cpp:117-136; Cell::genAMRmesh, Cell.cpp:168-243):
// lam = m / (sqrt(2) rho pi d^2) (T/T_inf)^(omega - 1/2) (the power by the
// FP32 log/exp above, the rest in FP64), dt = lam / (2 (|u| + sqrt(2kT/m))),
// level 2 of side lam/2, n = max(1, ceil(L/side)) at most AMR_MAX; level 3
// from the smallest old level-3 mean free path -- one per cell here, so
// lam -- or from twice the level-2 x side when the cell had no grid.
struct AmrGas {
    double Fn, mass, diam, T_inf, omega, boltz;
};

__device__ __forceinline__ int amr_div(double L, double side)
{
    const double q = L / side;
    if (!(q > 1.0)) return 1;
    const double n = ceil(q);
    return n > (double)AMR_MAX ? AMR_MAX : (int)n;
}

__device__ __forceinline__ uint32_t amr_grid(double dx, double dy, double lam_max, double lam_min, bool had)
{
    const int n2x = amr_div(dx, 0.5 * lam_max), n2y = amr_div(dy, 0.5 * lam_max);
    const double l2x = dx / (double)n2x, l2y = dy / (double)n2y;
    const double mn = had ? lam_min : 2.0 * l2x;
    const int n3x = amr_div(l2x, 0.5 * mn), n3y = amr_div(l2y, 0.5 * mn);
    return (uint32_t)n2x | ((uint32_t)n2y << 8) | ((uint32_t)n3x << 16) | ((uint32_t)n3y << 24);
}

__global__ __launch_bounds__(256) void k_amr_update(DevCase cs, AmrGas g, const double *__restrict__ acc,
                                                    const double *__restrict__ vol, const Ctrl *__restrict__ ctrl,
                                                    uint32_t *__restrict__ sub, float *__restrict__ lam,
                                                    double *__restrict__ dt)
{
    const int C = cs.ncell;
    const long long ns = (long long)ctrl->nsamples;
    const double pi_d2 = sqrt(2.0) * 3.14159265358979323846 * g.diam * g.diam;
    for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < C; c += gridDim.x * blockDim.x) {
        const double S0 = acc[c];
        lam[c] = 0.0f;
        dt[c] = 0.0;
        if (!(S0 > 0.0) || !(vol[c] > 0.0) || ns <= 0) { sub[c] = 0u; continue; }
        const double n = (S0 / (double)ns) * g.Fn / vol[c];
        const double ux = acc[C + c] / S0, uy = acc[2 * C + c] / S0, uz = acc[3 * C + c] / S0;
        const double uu = ux * ux + uy * uy + uz * uz;
        const double T = g.mass / (3.0 * g.boltz) * (acc[4 * C + c] / S0 - uu);
        const float tr = (float)(T / g.T_inf);
        if (!(tr > 0.0f)) { sub[c] = 0u; continue; }
        const float f = dm_expf((float)(g.omega - 0.5) * dm_logf(tr));
        const double l = g.mass / (pi_d2 * (n * g.mass)) * (double)f;
        lam[c] = (float)l;
        dt[c] = 0.5 * l / (sqrt(uu) + sqrt(2.0 * g.boltz * T / g.mass));
        sub[c] = amr_grid((double)cs.dx, (double)cs.dy, l, l, sub[c] != 0u);
    }
}

}  // namespace

hipError_t launch_cell(const DevCase &cs, const SoA &A, const SoA &B, const int *start, int *perm,
                       const int4 *runlist, const int *tile_first, float *sigcmax, float *mrem, int *ncoll_cell,
                       double *acc, Ctrl *ctrl, int sample_start, int tile_p, int grid, hipStream_t st)
{
    if (tile_p < CELL_P_MIN || tile_p > CELL_CAP) return hipErrorInvalidValue;
    CellArgs a{ cs, A, B, start, perm, runlist, tile_first, sigcmax, mrem, ncoll_cell, acc, ctrl, sample_start,
                nullptr, nullptr, 0, nullptr, tile_p };
    if (cs.zeta != 0) {
        if (!A.er || !B.er) return hipErrorInvalidValue;
        hipLaunchKernelGGL((k_cell<false, true>), dim3(grid), dim3(CELL_THREADS), 0, st, a);
    } else {
        hipLaunchKernelGGL((k_cell<false, false>), dim3(grid), dim3(CELL_THREADS), 0, st, a);
    }
    return hipGetLastError();
}

hipError_t launch_cell_batch(const DevCase &all, const CellConst *cc, const int *cbase, int ninst,
                             InstCtrl *ictl, const SoA &A, const SoA &B, const int *start, int *perm,
                             const int4 *runlist, const int *tile_first, float *sigcmax, float *mrem, int *ncoll_cell, double *acc,
                             Ctrl *ctrl, int tile_p, int grid, hipStream_t st)
{
    if (ninst < 1 || ninst > MAX_INST) return hipErrorInvalidValue;
    if (tile_p < CELL_P_MIN || tile_p > CELL_CAP) return hipErrorInvalidValue;
    CellArgs a{ all, A, B, start, perm, runlist, tile_first, sigcmax, mrem, ncoll_cell, acc, ctrl, -1, cc, cbase, ninst,
                ictl, tile_p };
    hipLaunchKernelGGL((k_cell<true, false>), dim3(grid), dim3(CELL_THREADS), 0, st, a);
    return hipGetLastError();
}

int occupancy_cell(bool rot)
{
    int nb = 0;
    const void *f = rot ? (const void *)k_cell<false, true> : (const void *)k_cell<false, false>;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, f, 256, 0) != hipSuccess || nb < 1) nb = 1;
    return nb;
}

int occupancy_cell_batch()
{
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, (const void *)k_cell<true, false>, 256, 0) != hipSuccess ||
        nb < 1)
        nb = 1;
    return nb;
}

hipError_t launch_amr_cell(const DevCase &cs, const SoA &A, const SoA &B, const int *start, float *sigcmax,
                           float *mrem, int *ncoll_cell, double *acc, Ctrl *ctrl, int sample_start, int grid,
                           hipStream_t st)
{
    if (!cs.sub) return hipErrorInvalidValue;
    AmrArgs a{ cs, A, B, start, sigcmax, mrem, ncoll_cell, acc, ctrl, sample_start };
    if (cs.zeta != 0) {
        if (!A.er || !B.er) return hipErrorInvalidValue;
        hipLaunchKernelGGL((k_amr_cell<true>), dim3(grid), dim3(AMR_THREADS), 0, st, a);
    } else {
        hipLaunchKernelGGL((k_amr_cell<false>), dim3(grid), dim3(AMR_THREADS), 0, st, a);
    }
    return hipGetLastError();
}

hipError_t launch_amr_update(const DevCase &cs, double Fn, double mass, double diam, double T_inf, double omega,
                             double boltz, const double *acc, const double *vol, const Ctrl *ctrl, uint32_t *sub,
                             float *lam, double *dt, hipStream_t st)
{
    const AmrGas g{ Fn, mass, diam, T_inf, omega, boltz };
    const int grid = std::max(1, std::min((cs.ncell + 255) / 256, 4096));
    hipLaunchKernelGGL(k_amr_update, dim3(grid), dim3(256), 0, st, cs, g, acc, vol, ctrl, sub, lam, dt);
    return hipGetLastError();
}

int occupancy_amr_cell(bool rot)
{
    int n = 0;
    if (rot) (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, (const void *)k_amr_cell<true>, AMR_THREADS, 0);
    else (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, (const void *)k_amr_cell<false>, AMR_THREADS, 0);
    return n > 0 ? n : 1;
}

}  // namespace dsmc
