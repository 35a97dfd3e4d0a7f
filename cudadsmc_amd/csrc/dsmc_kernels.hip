// HIP kernels of the DSMC hot path for gfx950 (wave64, 160 KB LDS per CU).
//
// One timestep = five launches, no host round trip:
//   k_move       free flight + walls + body + outflow, inflow creation, and
//                the binning: per wave (64 consecutive particles) one run per
//                cell, one packed atomic per run  (cudaDSMC GPUParticle.cu:229-363,
//                                                   :375-422, GPUCell.cu:169-198)
//   k_scan_runs  exclusive scan of the per-cell particle and run counts
//                within tiles of cells             (GPUCell.cu:104-120)
//   k_place      the tile offsets (every block, in LDS), then every cell's
//                runs listed contiguously
//   k_rank       per cell, runs ranked by wave: each run's destination base
//                (the stable order: ascending particle index within a cell),
//                the start tables and the k_cell tile boundaries
//                                                  (GPUParticle.cu:365-373)
//   k_cell       per tile of cells: the tile's runs expanded into the
//                cell-sorted permutation, gather into cell order (this is also the
//                outflow compaction), NTC/VHS collisions, FP64 sampling
//                                                  (GPUCell.cu:200-402,
//                                                   GPUParticle.cu:76-140)
// All launches have fixed grids (persistent, grid-stride over the
// device-resident particle count), so a step can be captured in a hipGraph.
#include <algorithm>

#include "dsmc_kernels.h"

namespace dsmc {

// ---------------------------------------------------------------------------
// Wave64 helpers
// ---------------------------------------------------------------------------
__device__ __forceinline__ int lane_id() { return (int)__lane_id(); }

// Ascending bitonic sort of one int per lane across the 64 lanes (flip
// formulation: every comparator is ascending, so +inf padding is free).
__device__ __forceinline__ int wave_sort64(int v)
{
    const int lane = lane_id();
#pragma unroll
    for (int size = 2; size <= 64; size <<= 1) {
        int partner = lane ^ (size - 1);
        int o = __shfl(v, partner, 64);
        v = lane < partner ? min(v, o) : max(v, o);
#pragma unroll
        for (int s = size >> 2; s > 0; s >>= 1) {
            partner = lane ^ s;
            o = __shfl(v, partner, 64);
            v = lane < partner ? min(v, o) : max(v, o);
        }
    }
    return v;
}

// Binning of one wave, which holds the 64 consecutive particles of chunk
// i >> 6 (lane = i & 63): the lanes are grouped by cell, each group being one
// run of its cell (the state is cell-sorted, so a wave touches only a few
// cells).  The lowest lane of a group adds (1 run, group size) to the cell's
// packed counter -- one atomic per run, all of the wave's going out together
// -- and records (cell, the run slot the atomic returned, lane mask) as the
// chunk's k-th run, k counting the groups in lane order; lane 0 stores the
// chunk's run count.  Lanes with c < 0 (left the domain) are in no run.
// agg_issue groups the lanes and sends the atomics; agg_finish writes the
// run records, which need the atomics' values.  A loop over chunks can
// finish a chunk after issuing the next one, so the atomic's round trip
// overlaps the next chunk's move.
struct RunPend {
    unsigned long long mine, old;   // the lane's run (leaders only), the counter it got
    int c, rec;                     // its cell, its record index (64 chunk + k)
};
__device__ __forceinline__ RunPend agg_issue(const Bins &bn, int c, uint32_t i)
{
    const int lane = lane_id();
    const unsigned long long lt = (1ull << lane) - 1ull;
    unsigned long long todo = __ballot(c >= 0);
    unsigned long long mine = 0ull;
    while (todo) {
        const int l = (int)__ffsll((long long)todo) - 1;
        const int lc = lane_value(c, l);
        const unsigned long long m = __ballot(c == lc);
        if (lane == l) mine = m;
        todo &= ~m;
    }
    const unsigned long long leaders = __ballot(mine != 0ull);
    const uint32_t q = i >> 6;
    RunPend r{ mine, 0ull, c, (int)(q * 64u) + (int)__popcll(leaders & lt) };
    if (mine) r.old = atomicAdd(&bn.cnt[c], (1ull << 32) | (unsigned long long)__popcll(mine));
    if (lane == 0) bn.chunk_runs[q] = __popcll(leaders);
    return r;
}
__device__ __forceinline__ void agg_finish(const Bins &bn, const RunPend &r)
{
    if (r.mine)
        bn.runrec[(size_t)r.rec] =
            make_int4(r.c, (int)(uint32_t)(r.old >> 32), (int)(uint32_t)r.mine, (int)(uint32_t)(r.mine >> 32));
}
__device__ __forceinline__ void agg_run(const Bins &bn, int c, uint32_t i) { agg_finish(bn, agg_issue(bn, c, i)); }

// The face of the q-th inflow particle of a step (the faces' particles
// follow one another, J[f] of face f): the number of faces whose cumulative
// count is <= q.
__device__ __forceinline__ int inflow_face(const int J[4], uint32_t q)
{
    const uint32_t a0 = (uint32_t)J[0], a1 = a0 + (uint32_t)J[1], a2 = a1 + (uint32_t)J[2];
    return (q >= a0 ? 1 : 0) + (q >= a1 ? 1 : 0) + (q >= a2 ? 1 : 0);
}

// ---------------------------------------------------------------------------
// k_move
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_move(DevCase cs, SoA A, SoA B, Bins bn, Ctrl *__restrict__ ctrl)
{
    const uint32_t step = ctrl->step;
    const SoA S = (step & 1u) ? B : A;
    const uint32_t n_live = ctrl->n_live;
    const uint32_t next_pid = ctrl->next_pid;
    int J[4];
    inject_counts(cs, step, J);
    uint32_t Jt = (uint32_t)(J[0] + J[1] + J[2] + J[3]);
    bool clamped = false;
    if (n_live + Jt > cs.cap) { Jt = cs.cap - n_live; clamped = true; }
    const uint32_t ntot = n_live + Jt;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        ctrl->n_move = ntot;
        if (clamped) atomicOr(&ctrl->err, (uint32_t)ERR_CAPACITY);
    }
    uint32_t err = 0;
    const uint32_t stride = gridDim.x * blockDim.x;
    const bool rot = cs.zeta != 0;
    // the live state of a thread's next particle is loaded one iteration
    // ahead, so those loads are in flight while the current particle moves
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    float px = 0.0f, py = 0.0f, pvx = 0.0f, pvy = 0.0f, pvz = 0.0f;
    if (i < n_live) { px = S.x[i]; py = S.y[i]; pvx = S.vx[i]; pvy = S.vy[i]; pvz = S.vz[i]; }
    RunPend pend{ 0ull, 0ull, 0, 0 };
    for (; i < ntot; i += stride) {
        float x = px, y = py, vx = pvx, vy = pvy, vz = pvz, er = 0.0f;
        const uint32_t j = i + stride;
        if (j < n_live) { px = S.x[j]; py = S.y[j]; pvx = S.vx[j]; pvy = S.vy[j]; pvz = S.vz[j]; }
        bool dead;
        if (i < n_live) {
            bool vchg = false, echg = false;
            dead = move_one(cs, x, y, vx, vy, vz, er, S.pid + i, 0u, false, step, cs.dt, err, vchg, echg);
            S.x[i] = x; S.y[i] = y;
            if (vchg) { S.vx[i] = vx; S.vy[i] = vy; S.vz[i] = vz; }
            if (echg) S.er[i] = er;     // re-drawn at a diffuse wall (zeta = 2 only)
        } else {
            const uint32_t q = i - n_live;
            const uint32_t pid = next_pid + q;
            dead = inject_one(cs, inflow_face(J, q), pid, step, x, y, vx, vy, vz, er, err);
            S.x[i] = x; S.y[i] = y; S.vx[i] = vx; S.vy[i] = vy; S.vz[i] = vz; S.pid[i] = pid;
            if (rot) S.er[i] = er;
        }
        const RunPend r = agg_issue(bn, dead ? -1 : cell_of(cs, x, y), i);
        agg_finish(bn, pend);
        pend = r;
    }
    agg_finish(bn, pend);
    if (err) atomicOr(&ctrl->err, err);
}

// ---------------------------------------------------------------------------
// k_scan: tile-local exclusive scan (SCAN_TILE cells per block), tile sums,
// and the last block to finish scans the tile sums.  count is zeroed for the
// next step.  start(c) = lstart[c] + toff[c / SCAN_TILE].
// ---------------------------------------------------------------------------
__device__ __forceinline__ int wave_incl_scan(int v) { return wave_prefix_sum(v); }

// Exclusive scan over the 256 threads of a block; returns the block total.
__device__ __forceinline__ int block_excl_scan(int v, int &excl, int *s_w)
{
    const int lane = lane_id(), w = threadIdx.x >> 6;
    int inc = wave_incl_scan(v);
    if (lane == 63) s_w[w] = inc;
    __syncthreads();
    int wofs = 0, total = 0;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        int t = s_w[k];
        if (k < w) wofs += t;
        total += t;
    }
    excl = wofs + inc - v;
    __syncthreads();
    return total;
}

__global__ __launch_bounds__(256) void k_scan(int *__restrict__ count, int *__restrict__ lstart,
                                              int *__restrict__ tsum, int *__restrict__ toff,
                                              int ncell, int ntile, Ctrl *__restrict__ ctrl)
{
    __shared__ int s_w[4];
    __shared__ int s_last;
    const int base = blockIdx.x * SCAN_TILE + threadIdx.x * SCAN_IPT;
    int v[SCAN_IPT];
    int run = 0;
#pragma unroll
    for (int k = 0; k < SCAN_IPT; ++k) {
        int c = base + k;
        v[k] = c < ncell ? count[c] : 0;
        run += v[k];
    }
#pragma unroll
    for (int k = 0; k < SCAN_IPT; ++k)
        if (base + k < ncell) count[base + k] = 0;
    int excl;
    int total = block_excl_scan(run, excl, s_w);
#pragma unroll
    for (int k = 0; k < SCAN_IPT; ++k) {
        int c = base + k;
        if (c < ncell) lstart[c] = excl;
        excl += v[k];
    }
    if (threadIdx.x == 0) tsum[blockIdx.x] = total;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    __syncthreads();
    if (threadIdx.x == 0) s_last = (atomicAdd(&ctrl->scan_done, 1u) == (uint32_t)(ntile - 1));
    __syncthreads();
    if (!s_last) return;
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    int carry = 0;
    for (int b0 = 0; b0 < ntile; b0 += 256) {
        int b = b0 + (int)threadIdx.x;
        int t = b < ntile ? __hip_atomic_load(&tsum[b], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : 0;
        int e;
        int tot = block_excl_scan(t, e, s_w);
        if (b < ntile) toff[b] = carry + e;
        carry += tot;
    }
    if (threadIdx.x == 0) {
        ctrl->n_next = (uint32_t)carry;
        ctrl->scan_done = 0u;
    }
}

// ---------------------------------------------------------------------------
// k_scan_runs: a tile-local scan of the packed per-cell counters (runs << 32
// | particles) and the tile totals.  Neither half can carry into the other --
// both totals stay below 2^32 -- so one 64-bit scan gives the particle starts
// in the low half and the run starts in the high half.  cnt is zeroed for the
// next step.  The tile offsets are scanned by every block of k_place in LDS
// (block 0 publishes them), so no block here waits on another: a last-block
// finish needs a device-scope release per block, which writes back the L2
// of its XCD.
// ---------------------------------------------------------------------------
__device__ __forceinline__ unsigned long long block_excl_scan64(unsigned long long v, unsigned long long &excl,
                                                                unsigned long long *s_w)
{
    const int lane = lane_id(), w = threadIdx.x >> 6;
    const unsigned long long inc = wave_prefix_sum_u64(v);
    if (lane == 63) s_w[w] = inc;
    __syncthreads();
    unsigned long long wofs = 0, total = 0;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const unsigned long long t = s_w[k];
        if (k < w) wofs += t;
        total += t;
    }
    excl = wofs + inc - v;
    __syncthreads();
    return total;
}

__global__ __launch_bounds__(256) void k_scan_runs(Bins bn, int ncell)
{
    __shared__ unsigned long long s_w[4];
    const int base = blockIdx.x * SCAN_TILE + threadIdx.x * SCAN_IPT;
    unsigned long long v[SCAN_IPT];
    unsigned long long run = 0;
#pragma unroll
    for (int k = 0; k < SCAN_IPT; ++k) {
        const int c = base + k;
        v[k] = c < ncell ? bn.cnt[c] : 0ull;
        run += v[k];
    }
#pragma unroll
    for (int k = 0; k < SCAN_IPT; ++k)
        if (base + k < ncell) bn.cnt[base + k] = 0ull;
    unsigned long long excl;
    const unsigned long long total = block_excl_scan64(run, excl, s_w);
#pragma unroll
    for (int k = 0; k < SCAN_IPT; ++k) {
        const int c = base + k;
        if (c < ncell) bn.lstart[c] = excl;
        excl += v[k];
    }
    if (threadIdx.x == 0) bn.tsum[blockIdx.x] = total;
}

// The tile offsets: exclusive scan of the tile totals into s_toff (n <=
// TOFF_LDS tiles), by one block; the totals end the step's particle and run
// counts.  With `publish` the block also writes them for the later kernels
// (bn.toff, ctrl n_next, n_runs, nruns).
constexpr int TOFF_LDS = 2048;

__device__ void tile_offsets(const Bins &bn, int ntile, unsigned long long *s_toff, unsigned long long *s_w,
                             bool publish, Ctrl *__restrict__ ctrl)
{
    unsigned long long carry = 0;
    for (int b0 = 0; b0 < ntile; b0 += 256) {
        const int b = b0 + (int)threadIdx.x;
        const unsigned long long t = b < ntile ? bn.tsum[b] : 0ull;
        unsigned long long e;
        const unsigned long long tot = block_excl_scan64(t, e, s_w);
        if (b < ntile) {
            s_toff[b] = carry + e;
            if (publish) bn.toff[b] = carry + e;
        }
        carry += tot;
    }
    if (publish && threadIdx.x == 0) {
        ctrl->n_next = (uint32_t)carry;
        ctrl->n_runs = (uint32_t)(carry >> 32);
        ctrl->nruns += carry >> 32;
    }
    __syncthreads();
}

// k_scan_tiles: the same in global memory, for grids of more than TOFF_LDS
// tiles (one block).
__global__ __launch_bounds__(256) void k_scan_tiles(Bins bn, int ntile, Ctrl *__restrict__ ctrl)
{
    __shared__ unsigned long long s_w[4];
    unsigned long long carry = 0;
    for (int b0 = 0; b0 < ntile; b0 += 256) {
        const int b = b0 + (int)threadIdx.x;
        const unsigned long long t = b < ntile ? bn.tsum[b] : 0ull;
        unsigned long long e;
        const unsigned long long tot = block_excl_scan64(t, e, s_w);
        if (b < ntile) bn.toff[b] = carry + e;
        carry += tot;
    }
    if (threadIdx.x == 0) {
        ctrl->n_next = (uint32_t)carry;
        ctrl->n_runs = (uint32_t)(carry >> 32);
        ctrl->nruns += carry >> 32;
    }
}

// Packed starts of cell c (particles low, runs high); c == ncell gives the
// totals.
__device__ __forceinline__ unsigned long long packed_start(const Bins &bn, int c, int ncell, uint32_t n_next,
                                                           uint32_t n_runs)
{
    return c < ncell ? bn.lstart[c] + bn.toff[c / SCAN_TILE]
                     : (((unsigned long long)n_runs << 32) | (unsigned long long)n_next);
}

// XCD-aware split of `ntask` wave tasks: workgroups are dispatched to the
// eight XCDs in turn (block b on XCD b % 8), each XCD with its own L2, so XCD
// x takes the x-th eighth of the tasks and its waves stride through that
// range -- neighbouring tasks (adjacent cells, and the runs that land in
// them) then share one L2.  Sets [first, end) and the stride; a wave
// starts at first and steps by stride while below end.
constexpr int NXCD = 8;

__device__ __forceinline__ void xcd_range(int ntask, int &first, int &end, int &stride)
{
    const int nx = min(NXCD, (int)gridDim.x);                    // every part has a block
    const int x = (int)blockIdx.x % nx, lb = (int)blockIdx.x / nx;
    const int nb = ((int)gridDim.x - x + nx - 1) / nx;          // blocks on part x
    const int wpb = (int)blockDim.x >> 6;
    const int lo = (int)((long long)ntask * x / nx), hi = (int)((long long)ntask * (x + 1) / nx);
    first = lo + lb * wpb + ((int)threadIdx.x >> 6);
    end = hi;
    stride = nb * wpb;
}

// ---------------------------------------------------------------------------
// k_place: every cell's runs listed contiguously -- record k of chunk q goes
// to run start(cell) + its slot.  A wave takes up to 64 chunks, scans their
// run counts and deals the runs out to its lanes, so every lane places a run.
// Within a cell the list is in the atomics' arrival order, which k_rank
// ranks away.
// ---------------------------------------------------------------------------
constexpr int PLACE_U = 4;

__global__ __launch_bounds__(256) void k_place(Bins bn, int ncell, Ctrl *__restrict__ ctrl)
{
    __shared__ unsigned long long s_toff[TOFF_LDS];
    __shared__ unsigned long long s_w[4];
    const int ntile = (ncell + SCAN_TILE - 1) / SCAN_TILE;
    const bool lds = ntile <= TOFF_LDS;     // else k_scan_tiles has published them
    if (lds) tile_offsets(bn, ntile, s_toff, s_w, blockIdx.x == 0, ctrl);
    const unsigned long long *const toff = lds ? s_toff : bn.toff;
    const int lane = lane_id();
    const uint32_t nch = (ctrl->n_move + 63u) >> 6;
    // chunks per task: 64, halved (down to 4) while that leaves fewer than
    // four tasks per wave of the grid -- a small case then keeps every wave
    // busy instead of a few waves running long chains
    const uint32_t waves = gridDim.x * (blockDim.x >> 6);
    uint32_t cpw = 64u;
    while (cpw > 4u && nch / cpw < 4u * waves) cpw >>= 1;
    int first, end, stride;
    xcd_range((int)((nch + cpw - 1u) / cpw), first, end, stride);
    for (int task = first; task < end; task += stride) {
        const uint32_t q0 = (uint32_t)task * cpw;
        const int nr = (uint32_t)lane < cpw && q0 + lane < nch ? bn.chunk_runs[q0 + lane] : 0;
        const int inc = wave_incl_scan(nr), excl = inc - nr;
        const int total = lane_value(inc, 63);
        // PLACE_U runs per lane per pass, their record loads in flight together
        for (int t0 = 0; t0 < total; t0 += 64 * PLACE_U) {
            size_t ri[PLACE_U];
            int4 r[PLACE_U];
#pragma unroll
            for (int u = 0; u < PLACE_U; ++u) {
                const int t = t0 + 64 * u + lane;
                // the chunk holding run t: the last lane whose exclusive count is <= t
                int j = 0;
#pragma unroll
                for (int d = 32; d > 0; d >>= 1) {
                    const int ex = __shfl(excl, j + d, 64);
                    if (ex <= t) j += d;
                }
                const int exj = __shfl(excl, j, 64);   // every lane takes part in the shuffles
                ri[u] = t < total ? (size_t)(q0 + j) * 64 + (t - exj) : ~(size_t)0;
            }
#pragma unroll
            for (int u = 0; u < PLACE_U; ++u)
                if (ri[u] != ~(size_t)0) r[u] = bn.runrec[ri[u]];
#pragma unroll
            for (int u = 0; u < PLACE_U; ++u)
                if (ri[u] != ~(size_t)0) {
                    const unsigned long long v = bn.lstart[r[u].x] + toff[r[u].x / SCAN_TILE];
                    bn.runlist[(uint32_t)(v >> 32) + (uint32_t)r[u].y] = make_int4((int)ri[u], 0, r[u].z, r[u].w);
                }
        }
    }
}

// ---------------------------------------------------------------------------
// k_rank: a wave takes RK_C consecutive cells (one per lane).  A run's id is
// 64q + k for the k-th run of chunk q; the runs of one cell lie in distinct
// chunks, and chunk q holds particles [64q, 64q + 64), so ordering a cell's
// runs by id orders them by particle index.  Each run's destination base is
// start(c) + (the lengths of the cell's runs with smaller id); it goes into
// the unused field of the run's runlist entry, which then holds (id, base,
// lane mask): k_cell expands the runs of its tile from there into the
// cell-sorted permutation, every cell's particles in ascending index order --
// the order of the serial binning, bit for bit.  The task's cells' runs are
// contiguous in runlist: the wave stages them in LDS as (id, length) pairs
// with coalesced loads (RK_LDS at a time, in segments of whole cells), then
// each lane takes a run, finds its cell by a binary search over the cells'
// first runs, and ranks it against that cell's runs.  Also materialises
// start[0..C] (particles) and start[C+1 .. 2C+1] (runs), and, for every tile
// of tile_p particles, the first cell that starts in it with that cell's
// particle and run starts (tiles holds int4 (cell, start, run start, 0)).
// ---------------------------------------------------------------------------
constexpr int RK_C = 16;                 // cells per wave task
constexpr int RK_LDS = 1024;             // runs staged in LDS per wave
static_assert(RK_C < 64, "a lane beyond the task's cells holds the cell after them");
constexpr int RK_U = 8;                  // runs compared per step of the ranking loop

__global__ __launch_bounds__(256) void k_rank(Bins bn, int *__restrict__ start, int4 *__restrict__ tiles,
                                              int ncell, int tile_p, const Ctrl *__restrict__ ctrl)
{
    __shared__ int2 s_run[4][RK_LDS];
    int2 *const sr = s_run[threadIdx.x >> 6];
    const int lane = lane_id();
    const uint32_t n_next = ctrl->n_next, n_runs = ctrl->n_runs;
    const int ntiles = ((int)n_next + tile_p - 1) / tile_p;
    int first, end, stride;
    xcd_range(ncell / RK_C + 1, first, end, stride);   // cells 0..ncell (the last: the totals)
    // lane j < RK_C holds cell task RK_C + j; the lanes from RK_C on hold the
    // cell after the task's last, whose starts end the task's runs.  The
    // packed starts of a task's cells (and of the cell before the first, lane
    // 0, for the tile table) are loaded one task ahead: the next task's loads
    // are in flight while this one ranks its runs
    auto starts = [&](int task, unsigned long long &a0, unsigned long long &ap) {
        const int cc = task * RK_C + min(lane, RK_C);
        a0 = packed_start(bn, min(cc, ncell), ncell, n_next, n_runs);
        ap = (lane == 0 && cc > 0) ? packed_start(bn, cc - 1, ncell, n_next, n_runs) : 0ull;
    };
    unsigned long long n0 = 0ull, np0 = 0ull;
    if (first < end) starts(first, n0, np0);
    for (int task = first; task < end; task += stride) {
        const int c = task * RK_C + lane;
        const bool own = lane < RK_C;
        const unsigned long long v0 = n0;
        const unsigned long long v1 = __shfl_down(v0, 1, 64), up = __shfl_up(v0, 1, 64);
        const unsigned long long vp = lane == 0 ? np0 : up;
        const int st = (int)(uint32_t)v0, rs = (int)(v0 >> 32);
        const int L = own && c < ncell ? (int)(v1 >> 32) - rs : 0;
        if (own && c <= ncell) {
            start[c] = st;
            start[ncell + 1 + c] = rs;
            const int tlo = c > 0 ? (int)(uint32_t)vp / tile_p + 1 : 0;
            const int thi = min(st / tile_p, ntiles - 1);
            for (int t = tlo; t <= thi; ++t) tiles[t] = make_int4(c, st, rs, 0);
        }
        if (task + stride < end) starts(task + stride, n0, np0);
        const int R0 = lane_value(rs, 0);
        const int rel = rs - R0, rend = rel + L;      // the lane's cell's runs [rel, rend), from R0
        // segments of consecutive cells whose runs fit the staging area (a
        // cell of more runs than that alone, read from runlist itself)
        for (int s0 = 0; s0 < RK_C;) {
            const int base = lane_value(rel, s0);
            const unsigned long long fit = __ballot(lane >= s0 && rend - base <= RK_LDS);
            const bool lds = fit != 0ull;                 // else cell s0 alone has too many runs
            const int e = lds ? 64 - (int)__clzll((long long)fit) : s0 + 1;
            const int nr = lane_value(rend, e - 1) - base;
            int4 *const rl = bn.runlist + R0 + base;
            if (lds) {
                for (int k0 = 0; k0 < nr; k0 += 256) {
                    int4 ev[4];
#pragma unroll
                    for (int u = 0; u < 4; ++u) {
                        const int k = k0 + 64 * u + lane;
                        if (k < nr) ev[u] = rl[k];
                    }
#pragma unroll
                    for (int u = 0; u < 4; ++u) {
                        const int k = k0 + 64 * u + lane;
                        if (k < nr) sr[k] = make_int2(ev[u].x, __popc((uint32_t)ev[u].z) + __popc((uint32_t)ev[u].w));
                    }
                }
                __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
                __builtin_amdgcn_wave_barrier();
            }
            // run k of the segment: the lane whose first run is the last one
            // <= k holds its cell; each lane ranks one run per pass against
            // its cell's runs, RK_U at a time (their reads in flight
            // together).  The staged and the unstaged form are separate
            // loops, so the staged one has no global load whose wait would
            // also wait for the stores of the passes before.
            auto cell_of = [&](int k, int &a, int &Lj, int &sj) {
                int j = 0;
#pragma unroll
                for (int d = 32; d > 0; d >>= 1)
                    if (__shfl(rel, j + d, 64) <= base + k) j += d;
                a = __shfl(rel, j, 64) - base;
                Lj = __shfl(L, j, 64);
                sj = __shfl(st, j, 64);
            };
            if (lds) {
                for (int k0 = 0; k0 < nr; k0 += 64) {
                    const int k = k0 + lane;
                    int a, Lj, sj;
                    cell_of(k, a, Lj, sj);
                    if (k < nr) {
                        const int me = sr[k].x;
                        int off = 0;
                        for (int t = a; t < a + Lj; t += RK_U) {
                            int2 f[RK_U];
#pragma unroll
                            for (int u = 0; u < RK_U; ++u) f[u] = sr[min(t + u, RK_LDS - 1)];
#pragma unroll
                            for (int u = 0; u < RK_U; ++u)
                                if (t + u < a + Lj && f[u].x < me) off += f[u].y;
                        }
                        reinterpret_cast<int *>(rl + k)[1] = sj + off;
                    }
                }
            } else {
                for (int k0 = 0; k0 < nr; k0 += 64) {
                    const int k = k0 + lane;
                    int a, Lj, sj;
                    cell_of(k, a, Lj, sj);
                    if (k < nr) {
                        const int me = rl[k].x;
                        int off = 0;
                        for (int t = a; t < a + Lj; ++t) {
                            const int4 f = rl[t];
                            if (f.x < me) off += __popc((uint32_t)f.z) + __popc((uint32_t)f.w);
                        }
                        reinterpret_cast<int *>(rl + k)[1] = sj + off;
                    }
                }
            }
            if (lds) {   // the staging area is free again
                __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
                __builtin_amdgcn_wave_barrier();
            }
            s0 = e;
        }
    }
}

// ---------------------------------------------------------------------------
// Initial fill (Run::assignParticle, Run.cpp:108-157)
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_fill_count(const double *__restrict__ expected, int ncell,
                                                    uint32_t seed, int *__restrict__ count)
{
    for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < ncell; c += gridDim.x * blockDim.x) {
        uint4 w = philox(0xFFFFFFFFu, 0u, SUB_INIT, 0u, (uint32_t)c, seed);
        double U = (double)w.x * 0x1p-32;
        double e = expected[c];
        count[c] = e > 0.0 ? (int)floor(e + U) : 0;
    }
}

__device__ __forceinline__ bool inside_body(const DevCase &cs, float px, float py)
{
    bool in = false;
    for (int k = 0; k < cs.nseg; ++k) {
        const float *S = cs.seg + 6 * k;
        float ax = S[0], ay = S[1], bx = S[2], by = S[3];
        if ((ay > py) != (by > py)) {
            float xi = ax + ((py - ay) * (bx - ax)) / (by - ay);
            if (px < xi) in = !in;
        }
    }
    return in;
}

// Index of the r-th set bit of m (r < popcount(m)).
__device__ __forceinline__ int nth_set_bit(uint64_t m, int r)
{
    for (int b = 0; b < 64; ++b)
        if ((m >> b) & 1ull) {
            if (r == 0) return b;
            --r;
        }
    return 63;
}

// One wave per cell; particles of cell c go to slots start(c) + k, pid = slot.
__global__ __launch_bounds__(256) void k_fill(DevCase cs, SoA A, const int *__restrict__ lstart,
                                              const int *__restrict__ toff, const uint8_t *__restrict__ cut,
                                              const float *__restrict__ ux, const float *__restrict__ uy,
                                              const float *__restrict__ sig, Ctrl *__restrict__ ctrl)
{
    const int lane = lane_id();
    const int waves = gridDim.x * (blockDim.x >> 6);
    const int n_next = (int)ctrl->n_next;
    uint32_t err = 0;
    for (int c = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); c < cs.ncell; c += waves) {
        const int st = lstart[c] + toff[c / SCAN_TILE];
        const int en = c + 1 < cs.ncell ? lstart[c + 1] + toff[(c + 1) / SCAN_TILE] : n_next;
        const int ci = c / cs.ny, cj = c % cs.ny;
        const bool ct = cut[c] != 0;
        // uncut cells: the cell; cut cells: the bounding box of the fluid part
        const float x0 = ct ? cs.fbox[4 * c] : (float)ci * cs.dx;
        const float y0 = ct ? cs.fbox[4 * c + 1] : (float)cj * cs.dy;
        const float wx = ct ? cs.fbox[4 * c + 2] - cs.fbox[4 * c] : cs.dx;
        const float wy = ct ? cs.fbox[4 * c + 3] - cs.fbox[4 * c + 1] : cs.dy;
        const uint64_t fm = ct ? cs.fmask[c] : ~0ull;
        const int nm = __popcll(fm);
        const float sw = wx * 0.125f, sh = wy * 0.125f;
        for (int k = lane; k < en - st; k += 64) {
            float px = x0, py = y0;
            bool ok = false;
            if (!ct) {
                const uint4 w = philox((uint32_t)k, 0u, SUB_INIT, 1u, (uint32_t)c, cs.seed);
                px = x0 + wx * u01(w.x);
                py = y0 + wy * u01(w.y);
                ok = true;
            }
            // cut cells: a sub-box of the fluid bounding box that touches
            // fluid, a point in it, rejected if inside the body
            for (int tr = 0; tr < FILL_TRIES && !ok; ++tr) {
                const uint4 w = philox((uint32_t)k, 0u, SUB_INIT, 1u + (uint32_t)tr, (uint32_t)c, cs.seed);
                const int sb = nth_set_bit(fm, mulhi_idx(w.x, nm));
                px = x0 + sw * ((float)(sb & 7) + u01(w.y));
                py = y0 + sh * ((float)(sb >> 3) + u01(w.z));
                ok = !inside_body(cs, px, py);
            }
            if (!ok) err |= ERR_FILL;
            uint4 w = philox((uint32_t)k, 0u, SUB_INIT, 0u, (uint32_t)c, cs.seed);
            float r1 = sqrtf(-2.0f * dm_logf(u01o(w.x)));
            float s1, c1, s2, c2;
            dm_sincos2pi(u01(w.y), s1, c1);
            float r2 = sqrtf(-2.0f * dm_logf(u01o(w.z)));
            dm_sincos2pi(u01(w.w), s2, c2);
            const int i = st + k;
            A.x[i] = px;
            A.y[i] = py;
            A.vx[i] = ux[c] + sig[c] * (r1 * c1);
            A.vy[i] = uy[c] + sig[c] * (r1 * s1);
            A.vz[i] = sig[c] * (r2 * c2);
            A.pid[i] = (uint32_t)i;
            // rotational energy at the cell's kT/m = sig^2 (Run.cpp:134)
            if (cs.zeta != 0)
                A.er[i] = rot_energy(philox((uint32_t)k, 0u, SUB_INIT, 0x80000000u, (uint32_t)c, cs.seed).x,
                                     sig[c] * sig[c]);
        }
    }
    if (err) atomicOr(&ctrl->err, err);
}

// Sets the control block after a fill or an upload.
__global__ void k_ctrl_reset(Ctrl *ctrl, int use_n_next, uint32_t n, uint32_t next_pid)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        uint32_t nl = use_n_next ? ctrl->n_next : n;
        ctrl->step = 0u;
        ctrl->n_live = nl;
        ctrl->n_move = nl;
        ctrl->next_pid = use_n_next ? nl : next_pid;
        if (!use_n_next) ctrl->err = 0u;  // after a fill, keep the fill's error bits
        ctrl->scan_done = 0u;
        ctrl->cell_done = 0u;
        ctrl->tile_next = 0u;
        ctrl->step_coll = ctrl->last_coll = ctrl->step_cand = ctrl->last_cand = 0u;
        ctrl->psteps = ctrl->ncoll = ctrl->nsamples = ctrl->ncand = 0ull;
    }
}

// ---------------------------------------------------------------------------
// k_sample: one sample of the state after a step (GPUCells::Sample,
// GPUCell.cu:147-167,336-402): the state is in cell order, cell c holding
// [start[c], start[c+1]).  The sums are k_cell's: SAMPLE_G lanes per cell,
// each over every SAMPLE_G-th particle, a fixed butterfly, one FP64 atomic
// per sum from one lane -- so a sample taken here equals the one a step
// takes by itself, bit for bit.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_sample(DevCase cs, SoA A, SoA B, const int *__restrict__ start,
                                                const int *__restrict__ ncoll_cell, double *__restrict__ acc,
                                                Ctrl *__restrict__ ctrl)
{
    const uint32_t step = ctrl->step;
    const SoA S = (step & 1u) ? B : A;
    const int lane = lane_id(), g = lane / SAMPLE_G, sub = lane % SAMPLE_G;
    constexpr int CPW = 64 / SAMPLE_G;
    const int waves = gridDim.x * (blockDim.x >> 6);
    const size_t C = (size_t)cs.ncell;
    for (int base = (blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6)) * CPW; base < cs.ncell;
         base += waves * CPW) {
        const int c = base + g;
        const int o = c < cs.ncell ? start[c] : 0;
        const int k = c < cs.ncell ? start[c + 1] - o : 0;
        const bool rot = cs.zeta != 0;
        double sx = 0.0, sy = 0.0, sz = 0.0, s2 = 0.0, se = 0.0;
        for (int q = sub; q < k; q += SAMPLE_G) {
            const double u = (double)S.vx[o + q], v = (double)S.vy[o + q], w = (double)S.vz[o + q];
            sx += u; sy += v; sz += w;
            s2 += u * u + v * v + w * w;
            if (rot) se += (double)S.er[o + q];
        }
#pragma unroll
        for (int d = SAMPLE_G / 2; d > 0; d >>= 1) {
            sx += __shfl_xor(sx, d, SAMPLE_G);
            sy += __shfl_xor(sy, d, SAMPLE_G);
            sz += __shfl_xor(sz, d, SAMPLE_G);
            s2 += __shfl_xor(s2, d, SAMPLE_G);
            if (rot) se += __shfl_xor(se, d, SAMPLE_G);
        }
        if (k > 0 && sub == 0) {
            unsafeAtomicAdd(&acc[c], (double)k);
            unsafeAtomicAdd(&acc[C + c], sx);
            unsafeAtomicAdd(&acc[2 * C + c], sy);
            unsafeAtomicAdd(&acc[3 * C + c], sz);
            unsafeAtomicAdd(&acc[4 * C + c], s2);
            unsafeAtomicAdd(&acc[5 * C + c], (double)ncoll_cell[c]);
            if (rot) unsafeAtomicAdd(&acc[6 * C + c], se);
        }
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) ctrl->nsamples += 1ull;
}

// ---------------------------------------------------------------------------
// k_move_batch: k_move for many independent instances in one launch.  Live
// particles are ordered by instance (instance i at [P_i, P_i + n_i) with P
// the prefix of the live counts); inflow particles follow all live ones in
// instance order (instance i from N_l + Q_i, Q the prefix of the clamped
// inflow counts).  Every block rebuilds both offset tables in LDS; a particle
// finds its instance by binary search and is moved with that instance's
// constants, one instance at a time per wave (wave-uniform loads).  Within
// an instance the work is exactly that of k_move, so each instance evolves
// bit-identically to a run of its own.
// ---------------------------------------------------------------------------
__device__ __forceinline__ int find_inst(const uint32_t *tab, int ninst, uint32_t key)
{
    int lo = 0, hi = ninst - 1;   // largest i with tab[i] <= key
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (tab[mid] <= key) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// The batch's instance tables of a step in LDS: s_p = exclusive prefix of
// the live counts, s_q = of the (clamped) inflow counts, each with the total
// at [ninst].  The writer (block 0 of k_move_batch) also sets the move
// counts and the capacity flags.
__device__ void batch_tables(const DevCase *__restrict__ cases, int ninst, Ctrl *__restrict__ ctrl,
                             InstCtrl *__restrict__ ictl, uint32_t step, uint32_t *s_p, uint32_t *s_q, int *s_w,
                             bool writer)
{
    const int t = threadIdx.x;
    uint32_t nl = 0, jt = 0;
    bool clamped = false;
    if (t < ninst) {
        nl = ictl[t].n_live;
        int J[4];
        inject_counts(cases[t], step, J);
        jt = (uint32_t)(J[0] + J[1] + J[2] + J[3]);
        const uint32_t cap = cases[t].cap;
        if (nl + jt > cap) { jt = cap > nl ? cap - nl : 0u; clamped = true; }
    }
    int ex_p = 0, ex_q = 0;
    const int tot_p = block_excl_scan((int)nl, ex_p, s_w);
    const int tot_q = block_excl_scan((int)jt, ex_q, s_w);
    if (t < ninst) { s_p[t] = (uint32_t)ex_p; s_q[t] = (uint32_t)ex_q; }
    if (t == 0) { s_p[ninst] = (uint32_t)tot_p; s_q[ninst] = (uint32_t)tot_q; }
    if (writer) {
        if (t < ninst) {
            ictl[t].n_move = nl + jt;
            if (clamped) atomicOr(&ictl[t].err, (uint32_t)ERR_CAPACITY);
        }
        if (t == 0) ctrl->n_move = (uint32_t)(tot_p + tot_q);
    }
    __syncthreads();
}

// k_move_batch: the chunks (64 consecutive particles, one wave) that hold
// live particles of one instance only -- all but a few.  The instance is
// then the same for the whole wave, so its case is read with scalar loads
// and the wave needs no loop over instances; the chunks across an instance
// boundary and those holding inflow particles are k_move_batch_rest's.
// Keeping the instance loop and the inflow out of this kernel keeps its
// registers low enough for more waves per SIMD.
__global__ __launch_bounds__(256) void k_move_batch(const DevCase *__restrict__ cases, int ninst,
                                                    const int *__restrict__ cbase, SoA A, SoA B,
                                                    Bins bn, Ctrl *__restrict__ ctrl, InstCtrl *__restrict__ ictl)
{
    __shared__ uint32_t s_p[MAX_INST + 1], s_q[MAX_INST + 1];
    __shared__ int s_w[4];
    const uint32_t step = ctrl->step;
    const SoA S = (step & 1u) ? B : A;
    batch_tables(cases, ninst, ctrl, ictl, step, s_p, s_q, s_w, blockIdx.x == 0);
    const int lane = lane_id();
    const int nq = (int)(s_p[ninst] >> 6);            // chunks wholly of live particles
    const int nw = (int)(gridDim.x * (blockDim.x >> 6));
    int q = (int)(blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6));
    uint32_t err = 0;
    int last = 0;                                      // the previous chunk's instance
    // the next chunk's state loaded one chunk ahead
    float px = 0.0f, py = 0.0f, pvx = 0.0f, pvy = 0.0f, pvz = 0.0f;
    if (q < nq) {
        const uint32_t i = (uint32_t)q * 64u + (uint32_t)lane;
        px = S.x[i]; py = S.y[i]; pvx = S.vx[i]; pvy = S.vy[i]; pvz = S.vz[i];
    }
    RunPend pend{ 0ull, 0ull, 0, 0 };
    for (; q < nq; q += nw) {
        const uint32_t i = (uint32_t)q * 64u + (uint32_t)lane;
        float x = px, y = py, vx = pvx, vy = pvy, vz = pvz, er = 0.0f;
        if (q + nw < nq) {
            const uint32_t j = i + 64u * (uint32_t)nw;
            px = S.x[j]; py = S.y[j]; pvx = S.vx[j]; pvy = S.vy[j]; pvz = S.vz[j];
        }
        const uint32_t lo = (uint32_t)q * 64u;
        int u = (s_p[last] <= lo && lo < s_p[last + 1]) ? last : find_inst(s_p, ninst, lo);
        u = __builtin_amdgcn_readfirstlane(u);
        last = u;
        if (s_p[u + 1] < lo + 64u) continue;           // an instance boundary inside: the rest kernel's
        const DevCase &cs = cases[u];
        bool vchg = false, echg = false;
        // batched instances are monatomic (zeta = 0): er is never used
        const bool dead = move_one(cs, x, y, vx, vy, vz, er, S.pid + i, 0u, false, step, cs.dt, err, vchg, echg);
        S.x[i] = x; S.y[i] = y;
        if (vchg) { S.vx[i] = vx; S.vy[i] = vy; S.vz[i] = vz; }
        if (err) { atomicOr(&ictl[u].err, err); err = 0; }
        const RunPend r = agg_issue(bn, dead ? -1 : cbase[u] + cell_of(cs, x, y), i);
        agg_finish(bn, pend);
        pend = r;
    }
    agg_finish(bn, pend);
}

// k_move_batch_rest: the chunks k_move_batch leaves -- those across an
// instance boundary (listed once, by the first instance that starts inside
// them) and those at and past the end of the live particles, where the
// inflow particles are created.  A wave takes a chunk and runs its lanes one
// instance at a time.
__global__ __launch_bounds__(256) void k_move_batch_rest(const DevCase *__restrict__ cases, int ninst,
                                                         const int *__restrict__ cbase, SoA A, SoA B,
                                                         Bins bn, Ctrl *__restrict__ ctrl,
                                                         InstCtrl *__restrict__ ictl)
{
    __shared__ uint32_t s_p[MAX_INST + 1], s_q[MAX_INST + 1];
    __shared__ int s_w[4];
    __shared__ int s_chunk[MAX_INST];
    __shared__ int s_nb;
    const int t = threadIdx.x;
    const uint32_t step = ctrl->step;
    const SoA S = (step & 1u) ? B : A;
    batch_tables(cases, ninst, ctrl, ictl, step, s_p, s_q, s_w, false);
    const uint32_t n_live = s_p[ninst], ntot = n_live + s_q[ninst];
    const int nq = (int)(n_live >> 6), nend = (int)((ntot + 63u) >> 6);
    // the boundary chunks below nq, in instance order
    if (t == 0) s_nb = 0;
    __syncthreads();
    if (t >= 1 && t < ninst) {
        const uint32_t s = s_p[t], sp = s_p[t - 1];
        const bool first = (s & 63u) != 0u && (int)(s >> 6) < nq &&
                           !(t >= 2 && (sp & 63u) != 0u && (sp >> 6) == (s >> 6));
        if (first) s_chunk[atomicAdd(&s_nb, 1)] = (int)(s >> 6);
    }
    __syncthreads();
    const int nb = s_nb, nitems = nb + (nend - nq);
    const int lane = lane_id();
    const int nw = (int)(gridDim.x * (blockDim.x >> 6));
    uint32_t err = 0;
    for (int it = (int)(blockIdx.x * (blockDim.x >> 6) + (t >> 6)); it < nitems; it += nw) {
        const int q = it < nb ? s_chunk[it] : nq + (it - nb);
        const uint32_t i = (uint32_t)q * 64u + (uint32_t)lane;
        int c = -1;
        if (i < ntot) {
            const bool inflow = i >= n_live;
            const uint32_t key = inflow ? i - n_live : i;
            const int inst = find_inst(inflow ? s_q : s_p, ninst, key);
            for (bool todo = true; todo;) {
                const int u = __builtin_amdgcn_readfirstlane(inst);
                if (inst != u) continue;
                todo = false;
                const DevCase &cs = cases[u];
                float x, y, vx, vy, vz, er = 0.0f;
                bool dead;
                if (!inflow) {
                    x = S.x[i]; y = S.y[i]; vx = S.vx[i]; vy = S.vy[i]; vz = S.vz[i];
                    bool vchg = false, echg = false;
                    dead = move_one(cs, x, y, vx, vy, vz, er, S.pid + i, 0u, false, step, cs.dt, err, vchg, echg);
                    S.x[i] = x; S.y[i] = y;
                    if (vchg) { S.vx[i] = vx; S.vy[i] = vy; S.vz[i] = vz; }
                } else {
                    const uint32_t qi = key - s_q[u];
                    int J[4];
                    inject_counts(cs, step, J);
                    const uint32_t pid = ictl[u].next_pid + qi;
                    dead = inject_one(cs, inflow_face(J, qi), pid, step, x, y, vx, vy, vz, er, err);
                    S.x[i] = x; S.y[i] = y; S.vx[i] = vx; S.vy[i] = vy; S.vz[i] = vz; S.pid[i] = pid;
                }
                c = dead ? -1 : cbase[u] + cell_of(cs, x, y);
                if (err) { atomicOr(&ictl[u].err, err); err = 0; }
            }
        }
        agg_run(bn, c, i);
    }
}

// ---------------------------------------------------------------------------
// Launchers
// ---------------------------------------------------------------------------
hipError_t launch_move(const DevCase &cs, const SoA &A, const SoA &B, const Bins &bn, Ctrl *ctrl, int grid,
                       hipStream_t st)
{
    hipLaunchKernelGGL(k_move, dim3(grid), dim3(256), 0, st, cs, A, B, bn, ctrl);
    return hipGetLastError();
}

hipError_t launch_scan(int *count, int *lstart, int *tsum, int *toff, int ncell, Ctrl *ctrl,
                       hipStream_t st)
{
    int ntile = (ncell + SCAN_TILE - 1) / SCAN_TILE;
    hipLaunchKernelGGL(k_scan, dim3(ntile), dim3(256), 0, st, count, lstart, tsum, toff, ncell, ntile,
                       ctrl);
    return hipGetLastError();
}

hipError_t launch_scan_runs(const Bins &bn, int ncell, Ctrl *ctrl, hipStream_t st)
{
    const int ntile = (ncell + SCAN_TILE - 1) / SCAN_TILE;
    hipLaunchKernelGGL(k_scan_runs, dim3(ntile), dim3(256), 0, st, bn, ncell);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess || ntile <= TOFF_LDS) return e;
    hipLaunchKernelGGL(k_scan_tiles, dim3(1), dim3(256), 0, st, bn, ntile, ctrl);
    return hipGetLastError();
}

hipError_t launch_scatter(const Bins &bn, int *start, int *tile_first, int ncell, int tile_p, const Ctrl *ctrl,
                          int grid, hipStream_t st)
{
    if (tile_p < CELL_P_MIN || tile_p > CELL_CAP) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_place, dim3(grid), dim3(256), 0, st, bn, ncell, const_cast<Ctrl *>(ctrl));
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_rank, dim3(grid), dim3(256), 0, st, bn, start, reinterpret_cast<int4 *>(tile_first), ncell,
                       tile_p, ctrl);
    return hipGetLastError();
}

hipError_t launch_fill(const DevCase &cs, const SoA &A, const double *expected, const uint8_t *cut,
                       const float *ux, const float *uy, const float *sig, int *count, int *lstart,
                       int *tsum, int *toff, Ctrl *ctrl, int grid, hipStream_t st)
{
    hipLaunchKernelGGL(k_fill_count, dim3(grid), dim3(256), 0, st, expected, cs.ncell, cs.seed, count);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    e = launch_scan(count, lstart, tsum, toff, cs.ncell, ctrl, st);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_fill, dim3(grid), dim3(256), 0, st, cs, A, lstart, toff, cut, ux, uy, sig, ctrl);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_ctrl_reset, dim3(1), dim3(64), 0, st, ctrl, 1, 0u, 0u);
    return hipGetLastError();
}

hipError_t launch_move_batch(const DevCase *cases, int ninst, const int *cbase, const SoA &A, const SoA &B,
                             const Bins &bn, Ctrl *ctrl, InstCtrl *ictl, int grid, hipStream_t st)
{
    if (ninst < 1 || ninst > MAX_INST) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_move_batch, dim3(grid), dim3(256), 0, st, cases, ninst, cbase, A, B, bn, ctrl, ictl);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_move_batch_rest, dim3(grid), dim3(256), 0, st, cases, ninst, cbase, A, B, bn, ctrl, ictl);
    return hipGetLastError();
}

hipError_t launch_sample(const DevCase &cs, const SoA &A, const SoA &B, const int *start, const int *ncoll_cell,
                         double *acc, Ctrl *ctrl, hipStream_t st)
{
    const int cells_per_block = 4 * (64 / SAMPLE_G);
    const int grid = std::max(1, std::min((cs.ncell + cells_per_block - 1) / cells_per_block, 8192));
    hipLaunchKernelGGL(k_sample, dim3(grid), dim3(256), 0, st, cs, A, B, start, ncoll_cell, acc, ctrl);
    return hipGetLastError();
}

hipError_t launch_ctrl_reset(Ctrl *ctrl, uint32_t n, uint32_t next_pid, hipStream_t st)
{
    hipLaunchKernelGGL(k_ctrl_reset, dim3(1), dim3(64), 0, st, ctrl, 0, n, next_pid);
    return hipGetLastError();
}

// Signed distance from every cell centre to its instance's body outline
// (negative inside the body), one thread per cell; blockIdx.y is the
// instance.  Distance to the nearest edge point and the even-odd crossing
// rule for inside, in FP64, evaluated in the same order as a plain
// vectorised host loop over the edges so both agree to the last bits but for
// hypot's rounding.  Instances without a body get +inf.
__global__ void __launch_bounds__(256) k_wall_distance(const WallInst *inst, const double *poly, double *out)
{
    const WallInst w = inst[blockIdx.y];
    const int c = (int)(blockIdx.x * 256 + threadIdx.x);
    if (c >= w.ncell) return;
    const int i = c / w.ny, j = c - (c / w.ny) * w.ny;
    const double px = ((double)i + 0.5) * w.dx, py = ((double)j + 0.5) * w.dy;
    const double *v = poly + 2 * (size_t)w.poly0;
    double d = INFINITY;
    bool inside = false;
    for (int k = 0; k < w.npoly; ++k) {
        const int k1 = (k + 1 == w.npoly) ? 0 : k + 1;
        const double ax = v[2 * k], ay = v[2 * k + 1], bx = v[2 * k1], by = v[2 * k1 + 1];
        const double ex = bx - ax, ey = by - ay;
        double t = ((px - ax) * ex + (py - ay) * ey) / (ex * ex + ey * ey);
        t = fmin(fmax(t, 0.0), 1.0);
        d = fmin(d, hypot(px - ax - t * ex, py - ay - t * ey));
        if ((ay > py) != (by > py)) {
            const double xint = ax + (py - ay) * ex / ey;
            if (px < xint) inside = !inside;
        }
    }
    out[(size_t)w.cbase + c] = inside ? -d : d;
}

hipError_t launch_wall_distance(const WallInst *inst, int ninst, int max_cells, const double *poly, double *out,
                                hipStream_t st)
{
    if (ninst < 1 || max_cells < 1) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_wall_distance, dim3((max_cells + 255) / 256, ninst), dim3(256), 0, st, inst, poly, out);
    return hipGetLastError();
}

}  // namespace dsmc

namespace dsmc {

static int occ(const void *fn)
{
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, fn, 256, 0) != hipSuccess || nb < 1) nb = 1;
    return nb;
}
int occupancy_move() { return occ((const void *)k_move); }
int occupancy_move_batch() { return occ((const void *)k_move_batch); }
int occupancy_scatter() { return occ((const void *)k_rank); }

}  // namespace dsmc
