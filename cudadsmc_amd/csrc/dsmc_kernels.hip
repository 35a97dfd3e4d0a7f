// HIP kernels of the DSMC hot path for gfx950 (wave64, 160 KB LDS per CU).
//
// One timestep = four launches, no host round trip:
//   k_move    free flight + walls + body + outflow, inflow creation and the
//             atomic per-cell count                       (cudaDSMC GPUParticle.cu:229-363,
//                                                          :375-422, GPUCell.cu:169-198)
//   k_scan    exclusive scan of the counts, two-level, last-block finish
//                                                         (GPUCell.cu:104-120)
//   k_scatter cell-sorted permutation + start table + tile boundaries
//                                                         (GPUParticle.cu:365-373)
//   k_cell    per tile of cells: stable (ascending index) order inside each
//             cell, gather into cell order (this is also the outflow
//             compaction), NTC/VHS collisions, FP64 sampling
//                                                         (GPUCell.cu:200-402,
//                                                          GPUParticle.cu:76-140)
// All launches have fixed grids (persistent, grid-stride over the
// device-resident particle count), so a step can be captured in a hipGraph.
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

// ---------------------------------------------------------------------------
// k_move
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_move(DevCase cs, SoA A, SoA B, int2 *__restrict__ cellslot,
                                              int *__restrict__ count, Ctrl *__restrict__ ctrl)
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
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < ntot; i += stride) {
        float x, y, vx, vy, vz;
        bool dead;
        if (i < n_live) {
            x = S.x[i]; y = S.y[i]; vx = S.vx[i]; vy = S.vy[i]; vz = S.vz[i];
            bool vchg = false;
            dead = move_one(cs, x, y, vx, vy, vz, S.pid + i, 0u, false, step, cs.dt, err, vchg);
            S.x[i] = x; S.y[i] = y;
            if (vchg) { S.vx[i] = vx; S.vy[i] = vy; S.vz[i] = vz; }
        } else {
            uint32_t q = i - n_live;
            int f = 0;
            uint32_t acc = (uint32_t)J[0];
            while (f < 3 && q >= acc) { ++f; acc += (uint32_t)J[f]; }
            uint32_t pid = next_pid + q;
            dead = inject_one(cs, f, pid, step, x, y, vx, vy, vz, err);
            S.x[i] = x; S.y[i] = y; S.vx[i] = vx; S.vy[i] = vy; S.vz[i] = vz; S.pid[i] = pid;
        }
        int2 cslot;
        if (dead) {
            cslot = make_int2(-1, 0);
        } else {
            int c = cell_of(cs, x, y);
            cslot = make_int2(c, atomicAdd(&count[c], 1));
        }
        cellslot[i] = cslot;
    }
    if (err) atomicOr(&ctrl->err, err);
}

// ---------------------------------------------------------------------------
// k_scan: tile-local exclusive scan (SCAN_TILE cells per block), tile sums,
// and the last block to finish scans the tile sums.  count is zeroed for the
// next step.  start(c) = lstart[c] + toff[c / SCAN_TILE].
// ---------------------------------------------------------------------------
__device__ __forceinline__ int wave_incl_scan(int v)
{
    const int lane = lane_id();
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        int t = __shfl_up(v, d, 64);
        if (lane >= d) v += t;
    }
    return v;
}

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
// k_scatter: perm[start(cell)+slot] = i, materialise start[0..C], and record
// for every tile of CELL_P particles the first cell that starts in it.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_scatter(const int2 *__restrict__ cellslot,
                                                 const int *__restrict__ lstart,
                                                 const int *__restrict__ toff, int *__restrict__ start,
                                                 int *__restrict__ perm, int *__restrict__ tile_first,
                                                 int ncell, const Ctrl *__restrict__ ctrl)
{
    const uint32_t n_move = ctrl->n_move;
    const int n_next = (int)ctrl->n_next;
    const int ntiles = (n_next + CELL_P - 1) / CELL_P;
    const uint32_t stride = gridDim.x * blockDim.x;
    const uint32_t gid = blockIdx.x * blockDim.x + threadIdx.x;
    for (uint32_t i = gid; i < n_move; i += stride) {
        int2 cs = cellslot[i];
        if (cs.x >= 0) perm[lstart[cs.x] + toff[cs.x / SCAN_TILE] + cs.y] = (int)i;
    }
    for (uint32_t c = gid; c <= (uint32_t)ncell; c += stride) {
        int st = (int)c < ncell ? lstart[c] + toff[c / SCAN_TILE] : n_next;
        start[c] = st;
        int tlo = 0;
        if (c > 0) {
            int prev = lstart[c - 1] + toff[(c - 1) / SCAN_TILE];
            tlo = prev / CELL_P + 1;
        }
        int thi = st / CELL_P;
        if (thi > ntiles - 1) thi = ntiles - 1;
        for (int t = tlo; t <= thi; ++t) tile_first[t] = (int)c;
    }
}

// ---------------------------------------------------------------------------
// k_cell
// ---------------------------------------------------------------------------
struct CellArgs {
    DevCase cs;
    SoA A, B;
    const int *start;
    int *perm;
    const int *tile_first;
    float *sigcmax, *mrem;
    int *ncoll_cell;
    double *acc;
    Ctrl *ctrl;
    int sample_start;
    int keep_perm;
};

// Block-wide ascending sort of p[0..k) (flip bitonic network, virtual +inf
// padding).  GLOBAL selects device-scope fences between stages.
template <bool GLOBAL>
__device__ void block_sort(int *p, int k)
{
    int P2 = 1;
    while (P2 < k) P2 <<= 1;
    const int half = P2 >> 1;
    for (int size = 2; size <= P2; size <<= 1) {
        for (int j = size >> 1; j > 0; j >>= 1) {
            for (int ci = threadIdx.x; ci < half; ci += blockDim.x) {
                int a, b;
                if (j == (size >> 1)) {
                    int hs = size >> 1, blk = ci / hs, pos = ci - blk * hs;
                    a = blk * size + pos;
                    b = blk * size + size - 1 - pos;
                } else {
                    a = (ci / j) * 2 * j + (ci % j);
                    b = a + j;
                }
                if (b < k) {
                    int va = p[a], vb = p[b];
                    if (vb < va) { p[a] = vb; p[b] = va; }
                }
            }
            if (GLOBAL) __threadfence();
            __syncthreads();
        }
    }
}

// NTC/VHS collisions of one cell by one wave, candidates in rounds: each
// round takes the longest prefix of the next <= 64 candidates with no
// particle shared with an earlier candidate of the round, so the result is
// exactly the serial loop of the reference (GPUCell.cu:230-330), including
// the running (sigma c_r)max, which becomes a prefix max.
template <bool GLOBAL>
__device__ __forceinline__ int ntc_cell(const DevCase &cs, int c, uint32_t step, int np, float *Vx,
                                        float *Vy, float *Vz, float *sigcmax, float *mrem,
                                        uint32_t &ncand)
{
    if (np < 2) return 0;
    const float vol = cs.vol_ntc[c];
    if (!(vol > 0.0f)) return 0;
    const int lane = lane_id();
    float smax = sigcmax[c];
    float e = (float)((long long)np * (long long)(np - 1)) * cs.ntc_coef;
    e = e * smax;
    e = e / vol;
    const float tot = e + mrem[c];
    const int M = (int)tot;
    const float rem = tot - (float)M;
    int nc = 0;
    for (int mb = 0; mb < M;) {
        const int m = mb + lane;
        const int nact = min(64, M - mb);
        const bool act = lane < nact;
        int i1 = -1, i2 = -2;
        uint4 w = make_uint4(0u, 0u, 0u, 0u);
        if (act) {
            w = philox((uint32_t)m, step, SUB_COLLIDE, 0u, (uint32_t)c, cs.seed);
            i1 = mulhi_idx(w.x, np);
            i2 = mulhi_idx(w.y, np - 1);
            if (i2 >= i1) i2 = i2 + 1;
        }
        bool conf = false;
        for (int j = 0; j < nact; ++j) {
            int a = __shfl(i1, j, 64), b = __shfl(i2, j, 64);
            if (j < lane && (a == i1 || a == i2 || b == i1 || b == i2)) conf = true;
        }
        const unsigned long long cm = __ballot(act && conf);
        const int f = cm ? (int)__ffsll((long long)cm) - 1 : nact;
        const bool proc = lane < f;
        float sc = 0.0f, g2 = 0.0f;
        float v1x = 0.f, v1y = 0.f, v1z = 0.f, v2x = 0.f, v2y = 0.f, v2z = 0.f;
        if (proc) {
            v1x = Vx[i1]; v1y = Vy[i1]; v1z = Vz[i1];
            v2x = Vx[i2]; v2y = Vy[i2]; v2z = Vz[i2];
            float gx = v1x - v2x, gy = v1y - v2y, gz = v1z - v2z;
            g2 = gx * gx + gy * gy + gz * gz;
            sc = g2 > 0.0f ? cs.K1 * dm_expf(cs.e_exp * dm_logf(g2)) : 0.0f;
        }
        float pm = sc;
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            float t = __shfl_up(pm, d, 64);
            if (lane >= d) pm = fmaxf(pm, t);
        }
        const float sml = fmaxf(smax, pm);
        const bool accd = proc && (u01(w.z) < sc / sml);
        if (accd) {
            float cosr = 2.0f * u01(w.w) - 1.0f;
            float sinr = sqrtf(1.0f - cosr * cosr);
            uint4 w2 = philox((uint32_t)m, step, SUB_COLLIDE, 1u, (uint32_t)c, cs.seed);
            float sp, cp;
            dm_sincos2pi(u01(w2.x), sp, cp);
            float g = sqrtf(g2);
            float rx = g * cosr, ry = g * (sinr * cp), rz = g * (sinr * sp);
            float mx = 0.5f * (v1x + v2x), my = 0.5f * (v1y + v2y), mz = 0.5f * (v1z + v2z);
            Vx[i1] = mx + 0.5f * rx; Vx[i2] = mx - 0.5f * rx;
            Vy[i1] = my + 0.5f * ry; Vy[i2] = my - 0.5f * ry;
            Vz[i1] = mz + 0.5f * rz; Vz[i2] = mz - 0.5f * rz;
        }
        nc += __popcll(__ballot(accd));
        smax = fmaxf(smax, __shfl(pm, f - 1, 64));
        mb += f;
        if (GLOBAL) __threadfence_block();
        else __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        __builtin_amdgcn_wave_barrier();
    }
    ncand += (uint32_t)M;
    if (lane == 0) { sigcmax[c] = smax; mrem[c] = rem; }
    return nc;
}

// FP64 moment sums of one cell by one wave, added to the accumulators.
__device__ __forceinline__ void sample_cell(int c, int np, int nc, const float *Vx, const float *Vy,
                                            const float *Vz, double *acc, int ncell)
{
    const int lane = lane_id();
    double sx = 0.0, sy = 0.0, sz = 0.0, s2 = 0.0;
    for (int k = lane; k < np; k += 64) {
        double a = (double)Vx[k], b = (double)Vy[k], d = (double)Vz[k];
        sx += a; sy += b; sz += d;
        s2 += a * a + b * b + d * d;
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        sx += __shfl_xor(sx, o, 64);
        sy += __shfl_xor(sy, o, 64);
        sz += __shfl_xor(sz, o, 64);
        s2 += __shfl_xor(s2, o, 64);
    }
    if (lane == 0) {
        acc[c] += (double)np;
        acc[(size_t)ncell + c] += sx;
        acc[2 * (size_t)ncell + c] += sy;
        acc[3 * (size_t)ncell + c] += sz;
        acc[4 * (size_t)ncell + c] += s2;
        acc[5 * (size_t)ncell + c] += (double)nc;
    }
}

// One tile: cells [ca, cb), particles [p0, p0+n) of the sorted order.
// LDS = true: staging arrays are the block's LDS; false: they alias the
// destination state and perm in HBM (tiles larger than the LDS capacity).
template <bool LDS>
__device__ void cell_tile(const CellArgs &a, const SoA &cur, const SoA &nxt, uint32_t step,
                          bool sampling, int ca, int cb, int p0, int n, int *Sp, float *Vx, float *Vy,
                          float *Vz, int *s_big, int *s_nbig, uint32_t &wcoll, uint32_t &wcand)
{
    const int tid = threadIdx.x, lane = lane_id(), wave = tid >> 6;
    const int *__restrict__ start = a.start;
    if (LDS)
        for (int k = tid; k < n; k += blockDim.x) Sp[k] = a.perm[p0 + k];
    if (tid == 0) *s_nbig = 0;
    __syncthreads();
    // stable order inside each cell: sort the particle indices ascending
    for (int c = ca + wave; c < cb; c += CELL_WAVES) {
        const int o = start[c] - p0, k = start[c + 1] - start[c];
        if (k <= 1) continue;
        if (k > 64) {
            if (lane == 0) {
                int s = atomicAdd(s_nbig, 1);
                if (s < CELL_BIG) s_big[s] = c;
            }
            continue;
        }
        int v = lane < k ? Sp[o + lane] : 0x7fffffff;
        v = wave_sort64(v);
        if (lane < k) Sp[o + lane] = v;
    }
    if (!LDS) __threadfence();
    __syncthreads();
    const int nbig = *s_nbig;
    if (nbig <= CELL_BIG) {
        for (int b = 0; b < nbig; ++b) {
            const int c = s_big[b];
            block_sort<!LDS>(Sp + (start[c] - p0), start[c + 1] - start[c]);
        }
    } else {
        for (int c = ca; c < cb; ++c) {
            const int k = start[c + 1] - start[c];
            if (k > 64) block_sort<!LDS>(Sp + (start[c] - p0), k);
        }
    }
    __syncthreads();
    if (LDS && a.keep_perm)
        for (int k = tid; k < n; k += blockDim.x) a.perm[p0 + k] = Sp[k];
    // gather into cell order (also drops deleted particles)
    for (int k = tid; k < n; k += blockDim.x) {
        const int q = Sp[k];
        nxt.x[p0 + k] = cur.x[q];
        nxt.y[p0 + k] = cur.y[q];
        nxt.pid[p0 + k] = cur.pid[q];
        Vx[k] = cur.vx[q];
        Vy[k] = cur.vy[q];
        Vz[k] = cur.vz[q];
    }
    if (!LDS) __threadfence();
    __syncthreads();
    for (int c = ca + wave; c < cb; c += CELL_WAVES) {
        const int o = start[c] - p0, k = start[c + 1] - start[c];
        int nc = ntc_cell<!LDS>(a.cs, c, step, k, Vx + o, Vy + o, Vz + o, a.sigcmax, a.mrem, wcand);
        wcoll += (uint32_t)nc;
        if (lane == 0) a.ncoll_cell[c] = nc;
        if (sampling && k > 0) sample_cell(c, k, nc, Vx + o, Vy + o, Vz + o, a.acc, a.cs.ncell);
    }
    __syncthreads();
    if (LDS)
        for (int k = tid; k < n; k += blockDim.x) {
            nxt.vx[p0 + k] = Vx[k];
            nxt.vy[p0 + k] = Vy[k];
            nxt.vz[p0 + k] = Vz[k];
        }
    __syncthreads();
}

__global__ __launch_bounds__(256) void k_cell(CellArgs a)
{
    __shared__ int s_perm[CELL_CAP];
    __shared__ float s_vx[CELL_CAP], s_vy[CELL_CAP], s_vz[CELL_CAP];
    __shared__ int s_big[CELL_BIG];
    __shared__ int s_nbig;
    __shared__ uint32_t s_coll, s_cand;
    __shared__ int s_last;
    Ctrl *ctrl = a.ctrl;
    const uint32_t step = ctrl->step;
    const SoA cur = (step & 1u) ? a.B : a.A;
    const SoA nxt = (step & 1u) ? a.A : a.B;
    const int n_next = (int)ctrl->n_next;
    const int ntiles = (n_next + CELL_P - 1) / CELL_P;
    const bool sampling = a.sample_start >= 0 && step >= (uint32_t)a.sample_start;
    if (threadIdx.x == 0) { s_coll = 0u; s_cand = 0u; }
    uint32_t wcoll = 0, wcand = 0;
    for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
        const int ca = a.tile_first[t];
        const int cb = (t + 1 < ntiles) ? a.tile_first[t + 1] : a.cs.ncell;
        if (ca >= cb) continue;
        const int p0 = a.start[ca], n = a.start[cb] - p0;
        if (n <= CELL_CAP)
            cell_tile<true>(a, cur, nxt, step, sampling, ca, cb, p0, n, s_perm, s_vx, s_vy, s_vz, s_big,
                            &s_nbig, wcoll, wcand);
        else
            cell_tile<false>(a, cur, nxt, step, sampling, ca, cb, p0, n, a.perm + p0, nxt.vx + p0,
                             nxt.vy + p0, nxt.vz + p0, s_big, &s_nbig, wcoll, wcand);
    }
    if (lane_id() == 0 && (wcoll | wcand)) {
        atomicAdd(&s_coll, wcoll);
        atomicAdd(&s_cand, wcand);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        if (s_coll) atomicAdd(&ctrl->step_coll, s_coll);
        if (s_cand) atomicAdd(&ctrl->step_cand, s_cand);
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        s_last = (atomicAdd(&ctrl->cell_done, 1u) == gridDim.x - 1);
        if (s_last) {
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
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
            ctrl->cell_done = 0u;
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
        const float x0 = (float)ci * cs.dx, y0 = (float)cj * cs.dy;
        const bool ct = cut[c] != 0;
        for (int k = lane; k < en - st; k += 64) {
            float px = x0, py = y0;
            bool ok = false;
            for (int tr = 0; tr < FILL_TRIES && !ok; tr += 2) {
                uint4 w = philox((uint32_t)k, 0u, SUB_INIT, 1u + (uint32_t)(tr / 2), (uint32_t)c, cs.seed);
                px = x0 + cs.dx * u01(w.x);
                py = y0 + cs.dy * u01(w.y);
                if (!ct || !inside_body(cs, px, py)) { ok = true; break; }
                px = x0 + cs.dx * u01(w.z);
                py = y0 + cs.dy * u01(w.w);
                if (!ct || !inside_body(cs, px, py)) { ok = true; break; }
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
        ctrl->err = 0u;
        ctrl->scan_done = 0u;
        ctrl->cell_done = 0u;
        ctrl->step_coll = ctrl->last_coll = ctrl->step_cand = ctrl->last_cand = 0u;
        ctrl->psteps = ctrl->ncoll = ctrl->nsamples = ctrl->ncand = 0ull;
    }
}

// ---------------------------------------------------------------------------
// Launchers
// ---------------------------------------------------------------------------
hipError_t launch_move(const DevCase &cs, const SoA &A, const SoA &B, int2 *cellslot, int *count,
                       Ctrl *ctrl, int grid, hipStream_t st)
{
    hipLaunchKernelGGL(k_move, dim3(grid), dim3(256), 0, st, cs, A, B, cellslot, count, ctrl);
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

hipError_t launch_scatter(const int2 *cellslot, const int *lstart, const int *toff, int *start,
                          int *perm, int *tile_first, int ncell, const Ctrl *ctrl, int grid,
                          hipStream_t st)
{
    hipLaunchKernelGGL(k_scatter, dim3(grid), dim3(256), 0, st, cellslot, lstart, toff, start, perm,
                       tile_first, ncell, ctrl);
    return hipGetLastError();
}

hipError_t launch_cell(const DevCase &cs, const SoA &A, const SoA &B, const int *start, int *perm,
                       const int *tile_first, float *sigcmax, float *mrem, int *ncoll_cell,
                       double *acc, Ctrl *ctrl, int sample_start, int keep_perm, int grid,
                       hipStream_t st)
{
    CellArgs a{ cs, A, B, start, perm, tile_first, sigcmax, mrem, ncoll_cell, acc, ctrl, sample_start,
                keep_perm };
    hipLaunchKernelGGL(k_cell, dim3(grid), dim3(256), 0, st, a);
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

hipError_t launch_ctrl_reset(Ctrl *ctrl, uint32_t n, uint32_t next_pid, hipStream_t st)
{
    hipLaunchKernelGGL(k_ctrl_reset, dim3(1), dim3(64), 0, st, ctrl, 0, n, next_pid);
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
int occupancy_cell() { return occ((const void *)k_cell); }
int occupancy_scatter() { return occ((const void *)k_scatter); }

}  // namespace dsmc
