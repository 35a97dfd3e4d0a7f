// Stateless binning on caller device arrays (include/dsmc_abi.h,
// dsmc_bin_sort): cell ids, per-cell counts, their exclusive scan, the
// stable cell-sorted permutation and every particle's rank in its cell --
// what GPUCells::CalparticleNum + CalparticleStartIndex and
// GPUParticles::Sort compute (cudaDSMC GPUCell.cu:93-120,169-198,
// GPUParticle.cu:142-147,365-373), with the order inside each cell made
// deterministic: ascending particle index, the serial atomicAdd order.
//
// Kernels, one stream, no host round trip:
//   k_bs_clear       zero the per-cell counters and the control word
//   k_bs_bin         cell id per particle, wave-aggregated counts
//   k_scan           exclusive scan (the step's own two-level scan)
//   k_bs_scatter     start table; each particle into its cell's segment
//   k_bs_sort_small  each segment sorted ascending, one wave per cell
//                    (up to 64 particles, register bitonic network)
//   k_bs_sort_big    longer segments, one block per cell: bitonic network
//                    in LDS up to BS_LDS particles, in HBM beyond
// A segment holds exactly the indices of its cell's particles, so sorting it
// ascending gives the stable order whatever the order of the atomics.
//
// Scratch (per-cell counters, scan partials, the long-cell list, a control
// word) comes from hipMalloc, one buffer per call, in dsmc_bin_sort; from a
// static device array in dsmc_bin_sort_noalloc, for callers that cannot
// allocate.
#include <hip/hip_runtime.h>

#include <array>
#include <climits>
#include <mutex>

#include "../../include/dsmc_abi.h"
#include "host_common.h"

using namespace dsmc;

namespace {

constexpr int BS_LDS = 4096;                 // longest segment sorted in LDS
constexpr int BS_BLOCK = 256;

// cell index of GPUCell.cu:184-189 (i = (int)(x/dx), cell = j + i*ny),
// clamped into the grid like the step's binning
struct BinGrid {
    int nx, ny, ncell;
    float dx, dy;
};

__device__ __forceinline__ int bs_cell(const BinGrid &g, float x, float y)
{
    int i = (int)(x / g.dx);
    int j = (int)(y / g.dy);
    i = i < 0 ? 0 : (i > g.nx - 1 ? g.nx - 1 : i);
    j = j < 0 ? 0 : (j > g.ny - 1 ? g.ny - 1 : j);
    return j + i * g.ny;
}

__device__ __forceinline__ int bs_lane() { return (int)__lane_id(); }

// ascending bitonic sort of one int per lane over the 64 lanes
__device__ __forceinline__ int bs_wave_sort(int v)
{
    const int lane = bs_lane();
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

// Block-wide ascending sort of p[0..k): flip bitonic network, every
// comparator ascending, so indices >= k act as +inf without storage.
template <bool GLOBAL>
__device__ void bs_block_sort(int *p, int k)
{
    int P2 = 1;
    while (P2 < k) P2 <<= 1;
    const int half = P2 >> 1;
    for (int size = 2; size <= P2; size <<= 1) {
        for (int j = size >> 1; j > 0; j >>= 1) {
            for (int ci = threadIdx.x; ci < half; ci += blockDim.x) {
                int a, b;
                if (j == (size >> 1)) {
                    const int hs = size >> 1, blk = ci / hs, pos = ci - blk * hs;
                    a = blk * size + pos;
                    b = blk * size + size - 1 - pos;
                } else {
                    a = (ci / j) * 2 * j + (ci % j);
                    b = a + j;
                }
                if (b < k) {
                    const int va = p[a], vb = p[b];
                    if (vb < va) { p[a] = vb; p[b] = va; }
                }
            }
            if (GLOBAL) __threadfence();
            __syncthreads();
        }
    }
}

__global__ __launch_bounds__(BS_BLOCK) void k_bs_clear(int *__restrict__ count, int ncell, Ctrl *__restrict__ ctrl)
{
    for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < ncell; c += gridDim.x * blockDim.x) count[c] = 0;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        ctrl->scan_done = 0u;
        ctrl->n_next = 0u;
        ctrl->tile_next = 0u;      // length of the long-segment list
    }
}

// cell[i] and one atomic per run of equal cells in a wave
__global__ __launch_bounds__(BS_BLOCK) void k_bs_bin(BinGrid g, const float *__restrict__ x, const float *__restrict__ y,
                                                     int n, int *__restrict__ cell, int *__restrict__ count)
{
    const int stride = gridDim.x * blockDim.x;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const int c = bs_cell(g, x[i], y[i]);
        cell[i] = c;
        const int lane = bs_lane();
        unsigned long long todo = __ballot(1);
        int leader = lane, cnt = 0;
        while (todo) {
            const int l = (int)__ffsll((long long)todo) - 1;
            const int lc = __shfl(c, l, 64);
            const unsigned long long m = __ballot(c == lc);
            if (c == lc) leader = l;
            if (lane == l) cnt = __popcll(m);
            todo &= ~m;
        }
        if (lane == leader) atomicAdd(&count[c], cnt);
    }
}

// start[0..C] and every particle into its cell's segment; count was zeroed
// by the scan and now hands out places
__global__ __launch_bounds__(BS_BLOCK) void k_bs_scatter(const int *__restrict__ cell, const int *__restrict__ lstart,
                                                         const int *__restrict__ toff, int *__restrict__ start,
                                                         int *__restrict__ perm, int *__restrict__ count, int n,
                                                         int ncell, const Ctrl *__restrict__ ctrl)
{
    const int stride = gridDim.x * blockDim.x, gid = blockIdx.x * blockDim.x + threadIdx.x;
    for (int i = gid; i < n; i += stride) {
        const int c = cell[i];
        perm[lstart[c] + toff[c / SCAN_TILE] + atomicAdd(&count[c], 1)] = i;
    }
    for (int c = gid; c <= ncell; c += stride) start[c] = c < ncell ? lstart[c] + toff[c / SCAN_TILE] : (int)ctrl->n_next;
}

__global__ __launch_bounds__(BS_BLOCK) void k_bs_sort_small(const int *__restrict__ start, int *__restrict__ perm,
                                                            int *__restrict__ local_id, int ncell, int *__restrict__ big,
                                                            Ctrl *__restrict__ ctrl)
{
    const int lane = bs_lane();
    const int waves = gridDim.x * (blockDim.x >> 6);
    for (int c = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); c < ncell; c += waves) {
        const int o = start[c], len = start[c + 1] - o;
        if (len > 64) {
            if (lane == 0) big[atomicAdd(&ctrl->tile_next, 1u)] = c;
            continue;
        }
        int v = lane < len ? perm[o + lane] : INT_MAX;
        if (len > 1) v = bs_wave_sort(v);
        if (lane < len) {
            perm[o + lane] = v;
            if (local_id) local_id[v] = lane;
        }
    }
}

__global__ __launch_bounds__(BS_BLOCK) void k_bs_sort_big(const int *__restrict__ start, int *__restrict__ perm,
                                                          int *__restrict__ local_id, const int *__restrict__ big,
                                                          const Ctrl *__restrict__ ctrl)
{
    __shared__ int s[BS_LDS];
    const int nb = (int)ctrl->tile_next;
    for (int q = blockIdx.x; q < nb; q += gridDim.x) {
        const int c = big[q];
        const int o = start[c], len = start[c + 1] - o;
        int *seg = perm + o;
        if (len <= BS_LDS) {
            for (int k = threadIdx.x; k < len; k += blockDim.x) s[k] = seg[k];
            __syncthreads();
            bs_block_sort<false>(s, len);
            for (int k = threadIdx.x; k < len; k += blockDim.x) {
                seg[k] = s[k];
                if (local_id) local_id[s[k]] = k;
            }
        } else {
            bs_block_sort<true>(seg, len);
            if (local_id)
                for (int k = threadIdx.x; k < len; k += blockDim.x) local_id[seg[k]] = k;
        }
        __syncthreads();
    }
}

// Scratch of one call, carved from one buffer: count, lstart and the long
// list (ncell ints each), the two scan tile arrays, and the control word.
struct SortScratch {
    int *count, *lstart, *big, *tsum, *toff;
    Ctrl *ctrl;
};

size_t scratch_ints(int ncell)
{
    const size_t nt = (size_t)(ncell + SCAN_TILE - 1) / SCAN_TILE;
    size_t ints = 3 * (size_t)ncell + 2 * nt;
    return (ints + 63) / 64 * 64;                 // the control word 256-byte aligned
}

size_t scratch_bytes(int ncell) { return sizeof(int) * scratch_ints(ncell) + sizeof(Ctrl); }

SortScratch carve(void *base, int ncell)
{
    int *p = static_cast<int *>(base);
    const size_t C = (size_t)ncell, nt = (C + SCAN_TILE - 1) / SCAN_TILE;
    SortScratch s;
    s.count = p;
    s.lstart = p + C;
    s.big = p + 2 * C;
    s.tsum = p + 3 * C;
    s.toff = p + 3 * C + nt;
    s.ctrl = reinterpret_cast<Ctrl *>(p + scratch_ints(ncell));
    return s;
}

int check_args(const dsmc_grid *g, const float *x, const float *y, int64_t n, const int *cell, const int *start,
               const int *perm, BinGrid &bg)
{
    if (!g || !start || (n > 0 && (!x || !y || !cell || !perm))) return fail(DSMC_ERR_ARG, "bin_sort: null argument");
    if (n < 0 || n > (int64_t)INT_MAX) return fail(DSMC_ERR_ARG, "bin_sort: n must be 0..2^31-1");
    if (g->N1 < 1 || g->N2 < 1 || (int64_t)g->N1 * g->N2 > (int64_t)INT_MAX / 4)
        return fail(DSMC_ERR_ARG, "bin_sort: N1, N2 must be >= 1 and their product below 2^29");
    if (!(g->L1 > 0) || !(g->L2 > 0)) return fail(DSMC_ERR_ARG, "bin_sort: L1, L2 must be positive");
    bg.nx = g->N1;
    bg.ny = g->N2;
    bg.ncell = g->N1 * g->N2;
    bg.dx = (float)(g->L1 / g->N1);
    bg.dy = (float)(g->L2 / g->N2);
    return DSMC_OK;
}

int launch_all(const BinGrid &bg, const float *x, const float *y, int n, int *cell, int *start, int *perm,
               int *local_id, const SortScratch &S, hipStream_t st)
{
    const int C = bg.ncell;
    const int gc = std::max(1, std::min((C + BS_BLOCK - 1) / BS_BLOCK, 4096));
    const int gn = std::max(1, std::min((n + BS_BLOCK - 1) / BS_BLOCK, 8192));
    const int gw = std::max(1, std::min((C + 3) / 4, 16384));
    hipLaunchKernelGGL(k_bs_clear, dim3(gc), dim3(BS_BLOCK), 0, st, S.count, C, S.ctrl);
    HIPCHK(hipGetLastError());
    if (n > 0) {
        hipLaunchKernelGGL(k_bs_bin, dim3(gn), dim3(BS_BLOCK), 0, st, bg, x, y, n, cell, S.count);
        HIPCHK(hipGetLastError());
    }
    HIPCHK(launch_scan(S.count, S.lstart, S.tsum, S.toff, C, S.ctrl, st));
    hipLaunchKernelGGL(k_bs_scatter, dim3(std::max(gn, gc)), dim3(BS_BLOCK), 0, st, cell, S.lstart, S.toff, start,
                       perm, S.count, n, C, S.ctrl);
    HIPCHK(hipGetLastError());
    hipLaunchKernelGGL(k_bs_sort_small, dim3(gw), dim3(BS_BLOCK), 0, st, start, perm, local_id, C, S.big, S.ctrl);
    HIPCHK(hipGetLastError());
    hipLaunchKernelGGL(k_bs_sort_big, dim3(1024), dim3(BS_BLOCK), 0, st, start, perm, local_id, S.big, S.ctrl);
    HIPCHK(hipGetLastError());
    return DSMC_OK;
}

// The allocation-free variant's scratch: one array per device (each GPU the
// library runs on holds its own copy until the process ends).
constexpr size_t kNoallocBytes = sizeof(int) * ((3 * (size_t)DSMC_BIN_SORT_MAX_CELLS +
                                                 2 * (((size_t)DSMC_BIN_SORT_MAX_CELLS + SCAN_TILE - 1) / SCAN_TILE) +
                                                 63) / 64 * 64) +
                                 sizeof(Ctrl);
static __device__ __align__(256) unsigned char scratch[kNoallocBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
std::array<void *, kMaxDevices> g_scratch_ptr{};   // device address of `scratch` on each device

int scratch_on_device(void **out)
{
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) return fail(DSMC_ERR_ARG, "bin_sort_noalloc: device index above 63");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(DSMC_ERR_HIP, std::string("bin_sort_noalloc: hipGetSymbolAddress: ") + hipGetErrorString(e));
        g_scratch_ptr[dev] = p;
    }
    *out = g_scratch_ptr[dev];
    return DSMC_OK;
}

}  // namespace

extern "C" {

int dsmc_bin_sort(const dsmc_grid *g, const float *x, const float *y, int64_t n, int *cell, int *start, int *perm,
                  int *local_id, void *stream)
{
    BinGrid bg;
    int r = check_args(g, x, y, n, cell, start, perm, bg);
    if (r) return r;
    hipStream_t st = static_cast<hipStream_t>(stream);
    void *buf = nullptr;
    HIPCHK(hipMalloc(&buf, scratch_bytes(bg.ncell)));
    r = launch_all(bg, x, y, (int)n, cell, start, perm, local_id, carve(buf, bg.ncell), st);
    hipError_t e = hipStreamSynchronize(st);
    (void)hipFree(buf);
    if (r) return r;
    if (e != hipSuccess) return fail(DSMC_ERR_HIP, std::string("bin_sort: ") + hipGetErrorString(e));
    return DSMC_OK;
}

int dsmc_bin_sort_noalloc(const dsmc_grid *g, const float *x, const float *y, int64_t n, int *cell, int *start,
                          int *perm, int *local_id, void *stream)
{
    BinGrid bg;
    int r = check_args(g, x, y, n, cell, start, perm, bg);
    if (r) return r;
    if (bg.ncell > DSMC_BIN_SORT_MAX_CELLS)
        return fail(DSMC_ERR_ARG, "bin_sort_noalloc: at most " + std::to_string(DSMC_BIN_SORT_MAX_CELLS) +
                                      " cells (N1*N2); dsmc_bin_sort takes larger grids");
    void *p = nullptr;
    if ((r = scratch_on_device(&p))) return r;
    return launch_all(bg, x, y, (int)n, cell, start, perm, local_id, carve(p, bg.ncell),
                      static_cast<hipStream_t>(stream));
}

}  // extern "C"
