// Coordinate-MLP surrogate on gfx950 (include/dsmc_mlp.h): BF16 MFMA
// forward and backward kernels with FP32 accumulation, FP32 Adam.
//
// Every matrix product of a training step is one shape, C[m][n] =
// sum_k A[m][k] B[n][k], with both operands K-contiguous in memory:
//   forward          Z_l   = Y_l  . W_l^T      A = Y_l [B][K],  B = W_l  [N][K]
//   backward, data   dY_l  = dZ_l . W_l        A = dZ_l [B][N], B = W_l^T [K][N]
//   backward, weight dW_l  = dZ_l^T . Y_l      A = dZ_l^T [N][B], B = Y_l^T [K][B]
// so activations and output gradients are written in both orientations by
// the producing kernel's epilogue, and W is kept with its transpose.  One
// kernel template covers all three: a 256-thread block stages a BM x 32 tile
// of A and a BN x 32 tile of B in LDS per k-step and each of its 4 waves
// issues v_mfma_f32_16x16x32_bf16 on a (BM/WM) x (BN/WN) sub-tile.  Lane l
// of an MFMA holds A row l%16, k 8(l/16)..+7; B column l%16, the same k;
// accumulator rows 4(l/16)+i, column l%16 (checked on the device).  Weight
// gradients reduce over the batch with split-K and FP32 atomics.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/dsmc_mlp.h"
#include "dsmc_device.h"

namespace {

typedef __bf16 bf16;
typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;
typedef __attribute__((__vector_size__(4 * sizeof(float)))) float f32x4;

constexpr int KSTEP = 32;        // k per MFMA
constexpr int LDS_ROW = 40;      // bf16 per staged row: 32 + 8 padding (80 B, 16-B aligned)
constexpr int PAD = 32;          // every width is padded to a multiple of this

thread_local std::string g_err;
int fail(int code, const std::string &m)
{
    g_err = m;
    return code;
}
#define HCHK(expr)                                                                           \
    do {                                                                                     \
        hipError_t e_ = (expr);                                                              \
        if (e_ != hipSuccess) return fail(-2, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

enum Epi { EPI_SILU = 0, EPI_LINEAR = 1, EPI_DSILU = 2, EPI_SPLIT = 3 };

struct GemmArgs {
    const bf16 *A;
    const bf16 *B;
    int lda, ldb;
    int M, N, K, kchunk;
    const float *bias;       // SILU / LINEAR: [N]
    bf16 *Z, *Y, *YT;        // SILU: pre-activation and SiLU [M][N] (ld N), SiLU^T [N][M] (ld ldt)
    float *C;                // LINEAR: fp32 [M][N];  SPLIT: fp32 partials [gridDim.z][M][N]
    const bf16 *Zp;          // DSILU: pre-activation of the layer below [M][N]
    bf16 *D, *DT;            // DSILU: d loss / d Z below [M][N] and transposed [N][M] (ld ldt)
    int ldt;
};

__device__ __forceinline__ float silu(float z) { return z / (1.0f + __expf(-z)); }
__device__ __forceinline__ float dsilu(float z)
{
    const float s = 1.0f / (1.0f + __expf(-z));
    return s * (1.0f + z * (1.0f - s));
}

template <int BM, int BN>
struct Shape {
    static constexpr int WM = (BM >= 64 && BN >= 64) ? 2 : (BM >= 64 ? 4 : (BN >= 64 ? 1 : 2));
    static constexpr int WN = 4 / WM;
    static constexpr int TM = BM / (16 * WM);
    static constexpr int TN = BN / (16 * WN);
    static_assert(TM >= 1 && TN >= 1, "tile too small for 4 waves");
};

template <int BM, int BN, int EPI>
__global__ __launch_bounds__(256) void k_gemm(GemmArgs g)
{
    using S = Shape<BM, BN>;
    // the SiLU and dSiLU epilogues stage their outputs in LDS (see below)
    constexpr bool TR = EPI == EPI_SILU || EPI == EPI_DSILU;
    constexpr int TROW = BM + 8;                    // bf16 per transposed row: 16-B aligned, padded
    constexpr int STAGE = (BM + BN) * LDS_ROW;      // one k-step's A and B tiles
    constexpr int TSIZE = TR ? (BN * TROW > BM * (BN + 8) ? BN * TROW : BM * (BN + 8)) : 0;
    __shared__ __attribute__((aligned(16))) bf16 smem[2 * STAGE > TSIZE ? 2 * STAGE : TSIZE];
    bf16 *const sT = smem;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave / S::WN, wn = wave % S::WN;
    const int m0 = blockIdx.x * BM, n0 = blockIdx.y * BN;
    const int kb = blockIdx.z * g.kchunk;
    const int ke = min(g.K, kb + g.kchunk);
    f32x4 acc[S::TM][S::TN];
#pragma unroll
    for (int a = 0; a < S::TM; ++a)
#pragma unroll
        for (int b = 0; b < S::TN; ++b) acc[a][b] = f32x4{ 0.f, 0.f, 0.f, 0.f };
    const bf16x8 zero8 = { (bf16)0.f, (bf16)0.f, (bf16)0.f, (bf16)0.f, (bf16)0.f, (bf16)0.f, (bf16)0.f, (bf16)0.f };
    // two LDS stages: the tiles of k-step k+1 are loaded into registers while
    // the MFMAs of step k run, then stored into the other stage (one barrier
    // per k-step)
    constexpr int NA = (BM * 4 + 255) / 256, NB = (BN * 4 + 255) / 256;
    bf16x8 ra[NA], rb[NB];
    auto fetch = [&](int k0) {
#pragma unroll
        for (int i = 0; i < NA; ++i) {
            const int p = tid + 256 * i, r = p >> 2, kk = k0 + 8 * (p & 3);
            bf16x8 v = zero8;
            if (p < BM * 4 && m0 + r < g.M) {
                const bf16 *src = g.A + (size_t)(m0 + r) * g.lda + kk;
                if (kk + 8 <= g.K) v = *reinterpret_cast<const bf16x8 *>(src);
                else
                    for (int j = 0; j < 8 && kk + j < g.K; ++j) v[j] = src[j];
            }
            ra[i] = v;
        }
#pragma unroll
        for (int i = 0; i < NB; ++i) {
            const int p = tid + 256 * i, r = p >> 2, kk = k0 + 8 * (p & 3);
            bf16x8 v = zero8;
            if (p < BN * 4 && n0 + r < g.N) {
                const bf16 *src = g.B + (size_t)(n0 + r) * g.ldb + kk;
                if (kk + 8 <= g.K) v = *reinterpret_cast<const bf16x8 *>(src);
                else
                    for (int j = 0; j < 8 && kk + j < g.K; ++j) v[j] = src[j];
            }
            rb[i] = v;
        }
    };
    if (kb < ke) fetch(kb);
    int buf = 0;
    for (int k0 = kb; k0 < ke; k0 += KSTEP) {
        bf16 *const sA = smem + buf * STAGE, *const sB = sA + BM * LDS_ROW;
#pragma unroll
        for (int i = 0; i < NA; ++i) {
            const int p = tid + 256 * i;
            if (p < BM * 4) *reinterpret_cast<bf16x8 *>(sA + (p >> 2) * LDS_ROW + 8 * (p & 3)) = ra[i];
        }
#pragma unroll
        for (int i = 0; i < NB; ++i) {
            const int p = tid + 256 * i;
            if (p < BN * 4) *reinterpret_cast<bf16x8 *>(sB + (p >> 2) * LDS_ROW + 8 * (p & 3)) = rb[i];
        }
        __syncthreads();
        if (k0 + KSTEP < ke) fetch(k0 + KSTEP);
        bf16x8 af[S::TM], bfr[S::TN];
#pragma unroll
        for (int a = 0; a < S::TM; ++a)
            af[a] = *reinterpret_cast<const bf16x8 *>(sA + (wm * S::TM * 16 + a * 16 + (lane & 15)) * LDS_ROW +
                                                      8 * (lane >> 4));
#pragma unroll
        for (int b = 0; b < S::TN; ++b)
            bfr[b] = *reinterpret_cast<const bf16x8 *>(sB + (wn * S::TN * 16 + b * 16 + (lane & 15)) * LDS_ROW +
                                                       8 * (lane >> 4));
#pragma unroll
        for (int a = 0; a < S::TM; ++a)
#pragma unroll
            for (int b = 0; b < S::TN; ++b)
                acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[a], bfr[b], acc[a][b], 0, 0, 0);
        buf ^= 1;
    }
    __syncthreads();      // every wave is done with the stages before the epilogue reuses them
    if constexpr (!TR) {
#pragma unroll
        for (int a = 0; a < S::TM; ++a)
#pragma unroll
            for (int b = 0; b < S::TN; ++b)
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const int row = m0 + wm * S::TM * 16 + a * 16 + 4 * (lane >> 4) + i;
                    const int col = n0 + wn * S::TN * 16 + b * 16 + (lane & 15);
                    if (row >= g.M || col >= g.N) continue;
                    const size_t o = (size_t)row * g.N + col;
                    if (EPI == EPI_LINEAR) g.C[o] = acc[a][b][i] + g.bias[col];
                    else g.C[(size_t)blockIdx.z * g.M * g.N + o] = acc[a][b][i];
                }
    } else {
        // SiLU / SiLU' epilogues: every bf16 output goes through LDS (the
        // k-loop stages, reused: row-major with rows of BN + 8, or
        // transposed with rows of BM + 8) and leaves as contiguous 16-byte
        // runs; the layer below's pre-activation comes in the same way.
        // N is a multiple of BN here, so only rows need a bound.
        bf16 *const sR = smem;
        constexpr int RROW = BN + 8;
        auto rl_of = [&](int a, int i) { return wm * S::TM * 16 + a * 16 + 4 * (lane >> 4) + i; };
        auto cl_of = [&](int b) { return wn * S::TN * 16 + b * 16 + (lane & 15); };
        auto store_rows = [&](bf16 *dst) {          // sR -> dst [M][N]
            __syncthreads();
            for (int p = tid; p < BM * (BN / 8); p += 256) {
                const int r = p / (BN / 8), c8 = (p % (BN / 8)) * 8;
                if (m0 + r < g.M)
                    *reinterpret_cast<bf16x8 *>(dst + (size_t)(m0 + r) * g.N + n0 + c8) =
                        *reinterpret_cast<const bf16x8 *>(sR + r * RROW + c8);
            }
            __syncthreads();
        };
        auto store_cols = [&](bf16 *dst) {          // sT -> dst [N][ldt], transposed
            __syncthreads();
            for (int p = tid; p < BN * (BM / 8); p += 256) {
                const int c = p / (BM / 8), r8 = (p % (BM / 8)) * 8;
                const int row = m0 + r8;
                if (row >= g.M) continue;
                bf16 *const o = dst + (size_t)(n0 + c) * g.ldt + row;
                if (row + 8 <= g.M) *reinterpret_cast<bf16x8 *>(o) = *reinterpret_cast<const bf16x8 *>(sT + c * TROW + r8);
                else
                    for (int j = 0; j < 8 && row + j < g.M; ++j) o[j] = sT[c * TROW + r8 + j];
            }
            __syncthreads();
        };
        if constexpr (EPI == EPI_SILU) {
#pragma unroll
            for (int a = 0; a < S::TM; ++a)
#pragma unroll
                for (int b = 0; b < S::TN; ++b) {
                    const float bias = g.bias[n0 + cl_of(b)];
#pragma unroll
                    for (int i = 0; i < 4; ++i) acc[a][b][i] += bias;        // acc now holds z
                }
#pragma unroll
            for (int a = 0; a < S::TM; ++a)
#pragma unroll
                for (int b = 0; b < S::TN; ++b)
#pragma unroll
                    for (int i = 0; i < 4; ++i) sR[rl_of(a, i) * RROW + cl_of(b)] = (bf16)silu(acc[a][b][i]);
            store_rows(g.Y);
            if (g.YT) {
#pragma unroll
                for (int a = 0; a < S::TM; ++a)
#pragma unroll
                    for (int b = 0; b < S::TN; ++b)
#pragma unroll
                        for (int i = 0; i < 4; ++i) sT[cl_of(b) * TROW + rl_of(a, i)] = (bf16)silu(acc[a][b][i]);
                store_cols(g.YT);
            }
#pragma unroll
            for (int a = 0; a < S::TM; ++a)
#pragma unroll
                for (int b = 0; b < S::TN; ++b)
#pragma unroll
                    for (int i = 0; i < 4; ++i) sR[rl_of(a, i) * RROW + cl_of(b)] = (bf16)acc[a][b][i];
            store_rows(g.Z);
        } else {
            // the layer below's pre-activation, row-major, into sR
            for (int p = tid; p < BM * (BN / 8); p += 256) {
                const int r = p / (BN / 8), c8 = (p % (BN / 8)) * 8;
                if (m0 + r < g.M)
                    *reinterpret_cast<bf16x8 *>(sR + r * RROW + c8) =
                        *reinterpret_cast<const bf16x8 *>(g.Zp + (size_t)(m0 + r) * g.N + n0 + c8);
            }
            __syncthreads();
#pragma unroll
            for (int a = 0; a < S::TM; ++a)
#pragma unroll
                for (int b = 0; b < S::TN; ++b)
#pragma unroll
                    for (int i = 0; i < 4; ++i)
                        acc[a][b][i] = (float)(bf16)(acc[a][b][i] * dsilu((float)sR[rl_of(a, i) * RROW + cl_of(b)]));
            __syncthreads();
#pragma unroll
            for (int a = 0; a < S::TM; ++a)
#pragma unroll
                for (int b = 0; b < S::TN; ++b)
#pragma unroll
                    for (int i = 0; i < 4; ++i) sR[rl_of(a, i) * RROW + cl_of(b)] = (bf16)acc[a][b][i];
            store_rows(g.D);
#pragma unroll
            for (int a = 0; a < S::TM; ++a)
#pragma unroll
                for (int b = 0; b < S::TN; ++b)
#pragma unroll
                    for (int i = 0; i < 4; ++i) sT[cl_of(b) * TROW + rl_of(a, i)] = (bf16)acc[a][b][i];
            store_cols(g.DT);
        }
    }
}

// Sum of the split-K partials: out[i] = sum over s of part[s][i] (n a
// multiple of 4).  A block covers 256 outputs as 64 float4 columns; its four
// waves take every fourth split (float4 loads, unrolled), and the four
// partial sums are added in LDS in wave order, so the result is
// deterministic.
__global__ __launch_bounds__(256) void k_sum_splits(const float *__restrict__ part, int splits, int n,
                                                    float *__restrict__ out)
{
    __shared__ float4 s_p[256];
    const int t = threadIdx.x, c = t & 63, ph = t >> 6;
    const int i4 = blockIdx.x * 64 + c;
    const bool in = 4 * i4 < n;
    float4 s = make_float4(0.f, 0.f, 0.f, 0.f);
    if (in) {
#pragma unroll 4
        for (int k = ph; k < splits; k += 4) {
            const float4 v = *reinterpret_cast<const float4 *>(part + (size_t)k * n + 4 * (size_t)i4);
            s.x += v.x; s.y += v.y; s.z += v.z; s.w += v.w;
        }
    }
    s_p[t] = s;
    __syncthreads();
    if (ph == 0 && in) {
        float4 r = s_p[c];
        for (int q = 1; q < 4; ++q) {
            const float4 v = s_p[q * 64 + c];
            r.x += v.x; r.y += v.y; r.z += v.z; r.w += v.w;
        }
        *reinterpret_cast<float4 *>(out + 4 * (size_t)i4) = r;
    }
}

// Minibatch rows: index from Philox (with replacement), or the identity.
__global__ void k_gather(const float *__restrict__ X, const float *__restrict__ T, int64_t S, int nin, int nout,
                         int B, int Kp, uint32_t seed_lo, uint32_t seed_hi, uint32_t step, int random,
                         bf16 *__restrict__ X0, bf16 *__restrict__ X0T, float *__restrict__ Tb)
{
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= B) return;
    int64_t s = b;
    if (random) {
        const uint4 w = dsmc::philox((uint32_t)b, step, 7u, 0u, seed_lo, seed_hi);
        s = (int64_t)(((uint64_t)w.x * (uint64_t)S) >> 32);   // S < 2^32
    }
    for (int k = 0; k < Kp; ++k) {
        const bf16 v = (bf16)(k < nin ? X[s * nin + k] : 0.0f);
        X0[(size_t)b * Kp + k] = v;
        X0T[(size_t)k * B + b] = v;
    }
    if (T)
        for (int j = 0; j < nout; ++j) Tb[(size_t)b * nout + j] = T[s * nout + j];
}

// Mean squared error over B rows and nout outputs; its gradient, zero-padded
// to Np columns, in both orientations.
__global__ void k_loss(const float *__restrict__ out, const float *__restrict__ Tb, int B, int nout, int Np,
                       bf16 *__restrict__ D, bf16 *__restrict__ DT, float *__restrict__ loss)
{
    __shared__ float s_w[4];
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    float l = 0.f;
    if (b < B) {
        const float scale = 2.0f / ((float)B * (float)nout);
        for (int c = 0; c < Np; ++c) {
            float d = 0.f;
            if (c < nout) {
                const float e = out[(size_t)b * Np + c] - Tb[(size_t)b * nout + c];
                l += e * e;
                d = scale * e;
            }
            D[(size_t)b * Np + c] = (bf16)d;
            DT[(size_t)c * B + b] = (bf16)d;
        }
    }
    for (int o = 32; o > 0; o >>= 1) l += __shfl_down(l, o, 64);
    if ((threadIdx.x & 63) == 0) s_w[threadIdx.x >> 6] = l;
    __syncthreads();
    if (threadIdx.x == 0) {
        float t = 0.f;
        for (int w = 0; w < (int)(blockDim.x >> 6); ++w) t += s_w[w];
        atomicAdd(loss, t / ((float)B * (float)nout));
    }
}

// Bias gradient, first stage: column sums of dZ [B][Np] (row-major, Np a
// multiple of 32) over `rows` rows per block, written to part[block][Np]
// (k_sum_splits adds the blocks).  Threads read 16-byte runs of 8 columns,
// Np/8 threads per row and 256/(Np/8) rows at a time, so every row is read
// contiguously; the row phases are reduced in LDS.
__global__ __launch_bounds__(256) void k_colsum(const bf16 *__restrict__ D, int B, int Np, int rows,
                                                float *__restrict__ part)
{
    __shared__ float s_p[2048];
    const int tid = threadIdx.x;
    const int r0 = blockIdx.x * rows, r1 = min(B, r0 + rows);
    const int vpr = Np / 8;
    float *const out = part + (size_t)blockIdx.x * Np;
    if (vpr > 256) {                      // wide layers: one column per thread
        for (int c = tid; c < Np; c += 256) {
            float s = 0.f;
            for (int r = r0; r < r1; ++r) s += (float)D[(size_t)r * Np + c];
            out[c] = s;
        }
        return;
    }
    const int phases = 256 / vpr, v = tid % vpr, ph = tid / vpr;
    float s[8] = { 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f };
    if (ph < phases)
#pragma unroll 4
        for (int r = r0 + ph; r < r1; r += phases) {
            const bf16x8 x = *reinterpret_cast<const bf16x8 *>(D + (size_t)r * Np + 8 * v);
#pragma unroll
            for (int j = 0; j < 8; ++j) s[j] += (float)x[j];
        }
    // phases * Np <= 2048 floats
    if (ph < phases)
#pragma unroll
        for (int j = 0; j < 8; ++j) s_p[ph * Np + 8 * v + j] = s[j];
    __syncthreads();
    for (int c = tid; c < Np; c += 256) {
        float t = 0.f;
        for (int k = 0; k < phases; ++k) t += s_p[k * Np + c];
        out[c] = t;
    }
}

// Adam on the real (unpadded) entries of one layer; writes the BF16 weights
// and their transpose, clears the weight-gradient accumulator.
__global__ void k_adam(float *__restrict__ w, float *__restrict__ m, float *__restrict__ v, float *__restrict__ gW,
                       float *__restrict__ gb, int nout, int nin, int Np, int Kp, float lr_t, float b1, float b2,
                       float eps, bf16 *__restrict__ W, bf16 *__restrict__ WT, float *__restrict__ bias)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    const int nw = nout * nin;
    if (p >= nw + nout) return;
    float g;
    int o = 0, i = 0;
    if (p < nw) {
        o = p / nin;
        i = p - o * nin;
        g = gW[(size_t)o * Kp + i];
        gW[(size_t)o * Kp + i] = 0.f;
    } else {
        g = gb[p - nw];
        gb[p - nw] = 0.f;
    }
    const float mm = b1 * m[p] + (1.f - b1) * g;
    const float vv = b2 * v[p] + (1.f - b2) * g * g;
    m[p] = mm;
    v[p] = vv;
    const float nwv = w[p] - lr_t * mm / (sqrtf(vv) + eps);
    w[p] = nwv;
    if (p < nw) {
        W[(size_t)o * Kp + i] = (bf16)nwv;
        WT[(size_t)i * Np + o] = (bf16)nwv;
    } else {
        bias[p - nw] = nwv;
    }
}

// FP32 master parameters of one layer -> BF16 weights, transpose, FP32 bias.
__global__ void k_publish(const float *__restrict__ w, int nout, int nin, int Np, int Kp, bf16 *__restrict__ W,
                          bf16 *__restrict__ WT, float *__restrict__ bias)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    const int nw = nout * nin;
    if (p >= nw + nout) return;
    if (p < nw) {
        const int o = p / nin, i = p - o * nin;
        W[(size_t)o * Kp + i] = (bf16)w[p];
        WT[(size_t)i * Np + o] = (bf16)w[p];
    } else {
        bias[p - nw] = w[p];
    }
}

int pad(int v) { return (v + PAD - 1) / PAD * PAD; }

}  // namespace

struct dsmc_mlp_s {
    int L = 0, maxB = 0;
    std::vector<int> width, Wp;           // real and padded widths, L + 1
    std::vector<int64_t> poff;            // offset of layer l in the flat parameter array
    int64_t nparams = 0;
    hipStream_t st = nullptr;
    float *w = nullptr, *m = nullptr, *v = nullptr;        // FP32 master, Adam moments (flat)
    std::vector<bf16 *> W, WT;                              // [Np][Kp], [Kp][Np]
    std::vector<float *> bias, gW, gb;                      // [Np], [Np][Kp], [Np]
    std::vector<bf16 *> Y, YT, Z, D, DT;                    // activations and gradients per level
    float *out = nullptr, *Tb = nullptr, *loss = nullptr;
    float *ws = nullptr;                                    // split-K partials of the weight gradients
    float *dX = nullptr, *dT = nullptr;                     // resident training set
    int64_t nset = 0;
    int64_t adam_t = 0;
    double ms_step = 0, flop_step = 0;
    std::vector<void *> allocs;
    hipEvent_t ev[2]{};

    template <class U>
    hipError_t alloc(U **p, size_t n)
    {
        hipError_t e = hipMalloc((void **)p, sizeof(U) * std::max<size_t>(n, 1));
        if (e == hipSuccess) {
            allocs.push_back((void *)*p);
            e = hipMemset(*p, 0, sizeof(U) * std::max<size_t>(n, 1));
        }
        return e;
    }
};

namespace {

template <int BM, int BN, int EPI>
hipError_t gemm(const GemmArgs &g, int ksplit, hipStream_t st)
{
    dim3 grid((g.M + BM - 1) / BM, (g.N + BN - 1) / BN, ksplit);
    hipLaunchKernelGGL((k_gemm<BM, BN, EPI>), grid, dim3(256), 0, st, g);
    return hipGetLastError();
}

// Picks the tile for the shape: 128-wide where the dimension allows.
template <int EPI>
hipError_t gemm_any(GemmArgs g, int ksplit, hipStream_t st)
{
    const bool bm = g.M % 128 == 0 || g.M > 4096, bn = g.N % 128 == 0;
    if (bm && bn) return gemm<128, 128, EPI>(g, ksplit, st);
    if (bm) return gemm<128, 32, EPI>(g, ksplit, st);
    if (bn) return gemm<32, 128, EPI>(g, ksplit, st);
    return gemm<32, 32, EPI>(g, ksplit, st);
}

int ksplit_for(int K)
{
    // weight gradients: split the batch so that the grid fills the device
    int s = K / 1024;
    return std::max(1, std::min(s, 256));
}

// Forward over B rows already gathered into Y[0] / YT[0]; fp32 output in m->out.
int forward(dsmc_mlp m, int B, bool keep_t)
{
    for (int l = 0; l < m->L; ++l) {
        GemmArgs g{};
        g.A = m->Y[l];
        g.lda = m->Wp[l];
        g.B = m->W[l];
        g.ldb = m->Wp[l];
        g.M = B;
        g.N = m->Wp[l + 1];
        g.K = m->Wp[l];
        g.kchunk = g.K;
        g.bias = m->bias[l];
        if (l + 1 < m->L) {
            g.Z = m->Z[l + 1];
            g.Y = m->Y[l + 1];
            g.YT = keep_t ? m->YT[l + 1] : nullptr;
            g.ldt = B;
            HCHK(gemm_any<EPI_SILU>(g, 1, m->st));
        } else {
            g.C = m->out;
            HCHK(gemm_any<EPI_LINEAR>(g, 1, m->st));
        }
    }
    return 0;
}

// Backward from D[L] / DT[L] (output gradient) down to the weight gradients.
int backward(dsmc_mlp m, int B)
{
    for (int l = m->L - 1; l >= 0; --l) {
        const int Np = m->Wp[l + 1], Kp = m->Wp[l];
        GemmArgs g{};
        g.A = m->DT[l + 1];
        g.lda = B;
        g.B = m->YT[l];
        g.ldb = B;
        g.M = Np;
        g.N = Kp;
        g.K = B;
        const int ks = ksplit_for(B);
        g.kchunk = ((B + ks - 1) / ks + KSTEP - 1) / KSTEP * KSTEP;
        const int splits = (B + g.kchunk - 1) / g.kchunk;
        g.C = m->ws;
        HCHK(gemm_any<EPI_SPLIT>(g, splits, m->st));
        hipLaunchKernelGGL(k_sum_splits, dim3((Np * Kp + 255) / 256), dim3(256), 0, m->st, m->ws, splits, Np * Kp,
                           m->gW[l]);      // Np, Kp multiples of 32
        HCHK(hipGetLastError());
        // bias gradient: per-block column sums into the workspace (free again
        // after the weight-gradient sum), then their sum
        const int rows = std::max(256, (B + 127) / 128), nblk = (B + rows - 1) / rows;
        hipLaunchKernelGGL(k_colsum, dim3(nblk), dim3(256), 0, m->st, m->D[l + 1], B, Np, rows, m->ws);
        HCHK(hipGetLastError());
        hipLaunchKernelGGL(k_sum_splits, dim3((Np + 255) / 256), dim3(256), 0, m->st, m->ws, nblk, Np, m->gb[l]);
        HCHK(hipGetLastError());
        if (l > 0) {
            GemmArgs d{};
            d.A = m->D[l + 1];
            d.lda = Np;
            d.B = m->WT[l];
            d.ldb = Np;
            d.M = B;
            d.N = Kp;
            d.K = Np;
            d.kchunk = Np;
            d.Zp = m->Z[l];
            d.D = m->D[l];
            d.DT = m->DT[l];
            d.ldt = B;
            HCHK(gemm_any<EPI_DSILU>(d, 1, m->st));
        }
    }
    return 0;
}

int gather(dsmc_mlp m, const float *X, const float *T, int64_t S, int B, uint64_t seed, uint32_t step, bool random)
{
    hipLaunchKernelGGL(k_gather, dim3((B + 255) / 256), dim3(256), 0, m->st, X, T, S, m->width[0],
                       m->width[m->L], B, m->Wp[0], (uint32_t)seed, (uint32_t)(seed >> 32), step, random ? 1 : 0,
                       m->Y[0], m->YT[0], m->Tb);
    HCHK(hipGetLastError());
    return 0;
}

int loss_and_grad(dsmc_mlp m, int B)
{
    HCHK(hipMemsetAsync(m->loss, 0, sizeof(float), m->st));
    hipLaunchKernelGGL(k_loss, dim3((B + 255) / 256), dim3(256), 0, m->st, m->out, m->Tb, B, m->width[m->L],
                       m->Wp[m->L], m->D[m->L], m->DT[m->L], m->loss);
    HCHK(hipGetLastError());
    return 0;
}

int adam_step(dsmc_mlp m, float lr, float b1, float b2, float eps)
{
    m->adam_t += 1;
    const double t = (double)m->adam_t;
    const float lr_t = (float)(lr * std::sqrt(1.0 - std::pow((double)b2, t)) / (1.0 - std::pow((double)b1, t)));
    for (int l = 0; l < m->L; ++l) {
        const int nout = m->width[l + 1], nin = m->width[l];
        const int64_t o = m->poff[l];
        const int n = nout * nin + nout;
        hipLaunchKernelGGL(k_adam, dim3((n + 255) / 256), dim3(256), 0, m->st, m->w + o, m->m + o, m->v + o,
                           m->gW[l], m->gb[l], nout, nin, m->Wp[l + 1], m->Wp[l], lr_t, b1, b2, eps, m->W[l],
                           m->WT[l], m->bias[l]);
        HCHK(hipGetLastError());
    }
    return 0;
}

double step_flop(dsmc_mlp m, int B)
{
    double macs = 0, back = 0;
    for (int l = 0; l < m->L; ++l) {
        const double mk = (double)m->width[l] * m->width[l + 1];
        macs += mk;
        back += (l > 0 ? 2.0 : 1.0) * mk;     // dW for every layer, dY for all but the first
    }
    return 2.0 * (double)B * (macs + back);
}

}  // namespace

extern "C" {

int dsmc_mlp_create(const int *widths, int nlayers, int max_batch, dsmc_mlp *out)
{
    if (!widths || !out || nlayers < 1 || max_batch < 1) return fail(-1, "bad argument");
    for (int l = 0; l <= nlayers; ++l)
        if (widths[l] < 1) return fail(-1, "widths must be positive");
    if (widths[0] > PAD || widths[nlayers] > PAD) return fail(-1, "at most 32 inputs and 32 outputs");
    dsmc_mlp m = new dsmc_mlp_s();
    m->L = nlayers;
    m->maxB = (max_batch + 127) / 128 * 128;
    m->width.assign(widths, widths + nlayers + 1);
    for (int l = 0; l <= nlayers; ++l) m->Wp.push_back(pad(widths[l]));
    int64_t off = 0;
    for (int l = 0; l < nlayers; ++l) {
        m->poff.push_back(off);
        off += (int64_t)widths[l] * widths[l + 1] + widths[l + 1];
    }
    m->nparams = off;
    auto bad = [&](hipError_t e) {
        std::string s = std::string("mlp allocation: ") + hipGetErrorString(e);
        dsmc_mlp_destroy(m);
        return fail(-2, s);
    };
    hipError_t e;
    if ((e = hipStreamCreateWithFlags(&m->st, hipStreamNonBlocking)) != hipSuccess) return bad(e);
    for (auto &ev : m->ev)
        if ((e = hipEventCreate(&ev)) != hipSuccess) return bad(e);
    if ((e = m->alloc(&m->w, off)) || (e = m->alloc(&m->m, off)) || (e = m->alloc(&m->v, off)) ||
        (e = m->alloc(&m->out, (size_t)m->maxB * m->Wp[nlayers])) ||
        (e = m->alloc(&m->Tb, (size_t)m->maxB * widths[nlayers])) || (e = m->alloc(&m->loss, 1)))
        return bad(e);
    m->W.resize(nlayers); m->WT.resize(nlayers); m->bias.resize(nlayers); m->gW.resize(nlayers); m->gb.resize(nlayers);
    for (int l = 0; l < nlayers; ++l) {
        const size_t wsz = (size_t)m->Wp[l + 1] * m->Wp[l];
        if ((e = m->alloc(&m->W[l], wsz)) || (e = m->alloc(&m->WT[l], wsz)) || (e = m->alloc(&m->bias[l], m->Wp[l + 1])) ||
            (e = m->alloc(&m->gW[l], wsz)) || (e = m->alloc(&m->gb[l], m->Wp[l + 1])))
            return bad(e);
    }
    {
        // split-K workspace: the largest layer at the largest split count
        size_t wsz = 1;
        for (int l = 0; l < nlayers; ++l) wsz = std::max(wsz, (size_t)m->Wp[l + 1] * m->Wp[l]);
        if ((e = m->alloc(&m->ws, wsz * (size_t)(ksplit_for(m->maxB) + 1)))) return bad(e);
    }
    m->Y.assign(nlayers + 1, nullptr); m->YT.assign(nlayers + 1, nullptr); m->Z.assign(nlayers + 1, nullptr);
    m->D.assign(nlayers + 1, nullptr); m->DT.assign(nlayers + 1, nullptr);
    for (int l = 0; l <= nlayers; ++l) {
        const size_t a = (size_t)m->maxB * m->Wp[l];
        if (l < nlayers && ((e = m->alloc(&m->Y[l], a)) || (e = m->alloc(&m->YT[l], a)))) return bad(e);
        if (l > 0 && l < nlayers && (e = m->alloc(&m->Z[l], a))) return bad(e);
        if (l > 0 && ((e = m->alloc(&m->D[l], a)) || (e = m->alloc(&m->DT[l], a)))) return bad(e);
    }
    *out = m;
    return 0;
}

int dsmc_mlp_destroy(dsmc_mlp m)
{
    if (!m) return 0;
    if (m->st) (void)hipStreamSynchronize(m->st);
    for (void *p : m->allocs) (void)hipFree(p);
    // the training set is owned apart from allocs: set_data replaces it
    if (m->dX) (void)hipFree(m->dX);
    if (m->dT) (void)hipFree(m->dT);
    for (auto &ev : m->ev)
        if (ev) (void)hipEventDestroy(ev);
    if (m->st) (void)hipStreamDestroy(m->st);
    delete m;
    return 0;
}

int dsmc_mlp_nparams(dsmc_mlp m, int64_t *n)
{
    if (!m || !n) return fail(-1, "null argument");
    *n = m->nparams;
    return 0;
}

int dsmc_mlp_set_params(dsmc_mlp m, const float *params)
{
    if (!m || !params) return fail(-1, "null argument");
    const size_t bytes = sizeof(float) * m->nparams;
    HCHK(hipMemcpyAsync(m->w, params, bytes, hipMemcpyHostToDevice, m->st));
    HCHK(hipMemsetAsync(m->m, 0, bytes, m->st));
    HCHK(hipMemsetAsync(m->v, 0, bytes, m->st));
    for (int l = 0; l < m->L; ++l) {
        const int nout = m->width[l + 1], nin = m->width[l], n = nout * nin + nout;
        hipLaunchKernelGGL(k_publish, dim3((n + 255) / 256), dim3(256), 0, m->st, m->w + m->poff[l], nout, nin,
                           m->Wp[l + 1], m->Wp[l], m->W[l], m->WT[l], m->bias[l]);
        HCHK(hipGetLastError());
        HCHK(hipMemsetAsync(m->gW[l], 0, sizeof(float) * (size_t)m->Wp[l + 1] * m->Wp[l], m->st));
        HCHK(hipMemsetAsync(m->gb[l], 0, sizeof(float) * (size_t)m->Wp[l + 1], m->st));
    }
    m->adam_t = 0;
    HCHK(hipStreamSynchronize(m->st));
    return 0;
}

int dsmc_mlp_get_params(dsmc_mlp m, float *params)
{
    if (!m || !params) return fail(-1, "null argument");
    HCHK(hipMemcpyAsync(params, m->w, sizeof(float) * m->nparams, hipMemcpyDeviceToHost, m->st));
    HCHK(hipStreamSynchronize(m->st));
    return 0;
}

int dsmc_mlp_forward(dsmc_mlp m, const float *x, int64_t n, float *y)
{
    if (!m || (n > 0 && (!x || !y)) || n < 0) return fail(-1, "bad argument");
    const int nin = m->width[0], nout = m->width[m->L], Np = m->Wp[m->L];
    float *dx = nullptr;
    HCHK(hipMalloc(&dx, sizeof(float) * (size_t)std::max<int64_t>(n, 1) * nin));
    int r = 0;
    std::vector<float> o((size_t)m->maxB * Np);
    for (int64_t s0 = 0; s0 < n && !r; s0 += m->maxB) {
        const int B = (int)std::min<int64_t>(m->maxB, n - s0);
        if (hipMemcpyAsync(dx, x + s0 * nin, sizeof(float) * (size_t)B * nin, hipMemcpyHostToDevice, m->st) != hipSuccess) {
            r = fail(-2, "upload");
            break;
        }
        if ((r = gather(m, dx, nullptr, B, B, 0, 0, false))) break;
        if ((r = forward(m, B, false))) break;
        if (hipMemcpyAsync(o.data(), m->out, sizeof(float) * (size_t)B * Np, hipMemcpyDeviceToHost, m->st) != hipSuccess ||
            hipStreamSynchronize(m->st) != hipSuccess) {
            r = fail(-2, "download");
            break;
        }
        for (int b = 0; b < B; ++b)
            for (int j = 0; j < nout; ++j) y[(s0 + b) * nout + j] = o[(size_t)b * Np + j];
    }
    (void)hipStreamSynchronize(m->st);
    (void)hipFree(dx);
    return r;
}

int dsmc_mlp_gradient(dsmc_mlp m, const float *x, const float *t, int n, float *loss, float *grad)
{
    if (!m || !x || !t || n < 8 || n > m->maxB || n % 8) return fail(-1, "bad argument (8 <= n <= max_batch, n % 8 == 0)");
    const int nin = m->width[0], nout = m->width[m->L];
    float *dx = nullptr, *dt = nullptr;
    HCHK(hipMalloc(&dx, sizeof(float) * (size_t)n * nin));
    HCHK(hipMalloc(&dt, sizeof(float) * (size_t)n * nout));
    int r = 0;
    do {
        if (hipMemcpyAsync(dx, x, sizeof(float) * (size_t)n * nin, hipMemcpyHostToDevice, m->st) != hipSuccess ||
            hipMemcpyAsync(dt, t, sizeof(float) * (size_t)n * nout, hipMemcpyHostToDevice, m->st) != hipSuccess) {
            r = fail(-2, "upload");
            break;
        }
        for (int l = 0; l < m->L && !r; ++l)
            if (hipMemsetAsync(m->gW[l], 0, sizeof(float) * (size_t)m->Wp[l + 1] * m->Wp[l], m->st) != hipSuccess ||
                hipMemsetAsync(m->gb[l], 0, sizeof(float) * (size_t)m->Wp[l + 1], m->st) != hipSuccess)
                r = fail(-2, "memset");
        if (r || (r = gather(m, dx, dt, n, n, 0, 0, false)) || (r = forward(m, n, true)) ||
            (r = loss_and_grad(m, n)) || (r = backward(m, n)))
            break;
        if (hipStreamSynchronize(m->st) != hipSuccess) { r = fail(-2, "sync"); break; }
        if (loss && hipMemcpy(loss, m->loss, sizeof(float), hipMemcpyDeviceToHost) != hipSuccess) { r = fail(-2, "loss"); break; }
        if (grad) {
            for (int l = 0; l < m->L && !r; ++l) {
                const int nout_l = m->width[l + 1], nin_l = m->width[l], Kp = m->Wp[l];
                std::vector<float> gw((size_t)m->Wp[l + 1] * Kp), gb(m->Wp[l + 1]);
                if (hipMemcpy(gw.data(), m->gW[l], sizeof(float) * gw.size(), hipMemcpyDeviceToHost) != hipSuccess ||
                    hipMemcpy(gb.data(), m->gb[l], sizeof(float) * gb.size(), hipMemcpyDeviceToHost) != hipSuccess) {
                    r = fail(-2, "gradient download");
                    break;
                }
                float *o = grad + m->poff[l];
                for (int a = 0; a < nout_l; ++a)
                    for (int b = 0; b < nin_l; ++b) o[(size_t)a * nin_l + b] = gw[(size_t)a * Kp + b];
                for (int a = 0; a < nout_l; ++a) o[(size_t)nout_l * nin_l + a] = gb[a];
            }
        }
    } while (false);
    (void)hipStreamSynchronize(m->st);
    (void)hipFree(dx);
    (void)hipFree(dt);
    return r;
}

int dsmc_mlp_adam(dsmc_mlp m, const float *grad, float lr, float beta1, float beta2, float eps)
{
    if (!m) return fail(-1, "null model");
    if (grad) {
        for (int l = 0; l < m->L; ++l) {
            const int nout = m->width[l + 1], nin = m->width[l], Kp = m->Wp[l];
            std::vector<float> gw((size_t)m->Wp[l + 1] * Kp, 0.f), gb(m->Wp[l + 1], 0.f);
            const float *gsrc = grad + m->poff[l];
            for (int a = 0; a < nout; ++a)
                for (int b = 0; b < nin; ++b) gw[(size_t)a * Kp + b] = gsrc[(size_t)a * nin + b];
            for (int a = 0; a < nout; ++a) gb[a] = gsrc[(size_t)nout * nin + a];
            HCHK(hipMemcpyAsync(m->gW[l], gw.data(), sizeof(float) * gw.size(), hipMemcpyHostToDevice, m->st));
            HCHK(hipMemcpyAsync(m->gb[l], gb.data(), sizeof(float) * gb.size(), hipMemcpyHostToDevice, m->st));
            HCHK(hipStreamSynchronize(m->st));
        }
    }
    int r = adam_step(m, lr, beta1, beta2, eps);
    if (r) return r;
    HCHK(hipStreamSynchronize(m->st));
    return 0;
}

int dsmc_mlp_set_data(dsmc_mlp m, const float *x, const float *t, int64_t n)
{
    if (!m || !x || !t || n < 1 || n >= (1LL << 32)) return fail(-1, "bad argument (1 <= n < 2^32)");
    const int nin = m->width[0], nout = m->width[m->L];
    // the previous set is dropped first; on any failure below the model is
    // left without data (dX = dT = null, nset = 0), never half-set
    if (m->st) (void)hipStreamSynchronize(m->st);
    if (m->dX) (void)hipFree(m->dX);
    if (m->dT) (void)hipFree(m->dT);
    m->dX = nullptr;
    m->dT = nullptr;
    m->nset = 0;
    float *dx = nullptr, *dt = nullptr;
    hipError_t e = hipMalloc(&dx, sizeof(float) * (size_t)n * nin);
    if (e == hipSuccess) e = hipMalloc(&dt, sizeof(float) * (size_t)n * nout);
    if (e == hipSuccess) e = hipMemcpy(dx, x, sizeof(float) * (size_t)n * nin, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(dt, t, sizeof(float) * (size_t)n * nout, hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        if (dx) (void)hipFree(dx);
        if (dt) (void)hipFree(dt);
        return fail(-2, std::string("set_data: ") + hipGetErrorString(e));
    }
    m->dX = dx;
    m->dT = dt;
    m->nset = n;
    return 0;
}

int dsmc_mlp_train(dsmc_mlp m, int steps, int batch, float lr, float beta1, float beta2, float eps, uint64_t seed,
                   float *loss_out)
{
    if (!m || steps < 0 || batch < 8 || batch > m->maxB || batch % 8)
        return fail(-1, "bad argument (8 <= batch <= max_batch, batch % 8 == 0)");
    if (!m->dX) return fail(-1, "no training data (dsmc_mlp_set_data)");
    std::vector<float> losses(std::max(steps, 1));
    HCHK(hipEventRecord(m->ev[0], m->st));
    for (int s = 0; s < steps; ++s) {
        int r;
        if ((r = gather(m, m->dX, m->dT, m->nset, batch, seed, (uint32_t)s, true)) || (r = forward(m, batch, true)) ||
            (r = loss_and_grad(m, batch)) || (r = backward(m, batch)) || (r = adam_step(m, lr, beta1, beta2, eps)))
            return r;
        if (loss_out) HCHK(hipMemcpyAsync(&losses[s], m->loss, sizeof(float), hipMemcpyDeviceToHost, m->st));
    }
    HCHK(hipEventRecord(m->ev[1], m->st));
    HCHK(hipEventSynchronize(m->ev[1]));
    float ms = 0.f;
    HCHK(hipEventElapsedTime(&ms, m->ev[0], m->ev[1]));
    m->ms_step = steps ? ms / steps : 0.0;
    m->flop_step = step_flop(m, batch);
    if (loss_out) std::memcpy(loss_out, losses.data(), sizeof(float) * steps);
    return 0;
}

int dsmc_mlp_train_stats(dsmc_mlp m, double *ms_per_step, double *flop_per_step)
{
    if (!m) return fail(-1, "null model");
    if (ms_per_step) *ms_per_step = m->ms_step;
    if (flop_per_step) *flop_per_step = m->flop_step;
    return 0;
}

const char *dsmc_mlp_last_error(void) { return g_err.c_str(); }

}  // extern "C"
