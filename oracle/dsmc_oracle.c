/*
 * dsmc_oracle.c -- CPU restatement of the reference DSMC hot path.
 *
 * TEST INFRASTRUCTURE ONLY (see dsmc_oracle.h): the checker for the HIP
 * kernels and the timed CPU baseline.  Never linked into the product.
 *
 * Build: oracle/Makefile (gcc -O3 -ffp-contract=off -fno-math-errno -fopenmp).
 * FP contraction must stay off: the GPU side rounds every product and sum
 * separately and bit-exact comparison depends on it.
 */
#include "dsmc_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------------- */
/* Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as   */
/* easy as 1, 2, 3", SC'11).  Replaces the reference's clock-seeded cuRAND   */
/* XORWOW streams (GPUParticle.cu:242,389; GPUCell.cu:219 -- defect D10).    */
/* ------------------------------------------------------------------------- */
void or_philox(const uint32_t ctr[4], uint32_t k0, uint32_t k1, uint32_t out[4])
{
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    for (int r = 0; r < 10; ++r) {
        if (r) { k0 += 0x9E3779B9u; k1 += 0xBB67AE85u; }
        uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
        uint32_t n1 = (uint32_t)p1;
        uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
        uint32_t n3 = (uint32_t)p0;
        c0 = n0; c1 = n1; c2 = n2; c3 = n3;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

/* uniform in [0,1) and (0,1] with 24 random bits (exact in FP32) */
static inline float u01(uint32_t w)  { return (float)(w >> 8) * 0x1p-24f; }
static inline float u01o(uint32_t w) { return (float)((w >> 8) + 1u) * 0x1p-24f; }
/* floor(w * n / 2^32): unbiased-enough integer in [0,n) without FP */
static inline int mulhi_idx(uint32_t w, int n) { return (int)(((uint64_t)w * (uint64_t)(uint32_t)n) >> 32); }
/* index of the r-th set bit of m (r < popcount(m)) */
static inline int nth_set_bit(uint64_t m, int r)
{
    for (int b = 0; b < 64; ++b)
        if ((m >> b) & 1u) {
            if (r == 0) return b;
            --r;
        }
    return 63;
}

/* ------------------------------------------------------------------------- */
/* Elementary functions from IEEE basic operations only.                      */
/* ------------------------------------------------------------------------- */
#define OR_LN2_HI  0x1.62e400p-1f      /* 0.693145751953125, 16 significant bits */
#define OR_LN2_LO  0x1.7f7d1cp-20f     /* ln2 - OR_LN2_HI */
#define OR_LOG2E   0x1.715476p+0f
#define OR_PIO2    0x1.921fb6p+0f

/* natural log for finite normal x > 0: x = 2^e m, m in [sqrt(1/2), sqrt(2)),
 * log m = 2 atanh(s), s = (m-1)/(m+1), |s| <= 0.1716, odd series to s^9. */
float or_logf(float x)
{
    uint32_t ix;
    memcpy(&ix, &x, 4);
    int e = (int)((ix >> 23) & 0xffu) - 127;
    uint32_t mb = (ix & 0x007fffffu) | 0x3f800000u;
    float m;
    memcpy(&m, &mb, 4);
    if (m > 0x1.6a09e6p+0f) { m = m * 0.5f; e = e + 1; }
    float f = m - 1.0f;
    float s = f / (2.0f + f);
    float s2 = s * s;
    float p = s2 * (0x1.555556p-2f + s2 * (0x1.99999ap-3f + s2 * (0x1.24924ap-3f + s2 * 0x1.c71c72p-4f)));
    float two_s = 2.0f * s;
    float logm = two_s + two_s * p;
    float ef = (float)e;
    return ef * OR_LN2_HI + (ef * OR_LN2_LO + logm);
}

/* exp for x in [-87, 88]: x = k ln2 + r (Cody-Waite), Taylor to r^7. */
float or_expf(float x)
{
    if (x < -87.0f) return 0.0f;
    if (x > 88.0f) x = 88.0f;
    float kf = floorf(x * OR_LOG2E + 0.5f);
    float r = (x - kf * OR_LN2_HI) - kf * OR_LN2_LO;
    float p = 1.0f + r * (1.0f + r * (0.5f + r * (0x1.555556p-3f + r * (0x1.555556p-5f +
              r * (0x1.111112p-7f + r * (0x1.6c16c2p-10f + r * 0x1.a01a02p-13f))))));
    int k = (int)kf;
    uint32_t sb;
    float scale;
    if (k > 127) {
        sb = (uint32_t)(127 + 127) << 23; memcpy(&scale, &sb, 4);
        return p * scale * 2.0f;
    }
    if (k < -126) {
        sb = (uint32_t)(k + 126 + 127) << 23; memcpy(&scale, &sb, 4);
        return p * scale * 0x1p-126f;
    }
    sb = (uint32_t)(k + 127) << 23;
    memcpy(&scale, &sb, 4);
    return p * scale;
}

/* sin and cos of 2*pi*u for u in [0,1): quadrant reduction is exact
 * (4u and its nearest integer are exact), |theta| <= pi/4, Taylor series. */
void or_sincos2pi(float u, float *sn, float *cs)
{
    float t = u * 4.0f;
    float q = floorf(t + 0.5f);
    float f = t - q;
    float th = f * OR_PIO2;
    float t2 = th * th;
    float sp = th + th * t2 * (-0x1.555556p-3f + t2 * (0x1.111112p-7f + t2 * (-0x1.a01a02p-13f + t2 * 0x1.71de3ap-19f)));
    float cp = 1.0f + t2 * (-0.5f + t2 * (0x1.555556p-5f + t2 * (-0x1.6c16c2p-10f + t2 * (0x1.a01a02p-16f + t2 * -0x1.27e4fcp-22f))));
    int qi = ((int)q) & 3;
    if (qi == 0)      { *sn = sp;  *cs = cp; }
    else if (qi == 1) { *sn = cp;  *cs = -sp; }
    else if (qi == 2) { *sn = -sp; *cs = -cp; }
    else              { *sn = -cp; *cs = sp; }
}

/* ------------------------------------------------------------------------- */
/* Binning index.  Reference: GPUCell.cu:184-189 and CartesianMesh.cpp:171-181 */
/* (i = (int)(x/dx), cell = k + j*nz + i*ny*nz with nz = 1).  Deviation D14:  */
/* out-of-range indices are clamped instead of leaving stale ids.            */
/* ------------------------------------------------------------------------- */
static inline int cell_of(const or_case *cs, float x, float y)
{
    int i = (int)(x / cs->dx);
    int j = (int)(y / cs->dy);
    if (i < 0) i = 0;
    if (i > cs->nx - 1) i = cs->nx - 1;
    if (j < 0) j = 0;
    if (j > cs->ny - 1) j = cs->ny - 1;
    return j + i * cs->ny;
}

/* Rotational energy per unit mass of a molecule with two rotational degrees
 * of freedom in equilibrium at kT/m = ekt: exponential with mean ekt, as
 * Random::RotationalenergySample (Random.cpp:48-51) draws -ln(U) k T. */
static inline float rot_energy(uint32_t w, float ekt) { return -or_logf(u01o(w)) * ekt; }

/* Diffuse re-emission from a wall with unit normal n (into the fluid).
 * Reference: GPUSegment.cuh:59-84 and WallBoundary.cpp:33-43.  D7 fix:
 * independent draws -- normal speed from the flux-weighted (Rayleigh) law
 * v_n = sqrt(-ln U) sqrt(2kTw/m), tangential and z components N(0, kTw/m).
 * With er != NULL the rotational energy is accommodated to the wall too
 * (drawn at kTw/m = ekt); the reference has no rotational wall model. */
static void diffuse_emit(float nx, float ny, float vmp, float sig, float ekt, uint32_t pid,
                         uint32_t seed, uint32_t step, uint32_t b,
                         float *vx, float *vy, float *vz, float *er)
{
    uint32_t ctr[4] = { b, step, OR_SUB_MOVE, 0u }, w[4];
    or_philox(ctr, pid, seed, w);
    float vn = sqrtf(-or_logf(u01o(w[0]))) * vmp;
    float r = sqrtf(-2.0f * or_logf(u01o(w[1]))) * sig;
    float s, c;
    or_sincos2pi(u01(w[2]), &s, &c);
    float vt = r * c;
    *vx = vn * nx + vt * (-ny);
    *vy = vn * ny + vt * nx;
    *vz = r * s;
    if (er) *er = rot_energy(w[3], ekt);
}

/* One particle's flight of duration t.  Reference: moveParticles
 * (GPUParticle.cu:244-350): x += v dt in 2D, cut-cell wall reflection, and
 * deletion outside [0,L1]x[0,L2].  Deviations: D6 (the hit point is found on
 * the actual path, not extrapolated forward), D8 (finite ray-segment test
 * against every body segment the path can reach, earliest hit first, repeated
 * for further hits), D17 (per-face boundary types: stream faces delete,
 * specular faces mirror as WallBoundary::apply in Boundary.cuh:40-62, diffuse
 * faces re-emit). */
static void move_one(const or_case *cs, float *px, float *py, float *pvx, float *pvy,
                     float *pvz, float *per, uint32_t pid, uint32_t step, float t,
                     uint8_t *dead, uint32_t *err)
{
    float x = *px, y = *py, vx = *pvx, vy = *pvy, vz = *pvz;
    uint8_t dd = 0;
    for (int b = 0;; ++b) {
        if (b == OR_MAX_BOUNCE) { *err |= OR_ERR_BOUNCE; break; }
        float best = t;
        int kind = -1;
        if (vx < 0.0f) {
            float s = (0.0f - x) / vx;
            if (s < best) { best = s; kind = 0; }
        } else if (vx > 0.0f) {
            float s = (cs->L1 - x) / vx;
            if (s < best) { best = s; kind = 1; }
        }
        if (vy < 0.0f) {
            float s = (0.0f - y) / vy;
            if (s < best) { best = s; kind = 2; }
        } else if (vy > 0.0f) {
            float s = (cs->L2 - y) / vy;
            if (s < best) { best = s; kind = 3; }
        }
        if (cs->nseg > 0) {
            float reach = sqrtf(vx * vx + vy * vy) * t;
            int all = reach > cs->margin;
            int cnt, c0 = 0;
            if (all) {
                cnt = cs->nseg;
            } else {
                int c = cell_of(cs, x, y);
                c0 = cs->cseg_start[c];
                cnt = cs->cseg_start[c + 1] - c0;
            }
            for (int q = 0; q < cnt; ++q) {
                int k = all ? q : cs->cseg_idx[c0 + q];
                const float *S = cs->seg + 6 * k;
                float vn = vx * S[4] + vy * S[5];
                if (!(vn < 0.0f)) continue;
                float ex = S[2] - S[0], ey = S[3] - S[1];
                float den = vx * ey - vy * ex;
                if (den == 0.0f) continue;
                float wx = S[0] - x, wy = S[1] - y;
                float s = (wx * ey - wy * ex) / den;
                if (!(s < best)) continue;
                if (!(s > 0.0f)) {
                    /* on or just behind the wall line and moving in: hit now --
                     * only at the start of a flight (a previous step's hit point
                     * rounded to the solid side).  After a reflection the particle
                     * leaves the face it hit, which for a convex body is leaving
                     * the body; a band hit on the neighbouring face of an acute
                     * vertex would trap it there. */
                    if (b > 0) continue;
                    float d = -(wx * S[4] + wy * S[5]);
                    if (d < -(cs->margin * OR_SEG_BAND)) continue;
                    s = 0.0f;
                }
                float u = (wx * vy - wy * vx) / den;
                if (u < -OR_SEG_TOL || u > 1.0f + OR_SEG_TOL) continue;
                /* within the end overlap, past the segment's own end: a hit
                 * only for a flight that also enters the solid side of the
                 * segment sharing that vertex (the body is one closed polygon,
                 * segment k from vertex k to k+1) -- only such a flight could
                 * slip into the body between the two rounded tests.  Without
                 * it, the overlaps of the two faces at an acute vertex bound a
                 * thin region outside the body in which a re-emitted particle
                 * can bounce from one overlap to the other without end. */
                if (u < 0.0f || u > 1.0f) {
                    int nb = u > 1.0f ? (k + 1 == cs->nseg ? 0 : k + 1) : (k == 0 ? cs->nseg - 1 : k - 1);
                    const float *Q = cs->seg + 6 * nb;
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
        if (kind < 4) {
            if (kind == 0) x = 0.0f;
            else if (kind == 1) x = cs->L1;
            else if (kind == 2) y = 0.0f;
            else y = cs->L2;
            int bc = cs->face_bc[kind];
            if (bc == OR_BC_STREAM) { dd = 1; break; }
            if (bc == OR_BC_SPECULAR) {
                if (kind < 2) vx = -vx; else vy = -vy;
            } else {
                float nx = kind == 0 ? 1.0f : (kind == 1 ? -1.0f : 0.0f);
                float ny = kind == 2 ? 1.0f : (kind == 3 ? -1.0f : 0.0f);
                diffuse_emit(nx, ny, cs->face_vmp[kind], cs->face_sig[kind], cs->face_ekt[kind], pid,
                             cs->seed, step, (uint32_t)b, &vx, &vy, &vz, per);
            }
        } else {
            const float *S = cs->seg + 6 * (kind - 4);
            if (cs->body_bc == OR_BC_SPECULAR) {
                /* specular law v' = v + 2 |v.n| n = v - 2 (v.n) n for v.n < 0
                 * (Segment.cpp:40, WallBoundary.cpp:46) */
                float vn = vx * S[4] + vy * S[5];
                vx = vx - 2.0f * vn * S[4];
                vy = vy - 2.0f * vn * S[5];
            } else {
                diffuse_emit(S[4], S[5], cs->body_vmp, cs->body_sig, cs->body_ekt, pid, cs->seed,
                             step, (uint32_t)b, &vx, &vy, &vz, per);
            }
        }
    }
    *px = x; *py = y; *pvx = vx; *pvy = vy; *pvz = vz;
    *dead = dd;
}

void or_move(const or_case *cs, int n, float *x, float *y, float *vx, float *vy,
             float *vz, float *er, const uint32_t *pid, uint32_t step, uint8_t *dead,
             uint32_t *err)
{
    uint32_t e = 0;
    for (int i = 0; i < n; ++i)
        move_one(cs, &x[i], &y[i], &vx[i], &vy[i], &vz[i], er ? &er[i] : NULL, pid[i], step,
                 cs->dt, &dead[i], &e);
    *err |= e;
}

/* ------------------------------------------------------------------------- */
/* Inflow.  Reference: Injet/InjectParticles (GPUParticle.cu:149-171,375-422) */
/* fills a slab of width V_jet*tau on x=0 with drifting Maxwellians (D18).    */
/* Restated as the Maxwellian flux through each stream face (Bird 1994 eq.    */
/* 4.22), number per step with exact remainder carry floor((s+1)e)-floor(s e)*/
/* and positions advanced by a random fraction of dt along their path.       */
/* ------------------------------------------------------------------------- */
void or_inject_counts(const or_case *cs, uint32_t step, int J[4])
{
    for (int f = 0; f < 4; ++f) {
        double e = cs->inj_e[f];
        if (cs->face_bc[f] == OR_BC_STREAM && e > 0.0)
            J[f] = (int)(floor(((double)step + 1.0) * e) - floor((double)step * e));
        else
            J[f] = 0;
    }
}

static void inject_one(const or_case *cs, int f, uint32_t pid, uint32_t step,
                       float *ox, float *oy, float *ovx, float *ovy, float *ovz, float *oer,
                       uint8_t *dead, uint32_t *err)
{
    uint32_t ctr[4] = { 0u, step, OR_SUB_INJECT, 0u }, w[4];
    or_philox(ctr, pid, cs->seed, w);
    float upos = u01(w[0]);
    float ufrac = u01(w[1]);
    float r = sqrtf(-2.0f * or_logf(u01o(w[2])));
    float sn, cn;
    or_sincos2pi(u01(w[3]), &sn, &cn);
    float g0 = r * cn, g1 = r * sn;

    /* normal speed ratio w = v_n/vmp with density ~ w exp(-(w-a)^2), w > 0:
     * acceptance-rejection against its maximum at w* = (a + sqrt(a^2+2))/2 */
    float a = cs->inj_a[f];
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
    int ok = 0;
    for (int tr = 0; tr < OR_INFLOW_TRIES && !ok; tr += 2) {
        uint32_t c2[4] = { 1u + (uint32_t)(tr / 2), step, OR_SUB_INJECT, 0u };
        or_philox(c2, pid, cs->seed, w);
        for (int h = 0; h < 2; ++h) {
            float cand = lo + (hi - lo) * u01(w[2 * h]);
            float d = cand - a;
            float acc = (cand / ws) * or_expf(ref - d * d);
            wn = cand;
            if (u01(w[2 * h + 1]) < acc) { ok = 1; break; }
        }
    }
    if (!ok) *err |= OR_ERR_INFLOW;
    float vn = wn * cs->inj_vmp;
    float x, y, vx, vy;
    float vz = cs->inj_sig * g1;
    if (f == 0)      { x = 0.0f;          y = upos * cs->L2; vx = vn;  vy = cs->inj_uy + cs->inj_sig * g0; }
    else if (f == 1) { x = cs->L1;        y = upos * cs->L2; vx = -vn; vy = cs->inj_uy + cs->inj_sig * g0; }
    else if (f == 2) { x = upos * cs->L1; y = 0.0f;          vy = vn;  vx = cs->inj_ux + cs->inj_sig * g0; }
    else             { x = upos * cs->L1; y = cs->L2;        vy = -vn; vx = cs->inj_ux + cs->inj_sig * g0; }
    /* rotational energy of the freestream (the draw GPUParticle.cu:409-420
     * makes and leaves unstored, D13) */
    float er = 0.0f;
    if (oer) {
        uint32_t c3[4] = { 0u, step, OR_SUB_INJECT, 1u };
        or_philox(c3, pid, cs->seed, w);
        er = rot_energy(w[0], cs->inj_ekt);
    }
    move_one(cs, &x, &y, &vx, &vy, &vz, oer ? &er : NULL, pid, step, ufrac * cs->dt, dead, err);
    *ox = x; *oy = y; *ovx = vx; *ovy = vy; *ovz = vz;
    if (oer) *oer = er;
}

void or_inject(const or_case *cs, uint32_t step, uint32_t next_pid, int J_total,
               float *x, float *y, float *vx, float *vy, float *vz, float *er,
               uint32_t *pid, uint8_t *dead, uint32_t *err)
{
    int J[4];
    or_inject_counts(cs, step, J);
    uint32_t e = 0;
    for (int q = 0; q < J_total; ++q) {
        int f = 0, acc = J[0];
        while (f < 3 && q >= acc) { ++f; acc += J[f]; }
        uint32_t p = next_pid + (uint32_t)q;
        pid[q] = p;
        inject_one(cs, f, p, step, &x[q], &y[q], &vx[q], &vy[q], &vz[q], er ? &er[q] : NULL,
                   &dead[q], &e);
    }
    *err |= e;
}

/* ------------------------------------------------------------------------- */
/* Bin / scan / scatter.  Reference: CalparticleNumAndLocalID                */
/* (GPUCell.cu:169-198), thrust::exclusive_scan (GPUCell.cu:104-120),        */
/* sortParticles (GPUParticle.cu:365-373).  The serial loop makes the        */
/* atomicAdd rank equal to index order, i.e. the scatter is stable.          */
/* ------------------------------------------------------------------------- */
void or_bin(const or_case *cs, int n, const float *x, const float *y,
            const uint8_t *dead, int *cell, int *count, int *slot)
{
    for (int i = 0; i < n; ++i) {
        if (dead && dead[i]) { cell[i] = -1; slot[i] = 0; continue; }
        int c = cell_of(cs, x[i], y[i]);
        cell[i] = c;
        slot[i] = count[c]++;
    }
}

void or_scan(const int *count, int ncell, int *start)
{
    int s = 0;
    for (int c = 0; c < ncell; ++c) { start[c] = s; s += count[c]; }
    start[ncell] = s;
}

void or_scatter(int n, const int *cell, const int *slot, const int *start, int *perm)
{
    for (int i = 0; i < n; ++i)
        if (cell[i] >= 0) perm[start[cell[i]] + slot[i]] = i;
}

/* ------------------------------------------------------------------------- */
/* NTC/VHS collisions of one cell.  Reference: collisionInCells              */
/* (GPUCell.cu:200-333) with the host semantics of Cell::collision           */
/* (Cell.cpp:335-409) and Particle::Collision (Particle.cpp:82-101).         */
/*  - D2/D25: remainder carry r' = (expected + r) - M  (the GPU kernel has   */
/*    none; the host version carries expected - M, which loses candidates);  */
/*  - D3: (sigma c_r)max persists per cell;                                  */
/*  - D4: distinct pair drawn without the off-by-one (i2 skips i1);          */
/*  - D5: Gamma(5/2 - omega) from tgamma, folded into K1 by the host;        */
/*  - candidate count uses the pair number np(np-1) (D24; np^2 over-counts   */
/*    collisions by (1 + 1/<np>)), T_ref in the VHS cross-section.          */
/* Rotational exchange (zeta = 2): the Larsen-Borgnakke scheme the reference */
/* writes out and leaves commented (GPUCell.cu:256-275,316-328, D1), for the */
/* pair at once -- the collision energy E_c = m_r g^2/2 + E_r1 + E_r2 is     */
/* split into translation, fraction x ~ Beta(5/2 - omega, zeta), and         */
/* rotation, shared between the two by Beta(zeta/2, zeta/2) (uniform for    */
/* zeta = 2).  Deviations: the translational energy uses the reduced mass    */
/* m_r = m/2 (the reference takes m, which breaks energy conservation), a    */
/* fraction rot_p = 1/Z_rot of the collisions exchanges (1 = every one, the  */
/* reference), and x = U1^(1/(5/2-omega)) U2^(1/(7/2-omega)) -- the product  */
/* of Beta(a,1) and Beta(a+1,1), which is Beta(a,2) -- in place of the       */
/* acceptance-rejection loop: the same law with a fixed number of draws.     */
/* Energies are per unit mass: E_c/m = g^2/4 + e1 + e2, g' = 2 sqrt(x E_c/m).*/
/* ------------------------------------------------------------------------- */
/* The candidate loop of one collision cell (or leaf): candidates m0 .. m0+M-1
 * of cell c's Philox stream, pairs among perm[s0 .. s0+np), running
 * (sigma c_r)max in *smax.  Returns the number of collisions. */
static int ntc_candidates(const or_case *cs, int c, uint32_t step, int s0, int np, int m0, int M,
                          const int *perm, float *vx, float *vy, float *vz, float *er, float *smax_io)
{
    float smax = *smax_io;
    int nc = 0;
    for (int m = m0; m < m0 + M; ++m) {
        uint32_t ctr[4] = { (uint32_t)m, step, OR_SUB_COLLIDE, 0u }, w[4];
        or_philox(ctr, (uint32_t)c, cs->seed, w);
        int i1 = mulhi_idx(w[0], np);
        int i2 = mulhi_idx(w[1], np - 1);
        if (i2 >= i1) i2 = i2 + 1;
        int p1 = perm[s0 + i1], p2 = perm[s0 + i2];
        float gx = vx[p1] - vx[p2], gy = vy[p1] - vy[p2], gz = vz[p1] - vz[p2];
        float g2 = gx * gx + gy * gy + gz * gz;
        float sc = g2 > 0.0f ? cs->K1 * or_expf(cs->e_exp * or_logf(g2)) : 0.0f;
        if (sc > smax) smax = sc;
        if (u01(w[2]) < sc / smax) {
            float cosr = 2.0f * u01(w[3]) - 1.0f;
            float sinr = sqrtf(1.0f - cosr * cosr);
            uint32_t c2[4] = { (uint32_t)m, step, OR_SUB_COLLIDE, 1u }, w2[4];
            or_philox(c2, (uint32_t)c, cs->seed, w2);
            float sp, cp;
            or_sincos2pi(u01(w2[0]), &sp, &cp);
            float g = sqrtf(g2);
            if (er) {
                uint32_t c3[4] = { (uint32_t)m, step, OR_SUB_COLLIDE, 2u }, w3[4];
                or_philox(c3, (uint32_t)c, cs->seed, w3);
                if (u01(w3[0]) < cs->rot_p) {
                    float ec = 0.25f * g2 + (er[p1] + er[p2]);
                    float x = or_expf(cs->lb_ia * or_logf(u01o(w3[1])) + cs->lb_ib * or_logf(u01o(w3[2])));
                    float et = x * ec;
                    float erot = ec - et;
                    float e1 = erot * u01(w3[3]);
                    er[p1] = e1;
                    er[p2] = erot - e1;
                    g = 2.0f * sqrtf(et);
                }
            }
            float rx = g * cosr, ry = g * (sinr * cp), rz = g * (sinr * sp);
            float mx = 0.5f * (vx[p1] + vx[p2]);
            float my = 0.5f * (vy[p1] + vy[p2]);
            float mz = 0.5f * (vz[p1] + vz[p2]);
            vx[p1] = mx + 0.5f * rx; vx[p2] = mx - 0.5f * rx;
            vy[p1] = my + 0.5f * ry; vy[p2] = my - 0.5f * ry;
            vz[p1] = mz + 0.5f * rz; vz[p2] = mz - 0.5f * rz;
            ++nc;
        }
    }
    *smax_io = smax;
    return nc;
}

/* NTC candidate number of np particles in volume vol with remainder carry. */
static int ntc_count(const or_case *cs, int np, float vol, float smax, float *mrem)
{
    float e = (float)((long long)np * (long long)(np - 1)) * cs->ntc_coef;
    e = e * smax;
    e = e / vol;
    float tot = e + *mrem;
    int M = (int)tot;
    *mrem = tot - (float)M;
    return M;
}

static int collide_cell(const or_case *cs, int c, uint32_t step, int s0, int np,
                        const int *perm, float *vx, float *vy, float *vz, float *er,
                        float *sigcmax, float *mrem)
{
    if (np < 2) return 0;
    float vol = cs->vol_ntc[c];
    if (!(vol > 0.0f)) return 0;
    float smax = sigcmax[c];
    int M = ntc_count(cs, np, vol, smax, &mrem[c]);
    int nc = ntc_candidates(cs, c, step, s0, np, 0, M, perm, vx, vy, vz, er, &smax);
    sigcmax[c] = smax;
    return nc;
}

/* ------------------------------------------------------------------------- */
/* Adaptive collision sub-cells and the variable time step.  Reference:      */
/* Cell::VTS (Cell.cpp:117-136), Cell::genAMRmesh (Cell.cpp:168-279),        */
/* Element::genAMRmesh (Element.cpp:158-182), Cell::sortParticle2children    */
/* (Cell.cpp:280-295), Cell::collision over the leaves (Cell.cpp:335-352)    */
/* and the Numchild/Numgrandchild datasets (Output.cpp:45-63).  The          */
/* reference defines these on the host and never calls them.  Restated:      */
/*  - mean free path lam = m / (sqrt(2) rho pi d^2) (T/T_inf)^(omega-1/2)   */
/*    (Cell.cpp:125) from the cell's time-averaged sums and its fluid        */
/*    volume (the reference samples each sub-cell instantaneously, dividing  */
/*    by the parent's volume); the power as exp/log in FP32 with this        */
/*    build's own routines, everything else in FP64;                         */
/*  - VTS: dt = lam / (2 (|u| + sqrt(2kT/m))) (Cell.cpp:128-129);            */
/*  - level 2: side Ch lam_max, Ch = 1/2, n = max(1, ceil(L1/L2)), side      */
/*    L1/n (Cell.cpp:174-197); level 3 in every level-2 cell: side           */
/*    Ch lam_min with lam_min the smallest old level-3 mean free path, or    */
/*    2 x (level-2 x side) when the cell had no grid (Cell.cpp:217-243).     */
/*    isIntersecting is always true (Element.cpp), so every level-2 cell of  */
/*    a cell gets the same level-3 grid and the leaves form one uniform      */
/*    nsx x nsy grid, nsx = n2x n3x; with one mean free path per cell,       */
/*    lam_max = lam_min = lam;                                               */
/*  - leaves: a cell's particles ordered stably by leaf, then NTC in every   */
/*    leaf with the leaf volume (cell fluid volume / leaves), a running      */
/*    (sigma c_r)max per leaf from the cell's value (the cell keeps the      */
/*    largest), the candidate remainder carried from leaf to leaf in leaf    */
/*    order, and the candidates numbered on through the leaves (one Philox   */
/*    stream per cell and step as without sub-cells);                        */
/*  - at most OR_AMR_MAX sub-cells per level and axis; a cell with more      */
/*    than OR_AMR_CAP particles or more than OR_AMR_LEAVES leaves collides   */
/*    without sub-cells.                                                     */
/* ------------------------------------------------------------------------- */
int or_amr_leaves(uint32_t sub, int *nsx, int *nsy)
{
    if (!sub) { *nsx = 1; *nsy = 1; return 1; }
    *nsx = (int)(sub & 0xffu) * (int)((sub >> 16) & 0xffu);
    *nsy = (int)((sub >> 8) & 0xffu) * (int)((sub >> 24) & 0xffu);
    return *nsx * *nsy;
}

static int amr_div(double L, double side)
{
    double q = L / side;
    if (!(q > 1.0)) return 1;
    double n = ceil(q);
    return n > (double)OR_AMR_MAX ? OR_AMR_MAX : (int)n;
}

uint32_t or_amr_grid(double dx, double dy, double lam_max, double lam_min, int had)
{
    int n2x = amr_div(dx, 0.5 * lam_max), n2y = amr_div(dy, 0.5 * lam_max);
    double l2x = dx / (double)n2x, l2y = dy / (double)n2y;
    double mn = had ? lam_min : 2.0 * l2x;
    int n3x = amr_div(l2x, 0.5 * mn), n3y = amr_div(l2y, 0.5 * mn);
    return (uint32_t)n2x | ((uint32_t)n2y << 8) | ((uint32_t)n3x << 16) | ((uint32_t)n3y << 24);
}

void or_amr_update(const or_case *cs, const or_amr_gas *g, const double *acc, long long nsamples,
                   const double *vol, uint32_t *sub, float *lam, double *dt)
{
    int C = cs->ncell;
    const double pi_d2 = sqrt(2.0) * 3.14159265358979323846 * g->diam * g->diam;
    for (int c = 0; c < C; ++c) {
        double S0 = acc[c];
        lam[c] = 0.0f;
        dt[c] = 0.0;
        if (!(S0 > 0.0) || !(vol[c] > 0.0) || nsamples <= 0) { sub[c] = 0u; continue; }
        double n = (S0 / (double)nsamples) * g->Fn / vol[c];
        double ux = acc[C + c] / S0, uy = acc[2 * C + c] / S0, uz = acc[3 * C + c] / S0;
        double uu = ux * ux + uy * uy + uz * uz;
        double T = g->mass / (3.0 * g->boltz) * (acc[4 * C + c] / S0 - uu);
        float tr = (float)(T / g->T_inf);
        if (!(tr > 0.0f)) { sub[c] = 0u; continue; }
        float f = or_expf((float)(g->omega - 0.5) * or_logf(tr));
        double l = g->mass / (pi_d2 * (n * g->mass)) * (double)f;
        lam[c] = (float)l;
        dt[c] = 0.5 * l / (sqrt(uu) + sqrt(2.0 * g->boltz * T / g->mass));
        sub[c] = or_amr_grid((double)cs->dx, (double)cs->dy, l, l, sub[c] != 0u);
    }
}

int or_amr_leaf(const or_case *cs, int c, float x, float y, int nsx, int nsy)
{
    int i = c / cs->ny, j = c % cs->ny;
    float fx = x / cs->dx - (float)i, fy = y / cs->dy - (float)j;
    int lx = (int)(fx * (float)nsx), ly = (int)(fy * (float)nsy);
    lx = lx < 0 ? 0 : (lx > nsx - 1 ? nsx - 1 : lx);
    ly = ly < 0 ? 0 : (ly > nsy - 1 ? nsy - 1 : ly);
    return ly + lx * nsy;
}

/* One cell with nsx x nsy leaves: perm[s0 .. s0+np) ordered stably by leaf,
 * then NTC leaf by leaf. */
static int collide_cell_leaves(const or_case *cs, int c, uint32_t step, int s0, int np, int nsx, int nsy,
                               int *perm, const float *x, const float *y, float *vx, float *vy,
                               float *vz, float *er, float *sigcmax, float *mrem)
{
    int nl = nsx * nsy;
    int *lf = (int *)malloc(sizeof(int) * (size_t)np);
    int *tmp = (int *)malloc(sizeof(int) * (size_t)np);
    int *ls = (int *)calloc((size_t)nl + 1, sizeof(int));
    for (int k = 0; k < np; ++k) {
        int p = perm[s0 + k];
        lf[k] = or_amr_leaf(cs, c, x[p], y[p], nsx, nsy);
        ls[lf[k] + 1] += 1;
    }
    for (int l = 0; l < nl; ++l) ls[l + 1] += ls[l];
    int *pos = (int *)malloc(sizeof(int) * (size_t)nl);
    for (int l = 0; l < nl; ++l) pos[l] = ls[l];
    for (int k = 0; k < np; ++k) tmp[pos[lf[k]]++] = perm[s0 + k];
    memcpy(perm + s0, tmp, sizeof(int) * (size_t)np);
    float vol = cs->vol_ntc[c];
    float vleaf = vol / (float)nl;
    float smax0 = sigcmax[c], best = smax0;
    int nc = 0, m0 = 0;
    if (vol > 0.0f)
        for (int l = 0; l < nl; ++l) {
            int npl = ls[l + 1] - ls[l];
            if (npl < 2) continue;
            int M = ntc_count(cs, npl, vleaf, smax0, &mrem[c]);
            float sm = smax0;
            nc += ntc_candidates(cs, c, step, s0 + ls[l], npl, m0, M, perm, vx, vy, vz, er, &sm);
            m0 += M;
            if (sm > best) best = sm;
        }
    sigcmax[c] = best;
    free(lf); free(tmp); free(ls); free(pos);
    return nc;
}

void or_collide(const or_case *cs, uint32_t step, const int *start, int *perm,
                const float *x, const float *y,
                float *vx, float *vy, float *vz, float *er, float *sigcmax, float *mrem,
                int *ncoll_cell, int nthreads)
{
    int C = cs->ncell;
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 256) num_threads(nthreads > 1 ? nthreads : 1) if (nthreads > 1)
#endif
    for (int c = 0; c < C; ++c) {
        int np = start[c + 1] - start[c], nsx = 1, nsy = 1;
        int nl = cs->sub ? or_amr_leaves(cs->sub[c], &nsx, &nsy) : 1;
        if (nl > 1 && nl <= OR_AMR_LEAVES && np >= 2 && np <= OR_AMR_CAP)
            ncoll_cell[c] = collide_cell_leaves(cs, c, step, start[c], np, nsx, nsy, perm, x, y,
                                                vx, vy, vz, er, sigcmax, mrem);
        else
            ncoll_cell[c] = collide_cell(cs, c, step, start[c], np, perm, vx, vy, vz, er, sigcmax, mrem);
    }
}

/* ------------------------------------------------------------------------- */
/* Sampling.  Reference: samplingInCells (GPUCell.cu:336-402) and            */
/* Cell::sample (Cell.cpp:410-447), restated as time-averaged FP64           */
/* accumulators (D15): counts, sum v, sum v^2, collisions per cell.          */
/* ------------------------------------------------------------------------- */
void or_sample(int ncell, const int *start, const int *perm, const float *vx,
               const float *vy, const float *vz, const float *er, const int *ncoll_cell,
               double *acc)
{
    for (int c = 0; c < ncell; ++c) {
        int s0 = start[c], np = start[c + 1] - s0;
        double sx = 0.0, sy = 0.0, sz = 0.0, s2 = 0.0, se = 0.0;
        for (int k = 0; k < np; ++k) {
            int p = perm[s0 + k];
            double a = (double)vx[p], b = (double)vy[p], d = (double)vz[p];
            sx += a; sy += b; sz += d;
            s2 += a * a + b * b + d * d;
            if (er) se += (double)er[p];
        }
        if (er) acc[6 * ncell + c] += se;
        acc[0 * ncell + c] += (double)np;
        acc[1 * ncell + c] += sx;
        acc[2 * ncell + c] += sy;
        acc[3 * ncell + c] += sz;
        acc[4 * ncell + c] += s2;
        acc[5 * ncell + c] += (double)(ncoll_cell ? ncoll_cell[c] : 0);
    }
}

void or_reorder(int n, const int *perm,
                const float *x, const float *y, const float *vx, const float *vy,
                const float *vz, const float *er, const uint32_t *pid,
                float *ox, float *oy, float *ovx, float *ovy, float *ovz, float *oer,
                uint32_t *opid)
{
    for (int k = 0; k < n; ++k) {
        int p = perm[k];
        ox[k] = x[p]; oy[k] = y[p]; ovx[k] = vx[p]; ovy[k] = vy[p]; ovz[k] = vz[p];
        if (er) oer[k] = er[p];
        opid[k] = pid[p];
    }
}

/* ------------------------------------------------------------------------- */
/* Initial fill.  Reference: Run::assignParticle (Run.cpp:108-157) with the  */
/* Maxwellian of Random::MaxwellDistribution (Random.cpp:30-46).  Per-cell   */
/* count floor(expected + U); positions uniform in the cell, rejected inside */
/* the body for cut cells (D16); velocities drift + N(0, kT/m) by            */
/* Box-Muller; pid = slot.                                                   */
/* ------------------------------------------------------------------------- */
static int inside_body(const or_case *cs, float px, float py)
{
    int in = 0;
    for (int k = 0; k < cs->nseg; ++k) {
        const float *S = cs->seg + 6 * k;
        float ax = S[0], ay = S[1], bx = S[2], by = S[3];
        if ((ay > py) != (by > py)) {
            float xi = ax + ((py - ay) * (bx - ax)) / (by - ay);
            if (px < xi) in ^= 1;
        }
    }
    return in;
}

int or_init_fill(const or_case *cs, const double *expected, const uint8_t *cut,
                 const float *ux, const float *uy, const float *sig, int cap,
                 float *x, float *y, float *vx, float *vy, float *vz, float *er,
                 uint32_t *pid, int *count, uint32_t *err)
{
    int C = cs->ncell;
    long long n = 0;
    for (int c = 0; c < C; ++c) {
        uint32_t ctr[4] = { 0xFFFFFFFFu, 0u, OR_SUB_INIT, 0u }, w[4];
        or_philox(ctr, (uint32_t)c, cs->seed, w);
        double U = (double)w[0] * 0x1p-32;
        int k = expected[c] > 0.0 ? (int)floor(expected[c] + U) : 0;
        count[c] = k;
        n += k;
    }
    if (n > cap) { *err |= OR_ERR_CAPACITY; return -1; }
    int i = 0;
    uint32_t e = 0;
    for (int c = 0; c < C; ++c) {
        int ci = c / cs->ny, cj = c % cs->ny;
        /* uncut cells: the cell; cut cells: the bounding box of the fluid part */
        const float *fb = cs->fbox + 4 * (size_t)c;
        float x0 = cut[c] ? fb[0] : (float)ci * cs->dx;
        float y0 = cut[c] ? fb[1] : (float)cj * cs->dy;
        float wx = cut[c] ? fb[2] - fb[0] : cs->dx;
        float wy = cut[c] ? fb[3] - fb[1] : cs->dy;
        const uint64_t fm = cut[c] ? cs->fmask[c] : ~0ull;
        const int nm = __builtin_popcountll(fm);
        const float sw = wx * 0.125f, sh = wy * 0.125f;
        for (int k = 0; k < count[c]; ++k, ++i) {
            uint32_t w[4];
            float px = x0, py = y0;
            int ok = 0;
            if (!cut[c]) {
                uint32_t ctr[4] = { (uint32_t)k, 0u, OR_SUB_INIT, 1u };
                or_philox(ctr, (uint32_t)c, cs->seed, w);
                px = x0 + wx * u01(w[0]);
                py = y0 + wy * u01(w[1]);
                ok = 1;
            }
            /* cut cells: a sub-box of the fluid bounding box that touches
             * fluid (host table fmask), a point in it, rejected inside the body */
            for (int tr = 0; tr < OR_FILL_TRIES && !ok; ++tr) {
                uint32_t ctr[4] = { (uint32_t)k, 0u, OR_SUB_INIT, 1u + (uint32_t)tr };
                or_philox(ctr, (uint32_t)c, cs->seed, w);
                int sb = nth_set_bit(fm, mulhi_idx(w[0], nm));
                px = x0 + sw * ((float)(sb & 7) + u01(w[1]));
                py = y0 + sh * ((float)(sb >> 3) + u01(w[2]));
                ok = !inside_body(cs, px, py);
            }
            if (!ok) e |= OR_ERR_FILL;
            uint32_t cv[4] = { (uint32_t)k, 0u, OR_SUB_INIT, 0u };
            or_philox(cv, (uint32_t)c, cs->seed, w);
            float r1 = sqrtf(-2.0f * or_logf(u01o(w[0])));
            float s1, c1, s2, c2;
            or_sincos2pi(u01(w[1]), &s1, &c1);
            float r2 = sqrtf(-2.0f * or_logf(u01o(w[2])));
            or_sincos2pi(u01(w[3]), &s2, &c2);
            x[i] = px; y[i] = py;
            vx[i] = ux[c] + sig[c] * (r1 * c1);
            vy[i] = uy[c] + sig[c] * (r1 * s1);
            vz[i] = sig[c] * (r2 * c2);
            pid[i] = (uint32_t)i;
            /* rotational energy at the cell's kT/m = sig^2 (Run.cpp:134) */
            if (er) {
                uint32_t cr[4] = { (uint32_t)k, 0u, OR_SUB_INIT, 0x80000000u };
                or_philox(cr, (uint32_t)c, cs->seed, w);
                er[i] = rot_energy(w[0], sig[c] * sig[c]);
            }
        }
    }
    *err |= e;
    return (int)n;
}

/* ------------------------------------------------------------------------- */
/* Full step: move -> inject -> bin -> scan -> scatter -> collide ->         */
/* (sample) -> reorder.  Reference phase order: Run::solver                  */
/* (Run.cpp:259-299): particlemove, ressignParticle, collision, Sample.      */
/* The parallel version gives bit-identical results to the serial one.       */
/* ------------------------------------------------------------------------- */
int or_step(const or_case *cs, or_state *st, int sample, int nthreads)
{
    int nt = nthreads > 1 ? nthreads : 1;
    int n = st->n;
    uint32_t err = 0;
#ifdef _OPENMP
#pragma omp parallel for schedule(static) num_threads(nt) reduction(| : err) if (nt > 1)
#endif
    for (int i = 0; i < n; ++i) {
        uint32_t e = 0;
        move_one(cs, &st->x[i], &st->y[i], &st->vx[i], &st->vy[i], &st->vz[i],
                 st->er ? &st->er[i] : NULL, st->pid[i], st->step, cs->dt, &st->dead[i], &e);
        err |= e;
    }
    int J[4];
    or_inject_counts(cs, st->step, J);
    int Jt = J[0] + J[1] + J[2] + J[3];
    if (n + Jt > st->cap) { Jt = st->cap - n; err |= OR_ERR_CAPACITY; }
    uint32_t einj = 0;
#ifdef _OPENMP
#pragma omp parallel for schedule(static) num_threads(nt) reduction(| : einj) if (nt > 1)
#endif
    for (int q = 0; q < Jt; ++q) {
        int f = 0, acc = J[0];
        while (f < 3 && q >= acc) { ++f; acc += J[f]; }
        uint32_t p = st->next_pid + (uint32_t)q;
        uint32_t e = 0;
        st->pid[n + q] = p;
        inject_one(cs, f, p, st->step, &st->x[n + q], &st->y[n + q], &st->vx[n + q],
                   &st->vy[n + q], &st->vz[n + q], st->er ? &st->er[n + q] : NULL, &st->dead[n + q], &e);
        einj |= e;
    }
    err |= einj;
    int ntot = n + Jt;
    int C = cs->ncell;

    /* bin: cell ids in parallel, ranks by a stable two-pass counting sort */
    if (nt == 1) {
        memset(st->count, 0, sizeof(int) * (size_t)C);
        or_bin(cs, ntot, st->x, st->y, st->dead, st->cell, st->count, st->slot);
    } else {
#ifdef _OPENMP
        int *hist = (int *)calloc((size_t)nt * (size_t)C, sizeof(int));
#pragma omp parallel num_threads(nt)
        {
            /* the team may be smaller than asked for (OMP_THREAD_LIMIT,
             * nested regions): split by the threads that actually run */
            const int T = omp_get_num_threads();
            int tid = omp_get_thread_num();
            int lo = (int)((long long)ntot * tid / T), hi = (int)((long long)ntot * (tid + 1) / T);
            int *h = hist + (size_t)tid * C;
            for (int i = lo; i < hi; ++i) {
                if (st->dead[i]) { st->cell[i] = -1; st->slot[i] = 0; continue; }
                int c = cell_of(cs, st->x[i], st->y[i]);
                st->cell[i] = c;
                st->slot[i] = h[c]++;
            }
#pragma omp barrier
#pragma omp for schedule(static)
            for (int c = 0; c < C; ++c) {
                int s = 0;
                for (int t = 0; t < T; ++t) { int v = hist[(size_t)t * C + c]; hist[(size_t)t * C + c] = s; s += v; }
                st->count[c] = s;
            }
#pragma omp barrier
            for (int i = lo; i < hi; ++i)
                if (st->cell[i] >= 0) st->slot[i] += h[st->cell[i]];
        }
        free(hist);
#endif
    }
    or_scan(st->count, C, st->start);
#ifdef _OPENMP
#pragma omp parallel for schedule(static) num_threads(nt) if (nt > 1)
#endif
    for (int i = 0; i < ntot; ++i)
        if (st->cell[i] >= 0) st->perm[st->start[st->cell[i]] + st->slot[i]] = i;

    or_collide(cs, st->step, st->start, st->perm, st->x, st->y, st->vx, st->vy, st->vz, st->er,
               st->sigcmax, st->mrem, st->ncoll_cell, nt);
    long long tc = 0;
    for (int c = 0; c < C; ++c) tc += st->ncoll_cell[c];
    st->total_coll += tc;
    if (sample) {
        or_sample(C, st->start, st->perm, st->vx, st->vy, st->vz, st->er, st->ncoll_cell, st->acc);
        st->nsamples += 1;
    }
    int nn = st->start[C];
#ifdef _OPENMP
#pragma omp parallel for schedule(static) num_threads(nt) if (nt > 1)
#endif
    for (int k = 0; k < nn; ++k) {
        int p = st->perm[k];
        st->bx[k] = st->x[p]; st->by[k] = st->y[p];
        st->bvx[k] = st->vx[p]; st->bvy[k] = st->vy[p]; st->bvz[k] = st->vz[p];
        st->bpid[k] = st->pid[p];
        if (st->er) st->ber[k] = st->er[p];
    }
    float *t;
    uint32_t *tp;
    t = st->x; st->x = st->bx; st->bx = t;
    t = st->y; st->y = st->by; st->by = t;
    t = st->vx; st->vx = st->bvx; st->bvx = t;
    t = st->vy; st->vy = st->bvy; st->bvy = t;
    t = st->vz; st->vz = st->bvz; st->bvz = t;
    tp = st->pid; st->pid = st->bpid; st->bpid = tp;
    t = st->er; st->er = st->ber; st->ber = t;
    st->n = nn;
    st->next_pid += (uint32_t)Jt;
    st->step += 1;
    st->err |= err;
    return nn;
}

/* ---- batch helpers for the tests ------------------------------------------ */
void or_logf_n(int n, const float *in, float *out) { for (int i = 0; i < n; ++i) out[i] = or_logf(in[i]); }
void or_expf_n(int n, const float *in, float *out) { for (int i = 0; i < n; ++i) out[i] = or_expf(in[i]); }
void or_sincos2pi_n(int n, const float *in, float *s, float *c)
{
    for (int i = 0; i < n; ++i) or_sincos2pi(in[i], &s[i], &c[i]);
}
/* out[4*i..] = philox(ctr[4*i..], k0[i], k1[i]) */
void or_philox_n(int n, const uint32_t *ctr, const uint32_t *k0, const uint32_t *k1, uint32_t *out)
{
    for (int i = 0; i < n; ++i) or_philox(ctr + 4 * i, k0[i], k1[i], out + 4 * i);
}
