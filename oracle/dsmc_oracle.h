/*
 * dsmc_oracle.h -- CPU restatement of the reference DSMC hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker that the HIP
 * kernels of cudadsmc_amd are compared against, and the timed CPU baseline
 * of bench.py.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it.  The product never links or calls it.
 *
 * Every routine cites the reference file:line whose behaviour it restates
 * (paths relative to the cudaDSMC source tree).  Deliberate deviations from
 * the reference are the defect fixes named D<n> in the comments; each one is
 * listed in oracle/README.md with the reason.
 *
 * Arithmetic contract shared with the GPU path (this is what makes
 * bit-exact comparison possible):
 *   - particle state is FP32 SoA; all per-particle arithmetic is FP32 with
 *     correctly rounded + - * / sqrt, no FMA contraction (-ffp-contract=off);
 *   - log, exp and sin/cos(2*pi*u) are the routines in this file, built only
 *     from IEEE basic operations (no libm), so both sides round identically;
 *   - sampling sums are FP64.
 */
#ifndef DSMC_ORACLE_H
#define DSMC_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* boundary condition of a domain face or of the body wall */
enum { OR_BC_STREAM = 0, OR_BC_SPECULAR = 1, OR_BC_DIFFUSE = 2 };

/* Philox substreams: counter word 2 */
enum { OR_SUB_MOVE = 1, OR_SUB_INJECT = 2, OR_SUB_COLLIDE = 3, OR_SUB_INIT = 4 };

/* error bits */
enum { OR_ERR_BOUNCE = 1, OR_ERR_INFLOW = 2, OR_ERR_CAPACITY = 4, OR_ERR_FILL = 8 };

#define OR_MAX_BOUNCE 64
#define OR_INFLOW_TRIES 256
#define OR_FILL_TRIES 1024
/* Segment ends overlap by this fraction so that a ray through a vertex shared
 * by two body segments cannot pass between the two rounded tests. */
#define OR_SEG_TOL 1e-5f
/* A particle within this band (fraction of a cell) behind a body segment,
 * moving into it, hits it at once: hit points rounded to the solid side of the
 * wall line cannot leak on a later step. */
#define OR_SEG_BAND 1e-4f

/* One DSMC case: grid, gas, walls, inflow and the cut-cell tables.  The
 * scalar fields are the derived constants the GPU path uses (see
 * dsmc_case_info in include/dsmc_abi.h); the tables are the ones the host
 * geometry layer builds. */
typedef struct or_case {
    int   nx, ny, ncell;
    float L1, L2, dx, dy;
    float dt;
    float ntc_coef;      /* Fn*dt/2 */
    float K1, e_exp;     /* sigma_T*c_r = K1 * (c_r^2)^e_exp  (VHS) */
    int   body_bc;
    float body_vmp, body_sig;          /* sqrt(2kTw/m), sqrt(kTw/m) */
    int   face_bc[4];                  /* x=0, x=L1, y=0, y=L2 */
    float face_vmp[4], face_sig[4];
    double inj_e[4];                   /* expected injections per step per face */
    float inj_a[4];                    /* inward normal speed ratio per face */
    float inj_vmp, inj_sig, inj_ux, inj_uy;
    float margin;                      /* reach covered by a cell's segment list */
    int   nseg;
    const float *seg;                  /* nseg*6: ax ay bx by nx ny (n into fluid) */
    const int   *cseg_start;           /* ncell+1 */
    const int   *cseg_idx;
    const float *vol_ntc;              /* per-cell volume used by NTC */
    const float *fbox;                 /* 4 per cell: fluid-part bounding box of cut cells */
    const uint64_t *fmask;             /* per cell: 8x8 sub-boxes of fbox that touch fluid */
    uint32_t seed;
    /* rotational energy (Larsen-Borgnakke), zeta = 2 rotational degrees of
     * freedom or 0 for a monatomic gas; energies are per unit mass, J/kg */
    int   zeta;
    float rot_p;                       /* fraction of collisions that exchange (1/Z_rot) */
    float lb_ia, lb_ib;                /* 1/(5/2-omega), 1/(7/2-omega) */
    float body_ekt, face_ekt[4], inj_ekt;  /* kT/m of the diffuse walls and the freestream */
    /* adaptive collision sub-cells per cell, packed n2x | n2y << 8 |
     * n3x << 16 | n3y << 24 (0: none); NULL: no cell has any */
    const uint32_t *sub;
} or_case;

/* ---- adaptive collision sub-cells and the variable time step -------------- */
#define OR_AMR_MAX 16          /* sub-cells per level and axis, at most */
#define OR_AMR_CAP 2048        /* cells with more particles collide without sub-cells */
#define OR_AMR_LEAVES 256      /* and so do cells whose grid has more leaves */
typedef struct or_amr_gas {
    double Fn, mass, diam, T_inf, omega, boltz;
} or_amr_gas;
/* Leaves of a packed grid (1 for 0). */
int or_amr_leaves(uint32_t sub, int *nsx, int *nsy);
/* The two-level grid of one cell of size dx x dy from its mean free path
 * (lam_max sizes level 2, lam_min level 3) and whether it had a grid. */
uint32_t or_amr_grid(double dx, double dy, double lam_max, double lam_min, int had);
/* Per cell from the time-averaged sums acc (nsamples samples) and fluid
 * volumes vol: mean free path lam, VTS time step dt (0 where the cell has no
 * gas), and the new grid sub (updated in place: its old value says whether
 * the cell had one). */
void or_amr_update(const or_case *cs, const or_amr_gas *g, const double *acc, long long nsamples,
                   const double *vol, uint32_t *sub, float *lam, double *dt);
/* Leaf of a particle at (x, y) in cell c with an nsx x nsy leaf grid. */
int or_amr_leaf(const or_case *cs, int c, float x, float y, int nsx, int nsy);

/* ---- primitives ---------------------------------------------------------- */
void  or_philox(const uint32_t ctr[4], uint32_t k0, uint32_t k1, uint32_t out[4]);
float or_logf(float x);
float or_expf(float x);
void  or_sincos2pi(float u, float *s, float *c);

/* ---- hot-path phases ----------------------------------------------------- */
/* Free flight + walls + body + outflow for particles [0,n).  er: rotational
 * energy per unit mass (re-drawn at diffuse walls), NULL for zeta = 0. */
void or_move(const or_case *cs, int n, float *x, float *y, float *vx, float *vy,
             float *vz, float *er, const uint32_t *pid, uint32_t step, uint8_t *dead,
             uint32_t *err);
/* Number of inflow particles created at `step` on each face. */
void or_inject_counts(const or_case *cs, uint32_t step, int J[4]);
/* Create this step's inflow particles at slots [0,J) of the output arrays. */
void or_inject(const or_case *cs, uint32_t step, uint32_t next_pid, int J_total,
               float *x, float *y, float *vx, float *vy, float *vz, float *er,
               uint32_t *pid, uint8_t *dead, uint32_t *err);
/* Binning: cell id and stable rank (serial atomicAdd order). */
void or_bin(const or_case *cs, int n, const float *x, const float *y,
            const uint8_t *dead, int *cell, int *count, int *slot);
void or_scan(const int *count, int ncell, int *start /* ncell+1 */);
void or_scatter(int n, const int *cell, const int *slot, const int *start, int *perm);
/* NTC/VHS collisions of every cell, through perm; updates velocities (and,
 * zeta = 2, rotational energies er) in place. */
/* perm is reordered in place within cells that have sub-cells (cs->sub):
 * stably by leaf, so each leaf's particles are contiguous; x, y locate the
 * leaves. */
void or_collide(const or_case *cs, uint32_t step, const int *start, int *perm,
                const float *x, const float *y,
                float *vx, float *vy, float *vz, float *er, float *sigcmax, float *mrem,
                int *ncoll_cell, int nthreads);
/* FP64 sampling accumulators: acc as SoA blocks of ncell: n, vx, vy, vz, v^2,
 * collisions, and with er != NULL a seventh, the rotational energy sum. */
void or_sample(int ncell, const int *start, const int *perm, const float *vx,
               const float *vy, const float *vz, const float *er, const int *ncoll_cell,
               double *acc);
void or_reorder(int n, const int *perm,
                const float *x, const float *y, const float *vx, const float *vy,
                const float *vz, const float *er, const uint32_t *pid,
                float *ox, float *oy, float *ovx, float *ovy, float *ovz, float *oer,
                uint32_t *opid);

/* Initial fill (Run::assignParticle): expected[c] = mean particle count of
 * cell c; cut[c] != 0 marks a partially solid cell (rejection sampling).
 * Writes into arrays of capacity cap; returns particle count (or -1). */
int or_init_fill(const or_case *cs, const double *expected, const uint8_t *cut,
                 const float *ux, const float *uy, const float *sig, int cap,
                 float *x, float *y, float *vx, float *vy, float *vz, float *er,
                 uint32_t *pid, int *count, uint32_t *err);

/* ---- batch helpers for the tests ------------------------------------------ */
void or_logf_n(int n, const float *in, float *out);
void or_expf_n(int n, const float *in, float *out);
void or_sincos2pi_n(int n, const float *in, float *s, float *c);
void or_philox_n(int n, const uint32_t *ctr, const uint32_t *k0, const uint32_t *k1, uint32_t *out);

/* ---- full step ------------------------------------------------------------ */
typedef struct or_state {
    int n, cap;
    uint32_t step, next_pid, err;
    float *x, *y, *vx, *vy, *vz; uint32_t *pid;      /* current state */
    float *bx, *by, *bvx, *bvy, *bvz; uint32_t *bpid; /* scratch */
    uint8_t *dead; int *cell, *slot, *perm;
    int *count, *start, *ncoll_cell;
    float *sigcmax, *mrem;
    double *acc; long long nsamples;
    long long total_coll;
    float *er, *ber;                                 /* rotational energy (zeta = 2), else NULL */
} or_state;

/* One timestep; sample != 0 accumulates sampling sums.  nthreads <= 1 runs
 * serially.  Returns the number of particles after the step. */
int or_step(const or_case *cs, or_state *st, int sample, int nthreads);

#ifdef __cplusplus
}
#endif
#endif
