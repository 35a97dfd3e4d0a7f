/*
 * dsmc_abi.h -- C ABI of the MI355X DSMC hot path (libcudadsmc.so).
 *
 * Replaces the two device-container classes the reference driver calls
 * directly (cudaDSMC src/particle/Particle.cuh:7-38 `GPUParticles`,
 * src/cell/Cell.cuh:7-35 `GPUCells`) and their kernels.  Plain pointers and
 * sizes only.  Every call returns 0 on success or a negative status; the
 * message of the last failure on the calling thread is dsmc_last_error().
 * A handle owns its device memory and one HIP stream; calls on one handle
 * are not thread-safe (one host thread per handle, as in the reference).
 * Host pointers are borrowed for the duration of the call only.
 */
#ifndef DSMC_ABI_H
#define DSMC_ABI_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define DSMC_OK             0
#define DSMC_ERR_ARG       -1   /* invalid argument or state */
#define DSMC_ERR_HIP       -2   /* HIP runtime failure */
#define DSMC_ERR_CAPACITY  -3   /* particle capacity exceeded */
#define DSMC_ERR_GEOMETRY  -4   /* invalid body polygon / cut-cell table */
#define DSMC_ERR_NODEVICE  -5   /* no HIP device */

/* boundary models (face_bc[], wall_bc) */
#define DSMC_BC_STREAM    0     /* freestream: particles leave, Maxwellian flux enters */
#define DSMC_BC_SPECULAR  1
#define DSMC_BC_DIFFUSE   2

/* device error word bits (dsmc_status.err) */
#define DSMC_DEV_ERR_BOUNCE   1  /* a particle needed more than 64 wall events in one step */
#define DSMC_DEV_ERR_INFLOW   2  /* inflow speed sampler hit its try limit */
#define DSMC_DEV_ERR_CAPACITY 4  /* injection clamped at capacity */
#define DSMC_DEV_ERR_FILL     8  /* fill could not place a particle in a cut cell's fluid part */

/* Case parameters.  Field names follow the reference's src/Param.h:5-37
 * (boltz, mass, diam, Vtl, L1, L2, L3, N1, N2, T, T_wall, V_jet, tau, Fn);
 * the rest are the run settings a case or a sweep of cases varies. */
typedef struct dsmc_params {
    double boltz;        /* J/K */
    double mass;         /* kg */
    double diam;         /* VHS reference diameter d_ref, m */
    double Vtl;          /* VHS viscosity index omega */
    double Tref;         /* VHS reference temperature, K */
    double L1, L2, L3;   /* domain x, y extent and unit depth, m */
    int    N1, N2;       /* cells in x and y */
    double n_inf;        /* freestream / initial number density, 1/m^3 */
    double T;            /* freestream / initial temperature, K */
    double V_jet;        /* freestream x velocity, m/s */
    double V_y;          /* freestream y velocity, m/s */
    double T_wall;       /* body wall temperature, K */
    int    wall_bc;      /* body wall model: DSMC_BC_DIFFUSE or DSMC_BC_SPECULAR */
    int    face_bc[4];   /* x=0, x=L1, y=0, y=L2 */
    double face_T[4];    /* wall temperature of diffuse faces */
    double tau;          /* time step, s */
    double ppc;          /* simulated particles per full fluid cell at n_inf (sets Fn) */
    double Fn;           /* real molecules per simulated particle; > 0 overrides ppc */
    uint64_t seed;       /* Philox key of this instance */
    int64_t capacity;    /* particle capacity; 0 = 4x the initial fill estimate */
    int    sample_start; /* first step whose state is sampled; < 0 = never */
    /* Rotational energy (src/Param.h:14 zeta, the reference's N2): 0 for a
     * monatomic gas, 2 for a diatomic one with Larsen-Borgnakke exchange in
     * collisions.  rot_prob = 1/Z_rot, the fraction of collisions that
     * exchange, in (0, 1] (1: every collision, as the reference's scheme);
     * ignored for zeta = 0. */
    int    zeta;
    double rot_prob;
} dsmc_params;

/* Immersed body: closed polygon, counter-clockwise, n_points >= 3
 * (reference src/object/Geom.h `m_points`, Circle.cpp:10-38). */
typedef struct dsmc_body {
    int n_points;
    const double *xy;    /* 2*n_points */
} dsmc_body;

/* Derived constants of a created case (what the kernels use). */
typedef struct dsmc_case_info {
    int    ncell, nseg, ncut;
    int    cseg_len;             /* length of the per-cell segment index list */
    double Fn, dt;
    float  L1, L2, dx, dy;
    float  ntc_coef, K1, e_exp;  /* NTC: Fn*dt/2; VHS sigma*c_r = K1*(c_r^2)^e_exp */
    float  sigcmax0;             /* initial (sigma*c_r)max, reference Vmax*pi*d^2 */
    float  body_vmp, body_sig;
    float  face_vmp[4], face_sig[4];
    double inj_e[4];             /* expected inflow particles per step per face */
    float  inj_a[4];             /* inward normal speed ratio per face */
    float  inj_vmp, inj_sig, inj_ux, inj_uy;
    float  margin;
    int    body_bc, face_bc[4];
    uint32_t seed;
    double cell_volume;          /* full cell volume dx*dy*L3 */
    double ppc_eff;              /* particles per full fluid cell at n_inf */
    int64_t capacity;            /* particle capacity */
    int    zeta;                 /* rotational degrees of freedom, 0 or 2 */
    int    sample_blocks;        /* SoA blocks of the sample accumulators: 6, 7 with zeta = 2 */
    float  rot_p;                /* Larsen-Borgnakke: fraction of collisions that exchange */
    float  lb_ia, lb_ib;         /* 1/(5/2-omega), 1/(7/2-omega): the translational-fraction draw */
    float  body_ekt, face_ekt[4], inj_ekt;  /* kT/m of diffuse walls and freestream, J/kg */
} dsmc_case_info;

typedef struct dsmc_status {
    uint32_t step;               /* steps completed */
    uint32_t n_live;             /* particles in the current state */
    uint32_t next_pid;           /* next particle id to assign */
    uint32_t err;                /* DSMC_DEV_ERR_* bits */
    uint64_t particle_steps;     /* sum over completed steps of particles moved */
    uint64_t collisions;         /* collisions over all completed steps */
    uint64_t nsamples;           /* sampled steps */
    uint32_t last_collisions;    /* collisions in the last completed step */
    uint32_t n_move;             /* particles processed by the last move (incl. inflow) */
} dsmc_status;

typedef struct dsmc_handle_s *dsmc_handle;

const char *dsmc_last_error(void);
int dsmc_device_count(int *n);
/* Compute units of the current device (sizes the launch settings). */
int dsmc_device_cus(int *n);

/* Lifecycle.  Replaces GPUParticles(N) / GPUCells(C) construction and
 * GPUCells::UploadFromHost's cut-cell tables (Particle.cuh:11, Cell.cuh:10-12;
 * Run.cpp:52-56,102-105,197-223).  body may be NULL (no immersed body). */
int dsmc_create(const dsmc_params *p, const dsmc_body *body, dsmc_handle *out);
int dsmc_destroy(dsmc_handle h);
/* Host-built cut-cell tables (the reference's GPUCells::UploadFromHost ifCut
 * and GPUSegment tables, src/cell/Cell.cuh:12, as Run::TransferCellsFromHostToDevice
 * builds them from its mesh, src/Run.cpp:197-223).  The case takes the body
 * from the wall segments instead of a polygon:
 *   seg[6*nseg]  ax ay bx by nx ny per wall segment (the reference GPUSegment's
 *                end points and normal, GPUSegment.cuh:11-19); the segments form
 *                one closed counter-clockwise outline in order -- segment k ends
 *                where k+1 starts, the last where the first starts -- and n is
 *                the unit normal (ey, -ex)/|e| into the fluid;
 *   frac[ncell]  fluid fraction of every cell (1 uncut, 0 solid; the fluid
 *                volume is frac*dx*dy*L3), or NULL to clip the outline here;
 *   cut[ncell]   the ifCut flags (1 exactly where 0 < frac < 1), or NULL.
 * nseg = 0 is a case without a body.  The per-cell reach lists of the move
 * and the fill helpers of cut cells are derived from the outline.  Invalid
 * tables return DSMC_ERR_GEOMETRY before any device is touched. */
typedef struct dsmc_cut_tables {
    int nseg;
    const double *seg;
    const double *frac;
    const uint8_t *cut;
} dsmc_cut_tables;
int dsmc_create_from_tables(const dsmc_params *p, const dsmc_cut_tables *t, dsmc_handle *out);
int dsmc_case_info_get(dsmc_handle h, dsmc_case_info *out);
/* Host-built cut-cell tables: seg[6*nseg] (ax ay bx by nx ny, n into the
 * fluid), cseg_start[ncell+1], cseg_idx[cseg_len], vol[ncell] fluid volume,
 * vol_ntc[ncell] NTC volume, cut[ncell] partial-cell flag, fbox[4*ncell]
 * bounding box (x0 y0 x1 y1) of the fluid part of cut cells.  NULL skips. */
int dsmc_case_tables(dsmc_handle h, float *seg, int *cseg_start, int *cseg_idx,
                     double *vol, float *vol_ntc, uint8_t *cut, float *fbox, uint64_t *fmask);
/* Host only (no device needed): the same derived constants and tables for a
 * parameter set, plus frac[ncell] fluid fraction.  Call with NULL tables to
 * get the sizes in info first. */
int dsmc_case_describe(const dsmc_params *p, const dsmc_body *body, dsmc_case_info *info,
                       float *seg, int *cseg_start, int *cseg_idx, double *vol,
                       float *vol_ntc, uint8_t *cut, double *frac, float *fbox, uint64_t *fmask);

/* Initial state.  Replaces Run::assignParticle + TransferParticlesFromHostToDevice
 * (Run.cpp:108-191).  init_freestream fills every fluid cell at n_inf, T,
 * (V_jet, V_y); init_fields fills from per-cell n, ux, uy, T (warm start).
 * Both reset the step counter, the particle ids and the cell state. */
int dsmc_init_freestream(dsmc_handle h);
int dsmc_init_fields(dsmc_handle h, const double *n, const double *ux,
                     const double *uy, const double *T);

/* Particle state I/O (GPUParticles::UploadFromHost, GPUParticle.cu:41-58).
 * upload replaces the state with n particles; pid may be NULL (0..n-1). */
int dsmc_upload_particles(dsmc_handle h, int64_t n, const float *x, const float *y,
                          const float *vx, const float *vy, const float *vz,
                          const uint32_t *pid);
int dsmc_download_particles(dsmc_handle h, int64_t cap, float *x, float *y, float *vx,
                            float *vy, float *vz, uint32_t *pid, int64_t *n_out);
/* Rotational energy of a diatomic gas (zeta = 2; the reference's d_Erot,
 * Particle.cuh:30, GPUParticle.cu:19,44-53), as energy per unit mass e in
 * J/kg, E_rot = mass * e, one float per particle in the order of the
 * particle arrays.  dsmc_upload_particles sets e = 0; dsmc_upload_rotational
 * then sets the n = n_live energies of the current state (finite, >= 0).
 * download: the current state (moved = 0) or the n_move particles of the
 * last dsmc_move (moved = 1).  DSMC_ERR_ARG for a monatomic case. */
int dsmc_upload_rotational(dsmc_handle h, int64_t n, const float *er);
int dsmc_download_rotational(dsmc_handle h, int moved, int64_t cap, float *er, int64_t *n_out);

/* The three phases of Run::solver (Run.cpp:159-162,245-256):
 *  dsmc_move     = Run::particlemove (GPUParticles::Move + DeleteInvalid) plus
 *                  the inflow of GPUParticles::Injet and the atomic per-cell
 *                  count of GPUCells::CalparticleNum -- one fused kernel;
 *  dsmc_reassign = GPUCells::CalparticleStartIndex + GPUParticles::Sort
 *                  (exclusive scan, scatter into the cell-sorted permutation);
 *  dsmc_collide  = GPUCells::Collision (+ GPUCells::Sample when the step is
 *                  >= sample_start): stable order within each cell, state
 *                  reordered and compacted into cell order, NTC/VHS collisions.
 * A step is move, reassign, collide, in that order. */
int dsmc_move(dsmc_handle h);
int dsmc_reassign(dsmc_handle h);
int dsmc_collide(dsmc_handle h);
/* nsteps full steps, launched eagerly (graph = 0) or by replaying a captured
 * hipGraph of the step (graph = 1).  Asynchronous; dsmc_sync waits. */
int dsmc_run(dsmc_handle h, int nsteps, int graph);
int dsmc_sync(dsmc_handle h);
int dsmc_status_get(dsmc_handle h, dsmc_status *out);

/* Intermediate buffers for parity checks.  cell is valid after dsmc_move
 * for the n_move processed particles (cell -1 = deleted); count after
 * dsmc_move; start[ncell+1] and perm (ascending particle index per cell)
 * after dsmc_reassign, perm until the next dsmc_reassign.  perm is always
 * kept: dsmc_set_keep_perm is accepted for callers that set it and changes
 * nothing. */
int dsmc_set_keep_perm(dsmc_handle h, int on);
/* The n_move particles processed by the last dsmc_move (moved + inflow, in
 * processing order, deleted ones included); valid until dsmc_collide. */
int dsmc_download_moved(dsmc_handle h, int64_t cap, float *x, float *y, float *vx, float *vy,
                        float *vz, uint32_t *pid, int64_t *n_out);
int dsmc_download_bins(dsmc_handle h, int64_t cap, int *cell, int *count);
int dsmc_download_start(dsmc_handle h, int *start);
int dsmc_download_perm(dsmc_handle h, int64_t cap, int *perm);
int dsmc_cell_state(dsmc_handle h, float *sigcmax, float *mrem, int *ncoll);
int dsmc_set_cell_state(dsmc_handle h, const float *sigcmax, const float *mrem);

/* Checkpoint / resume.  A run continues bit-identically from a saved state
 * given: the particles in their stored order (dsmc_upload_particles with
 * pids), the cell state (dsmc_set_cell_state), the counters below and the
 * sample accumulators.  dsmc_restore_counters sets step, next_pid,
 * particle_steps, collisions and nsamples from *s (n_live stays that of the
 * last upload; the other fields are ignored).  The reference writes fields
 * only (Output.cpp:12-154) and has no resume; these replace nothing. */
int dsmc_restore_counters(dsmc_handle h, const dsmc_status *s);
int dsmc_upload_samples(dsmc_handle h, const double *acc, uint64_t nsamples);

/* Sampling (GPUCells::Sample + DownloadToHost, GPUCell.cu:83-91,147-167).
 * acc[S*ncell], S = dsmc_case_info.sample_blocks: SoA blocks sum n, sum vx,
 * sum vy, sum vz, sum v^2, collisions and, for zeta = 2, the sum of the
 * rotational energy per unit mass.  Fields: number density n, velocity,
 * translational temperature, pressure; dsmc_download_rot_temperature gives
 * T_rot = 2 m <e> / (zeta k) per cell (DSMC_ERR_ARG for a monatomic case),
 * and the overall temperature is (3 T + zeta T_rot) / (3 + zeta). */
/* One sample of the current state into the accumulators (GPUCells::Sample,
 * src/cell/Cell.cuh:22, GPUCell.cu:147-167,336-402, which Run::solver calls
 * every 100 steps, src/Run.cpp:294-296): per cell, the particle count, the
 * velocity and v^2 sums and the last step's collisions, added as one sampled
 * step -- the same sums, in the same order, as the sampling a step does by
 * itself from params.sample_start on.  The state is in cell order after a
 * completed step; after an upload or a fill (no step yet) this returns
 * DSMC_ERR_ARG.  Asynchronous on the handle's stream. */
int dsmc_sample(dsmc_handle h);
int dsmc_reset_samples(dsmc_handle h);

/* Adaptive collision sub-cells and the variable time step: Cell::genAMRmesh
 * (src/cell/Cell.h:55, Cell.cpp:168-279), Element::genAMRmesh
 * (src/meshes/Element.h:44, Element.cpp:158-182), Cell::VTS (Cell.h:52,
 * Cell.cpp:117-136) and Cell::collision in the leaves (Cell.cpp:335-352).
 * A cell's grid is packed n2x | n2y << 8 | n3x << 16 | n3y << 24 -- n2x x n2y
 * level-2 cells, each with n3x x n3y level-3 cells -- and 0 means none; its
 * leaves form a uniform (n2x n3x) x (n2y n3y) grid.  From the next step on,
 * a cell with more than one and at most 256 leaves and 2 .. 2048 particles
 * orders its particles stably by leaf and collides leaf by leaf with the
 * leaf volume (fluid volume / leaves).
 * dsmc_set_subcells: ncell packed grids (NULL: none at all; every level of a
 * non-zero grid needs at least one sub-cell, else DSMC_ERR_ARG).
 * dsmc_amr_update: per cell from the sampling sums (time-averaged; at least
 * one sampled step, else DSMC_ERR_ARG): mean free path
 * lam = m / (sqrt(2) rho pi d^2) (T/T_inf)^(omega - 1/2), VTS time step
 * dt = lam / (2 (|u| + sqrt(2kT/m))), and the grid by genAMRmesh's rule --
 * level 2 of side lam/2, n = max(1, ceil(L/side)), at most 16 per axis;
 * level 3 from twice the level-2 x side for a cell without a grid, else from
 * lam -- a cell without gas gets none.  Switches the sub-cells on.
 * dsmc_download_amr: the grids, mean free paths (m) and VTS steps (s); any
 * pointer may be NULL.  Zeros before the first update.  Single handles only
 * (a batched sweep has no sub-cells). */
int dsmc_set_subcells(dsmc_handle h, const uint32_t *sub);
int dsmc_amr_update(dsmc_handle h);
int dsmc_download_amr(dsmc_handle h, uint32_t *sub, float *lam, double *dt);
int dsmc_download_samples(dsmc_handle h, double *acc, uint64_t *nsamples);
int dsmc_download_fields(dsmc_handle h, double *n, double *ux, double *uy,
                         double *uz, double *T, double *p);
int dsmc_download_rot_temperature(dsmc_handle h, double *Trot);

/* ---- Binning on caller device arrays -----------------------------------
 * The reference's GPUCells::CalparticleNum -> CalparticleStartIndex ->
 * GPUParticles::Sort chain (src/cell/Cell.cuh:14-20, GPUCell.cu:93-120,169-198;
 * src/particle/Particle.cuh, GPUParticle.cu:142-147,365-373) on arrays the
 * caller owns, with no handle: for n particles at FP32 (x, y), device
 * pointers all,
 *   cell[n]      cell id j + i*N2, i = (int)(x/dx), j = (int)(y/dy), clamped
 *                into the grid (the reference leaves out-of-grid ids stale);
 *   start[C+1]   exclusive scan of the per-cell counts (thrust::exclusive_scan
 *                of d_particleNum), start[C] = n;
 *   perm[n]      cell-sorted permutation, ascending particle index inside
 *                each cell -- the serial order of the reference's atomicAdd
 *                ranks, where the GPU's own order is arbitrary;
 *   local_id[n]  rank of each particle in its cell in that order (d_localID);
 *                may be NULL.
 * dx = L1/N1, dy = L2/N2 in double, then rounded to FP32 (as a case does).
 * stream: a hipStream_t, or NULL for the null stream; the call returns when
 * the results are ready.  Scratch is allocated and freed per call. */
typedef struct dsmc_grid {
    int N1, N2;
    double L1, L2;
} dsmc_grid;
int dsmc_bin_sort(const dsmc_grid *g, const float *x, const float *y, int64_t n, int *cell, int *start, int *perm,
                  int *local_id, void *stream);
/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): dsmc_bin_sort with the same arguments and results, whose
 * scratch is a static device array of the library (about 12.6 MB per GPU,
 * held until the process ends) instead of a per-call hipMalloc.  It calls no
 * allocating routine; grids above DSMC_BIN_SORT_MAX_CELLS cells are rejected
 * (DSMC_ERR_ARG).  It is asynchronous on `stream`: the results are ready when
 * the stream reaches the end of the call.  All calls on one GPU share that
 * GPU's scratch, so calls that overlap on one GPU (two streams, two threads)
 * race; each call clears what it needs. */
#define DSMC_BIN_SORT_MAX_CELLS 1048576
int dsmc_bin_sort_noalloc(const dsmc_grid *g, const float *x, const float *y, int64_t n, int *cell, int *start,
                          int *perm, int *local_id, void *stream);

/* Launch settings (no reference counterpart; the reference hard-codes
 * 128-thread blocks, GPUParticle.cu:60-74).  Results never depend on them --
 * only the time does -- which is what lets the autotuner change them.
 * Changing them drops the captured graph. */
typedef struct dsmc_launch {
    int move_blocks;     /* k_move grid, blocks of 256 threads */
    int scatter_blocks;  /* k_place and k_rank grid */
    int cell_blocks;     /* k_cell grid (persistent blocks pulling tiles) */
    int cell_tile;       /* particles per k_cell tile, 256..2048 */
    int graph_steps;     /* steps per captured hipGraph, 1..64 */
} dsmc_launch;
int dsmc_get_launch(dsmc_handle h, dsmc_launch *out);
int dsmc_set_launch(dsmc_handle h, const dsmc_launch *cfg);

/* Per-kernel timing with HIP events on the handle's stream (eager steps
 * only).  ms[k] = mean duration of kernel k over timed steps, k = 0 move,
 * 1 scan, 2 scatter, 3 cell; bytes[k] = mean algorithmic bytes per launch of
 * kernel k over the same steps.  on = 1 events and bytes, 2 bytes only (no
 * events: ms stays 0), 0 off; any other value is DSMC_ERR_ARG. */
int dsmc_set_timing(dsmc_handle h, int on);
int dsmc_kernel_times(dsmc_handle h, double ms[4], double bytes[4], int64_t *launches);

/* ---- Batched instances ------------------------------------------------
 * No counterpart in the reference, which runs one case per process.  n
 * independent cases (own grid, body, gas state, walls, seed, sampling start;
 * uneven particle and cell counts) advance together: every phase of a step
 * is one launch over all instances, driven by per-instance offset tables,
 * and dsmc_batch_run replays the whole step as a hipGraph.  Each instance
 * evolves bit-identically to a dsmc_handle of its own.
 *   bodies: n entries (n_points = 0 for none), or NULL for no bodies;
 *   init_fields: per-cell n, ux, uy, T of all instances concatenated in
 *     instance order (sum of ncell values);
 *   status / downloads take the instance index i; dsmc_batch_status with
 *     i = -1 gives the totals (particle_steps, collisions, n_live). */
typedef struct dsmc_batch_s *dsmc_batch;
int dsmc_batch_create(const dsmc_params *params, const dsmc_body *bodies, int n, dsmc_batch *out);
int dsmc_batch_destroy(dsmc_batch b);
int dsmc_batch_case_info(dsmc_batch b, int i, dsmc_case_info *out);
int dsmc_batch_init_freestream(dsmc_batch b);
int dsmc_batch_init_fields(dsmc_batch b, const double *n, const double *ux, const double *uy,
                           const double *T);
int dsmc_batch_run(dsmc_batch b, int nsteps, int graph);
int dsmc_batch_sync(dsmc_batch b);
int dsmc_batch_status(dsmc_batch b, int i, dsmc_status *out);
int dsmc_batch_download_particles(dsmc_batch b, int i, int64_t cap, float *x, float *y, float *vx,
                                  float *vy, float *vz, uint32_t *pid, int64_t *n_out);
int dsmc_batch_cell_state(dsmc_batch b, int i, float *sigcmax, float *mrem, int *ncoll);
int dsmc_batch_reset_samples(dsmc_batch b);
int dsmc_batch_download_samples(dsmc_batch b, int i, double *acc, uint64_t *nsamples);
int dsmc_batch_download_fields(dsmc_batch b, int i, double *n, double *ux, double *uy, double *uz,
                               double *T, double *p);
/* Steady-state monitoring of a sweep (no counterpart in the reference, whose
 * Run::solver samples one case and writes its fields at the end): the
 * particle-count sums of every cell of every instance (counts[total cells],
 * instance order) and, when nsamples is not NULL, each instance's number of
 * sampled steps (nsamples[n]), in one transfer.  Sums only grow between
 * dsmc_batch_reset_samples calls, so differences of two reads give the
 * counts of the steps in between. */
int dsmc_batch_download_counts(dsmc_batch b, double *counts, uint64_t *nsamples);
/* The same for the five moment sums (no counterpart in the reference): sums[5
 * * total cells] as five blocks -- sum 1, sum vx, sum vy, sum vz, sum v^2 --
 * each over every cell of every instance in instance order, one transfer.
 * Differences of two reads give the moments of the steps in between, from
 * which a steady-state test of density, velocity and temperature follows. */
int dsmc_batch_download_moments(dsmc_batch b, double *sums, uint64_t *nsamples);
/* Signed distance (m) from every cell centre to its instance's body outline,
 * negative inside the body, +inf without a body (d[total cells], instance
 * order; no counterpart in the reference: an input feature of the learned
 * warm start). */
int dsmc_batch_wall_distance(dsmc_batch b, double *d);
/* Retires instances from a sweep (no counterpart in the reference, which
 * runs one case per process): keeps instances keep[0..nkeep) -- strictly
 * ascending indices -- which become instances 0..nkeep-1, and frees the
 * rest.  Each kept instance continues bit-identically (particles, cell
 * state, sample sums, counters and random streams carry over); the totals
 * of dsmc_batch_status(b, -1, ...) become the kept instances' sums.  The
 * handle stays valid; on failure it is unchanged. */
int dsmc_batch_compact(dsmc_batch b, const int *keep, int nkeep);
int dsmc_batch_get_launch(dsmc_batch b, dsmc_launch *out);
int dsmc_batch_set_launch(dsmc_batch b, const dsmc_launch *cfg);
/* Per-kernel timing of batched steps (eager), as dsmc_set_timing. */
int dsmc_batch_set_timing(dsmc_batch b, int on);
int dsmc_batch_kernel_times(dsmc_batch b, double ms[4], double bytes[4], int64_t *launches);

#ifdef __cplusplus
}
#endif
#endif
