/*
 * si3d_gpu.h -- C-ABI of libsi3dgpu.so, the MI355X (gfx950) time-step core of
 * pSi3D.
 *
 * The Fortran driver (si3d.f90) keeps reading the input files, computing the
 * grid (bathy) and the initial state (init), and writing every output file.
 * It hands the grid and state to the library once, then replaces its calls to
 * fd / settrap / settrap2 / save with the entry points below, and pulls fields
 * back with s3g_get_field_* before an output is due.  All pointers are host
 * pointers owned by the caller; the library copies from / to them and never
 * keeps them.  Arrays use the reference layout: 3-D fields are (km1, lm1) with k
 * fastest, 2-D fields are (lm1), columns numbered l = 1..lm plus the dry
 * sentinel column lm1 = lm+1 (si3d_types.f90:117-170, si3d_utils.f90:757-806).
 *
 * Return codes follow the reference's error convention (si3d_procedures.f90:
 * 2002-2016): 0 ok, > 0 warning (1 = free-surface solve hit its iteration
 * limit, like NSPCG ier=1), < 0 fatal (the Fortran shim prints
 * s3g_last_error() and STOPs).
 *
 * One host thread calls the library (the OpenMP driver wraps each call in
 * `!$omp masked` + `!$omp barrier`).  All device work of one context runs in
 * order on one HIP stream.
 */
#ifndef SI3D_GPU_H
#define SI3D_GPU_H

#ifdef __cplusplus
extern "C" {
#endif

#define S3G_OK 0
#define S3G_WARN_ITMAX 1
#define S3G_ERR_ARG (-1)
#define S3G_ERR_HIP (-2)
#define S3G_ERR_UNSUPPORTED (-3)
#define S3G_ERR_SOLVER (-6)

/* Precision of the device state. */
#define S3G_FP64 0  /* FP64 everywhere                                         */
#define S3G_MIXED 1 /* FP32 stencils, column solves and SpMV; FP64 dots, true
                       residual, refinement, zeta and continuity             */

/* Free-surface preconditioners. */
#define S3G_PREC_JACOBI 0
#define S3G_PREC_LINE 1 /* exact tridiagonal solve along each N-S line of wet cells */
#define S3G_PREC_MG 2   /* V(1,1) aggregation multigrid on 2x2 blocks of wet cells */

typedef struct s3g_ctx s3g_ctx;

/* Grid as left by bathy (si3d_utils.f90:595-892) and init
 * (si3d_procedures.f90:67-134).  Index arrays are 1-based like the Fortran. */
typedef struct {
  int im1, jm1, km1, lm;
  const int *l2i, *l2j;             /* [lm1]                                 */
  const int *lEC, *lWC, *lNC, *lSC; /* [lm1]; lm1 where the neighbour is dry */
  const int *kmz, *k1z, *k1u, *k1v; /* [lm1]                                 */
  const float *zlevel;              /* [km1]                                 */
  const float *hhs, *hhu, *hhv;     /* [lm1]                                 */
} s3g_grid;

/* Run parameters read by `input` (si3d_utils.f90:21-315). */
typedef struct {
  double dt, dx, dy;
  int itrap, niter, ismooth, iturb, iadv, itrmom, ihd, ibc, isal, ifsurfbc;
  double beta, Av0, Dv0, Ax0, Ay0, f, theta, cd, cw, wa, phi, dtsurfbc;
  int n_surfbc;          /* records of si3d_surfbc_W.txt / si3d_surfbc.txt */
  const double *surfbc;  /* nv*n_surfbc values, nv per record in the order
                            of the file (si3d_boundaryconditions.f90:
                            4160-4300, 4450-4484):
                            ifsurfbc 20: cdw uair vair (nv = 3);
                            1: eta Qsw Qn cdw uair vair (nv = 6);
                            2: eta Qsw Ta Pa Rh Cc cdw uair vair (nv = 9);
                            3: eta Qsw Ta Pa Rh Qlw cdw uair vair (nv = 9);
                            10, 11: eta Pa, then per met station Qsw Ta Rh
                            Cc (10) or Qlw (11) uair vair
                            (nv = 6*nmetstat + 2; 4301-4450)            */
  int precision;         /* S3G_FP64 or S3G_MIXED                          */
  int preconditioner;    /* S3G_PREC_*                                     */
  double pcg_tol;        /* relative FP64 true-residual tolerance          */
  int pcg_maxit;
  int ntr;               /* conservative tracers (ecomod = 0; exTracer/ImTracer,
                            si3d_procedures.f90:4090-4484), fields "tracer1".. and
                            "tracerpp1".. */
  /* met stations of ifsurfbc 10 and 11 (si3d_boundaryconditions.f90:
     4301-4450, 5510-5748): ignored for the other modes                    */
  int icw;               /* drag: 0 cw, 1 Amorocho & DeVries, 2 Andreas    */
  int nmetstat;          /* stations, at most 16                           */
  const double *metxy;   /* 2*nmetstat: x, y of each station (grid units)  */
  int iinterp;           /* 1 Barnes, 2 two-pass Barnes, 3 inverse distance*/
  double gammaB, delNfactor; /* iinterp 2 (si3d_utils.f90:266-275)        */
} s3g_params;

/* Input deck for the standalone driver: the content of si3d_inp.txt,
 * si3d_bathy, si3d_layer.txt and si3d_init.txt (same meaning as above). */
typedef struct {
  int imm, jmm;          /* interior cells (si3d_bathy header)              */
  int ibathyf, km1_file; /* ibathyf < 0: km1_file zlevel records are given  */
  double ddz, dzmin, datadj, zetainit, zl;
  const double *zlevel_file; /* [km1_file]                                  */
  const int *bathy_dm;       /* imm*jmm depths in dm, index (i-2)*jmm+(j-2) */
  const double *init_profile;/* [km1] active scalar per layer               */
  const double *tracer_profiles; /* [ntr*km1] initial tracers (si3d_init.txt
                                    columns 2.., si3d_procedures.f90:330-356) */
} s3g_deck;

typedef struct {
  long steps, fd_calls, solves;
  long pcg_iters_total;      /* over all solves                              */
  int pcg_iters_last;        /* last solve                                   */
  double pcg_relres_last;    /* FP64 true relative residual of the last solve*/
  double zeta_volume;        /* sum of zeta over wet columns (m)             */
  int lm, km1;
  long n_wet;                /* sum of wet layers at creation                */
  long graph_replays;        /* kernel sequences of fd / settrap / settrap2 /
                                save replayed from captured HIP graphs       */
} s3g_stats;

/* --- Drop-in entry points (replace fd/settrap/settrap2/save) ------------- */

/* Replaces the state hand-over after init (si3d.f90:48-68). */
int s3g_init(const s3g_grid *grid, const s3g_params *params, s3g_ctx **out);

/* Copy one state array in / out (names as in si3d_types.f90: "uh", "sal",
 * "q2p", "s", "k1z", ...).  *_f32 for REAL arrays, *_f64 for REAL*8 arrays
 * (q2, q2l and their time levels) or for FP64 callers, *_i32 for indices.
 * "agx", "arx", "agy", "ary" can be read (matmom's per-layer solution, formed
 * on request from the factors the library keeps) but not set. */
int s3g_set_field_f32(s3g_ctx *c, const char *name, const float *host);
int s3g_set_field_f64(s3g_ctx *c, const char *name, const double *host);
int s3g_set_field_i32(s3g_ctx *c, const char *name, const int *host);
int s3g_get_field_f32(s3g_ctx *c, const char *name, float *host);
int s3g_get_field_f64(s3g_ctx *c, const char *name, double *host);
int s3g_get_field_i32(s3g_ctx *c, const char *name, int *host);

/* fd(n, ..., istep, lastiter, ..., iter, ..., thrs): si3d_procedures.f90:382-575 */
int s3g_fd(s3g_ctx *c, int n, int istep, int iter, int lastiter, double thrs);
/* settrap: si3d_procedures.f90:3418-3618 */
int s3g_settrap(s3g_ctx *c);
/* settrap2: si3d_procedures.f90:3944-4087 */
int s3g_settrap2(s3g_ctx *c);
/* save(istep): si3d_procedures.f90:3621-3941 */
int s3g_save(s3g_ctx *c, int istep);

/* One whole time step of the si3d.f90 loop (324-393) / nsteps of them. */
int s3g_step(s3g_ctx *c, int n);
int s3g_run(s3g_ctx *c, int n0, int nsteps);

/* Run one phase of fd alone (phase-level checking): 0 surfbc, 1 horizontal
 * mixing, 2 exmom x, 3 matmom x, 4 exmom y, 5 matmom y, 6 matcon, 7 free
 * surface solve + slayer_h, 8 vel + continuity, 9 exsal, 10 imsal, 11 vel2,
 * 12 smooth, 13 turbulence.  Phases 9 and 10 run fused in fd; phase 9 alone
 * computes the explicit transport terms, phase 10 the implicit solve. */
int s3g_phase(s3g_ctx *c, int phase, int n, int istep, double thrs);

/* Basin energy budget of the turbulence closure, the TKE columns of
 * ScalarBalance.txt (OutScalarEnergyBalance, si3d_utils.f90:3698-3814, from
 * the sums SolveTKEandLengthScale keeps, si3d_mixing.f90:333-338, 466-475,
 * 539-548): out[0] TKinE (sum of q2 (rho+1000) h over the interfaces of the
 * latest trapezoidal call), out[1..3] ShearCum, BuocyCum, DisspCum (shear
 * and buoyancy production and dissipation, times dx dy dt, added on each
 * leapfrog call).  reset != 0 then clears the three integrals, as the
 * reference does after writing them.  Needs iturb >= 1 and one lake. */
int s3g_energy_balance(s3g_ctx *c, double *out, int reset);

int s3g_synchronize(s3g_ctx *c);
int s3g_get_stats(s3g_ctx *c, s3g_stats *out);
/* Average device duration (ms) of each timed kernel class since the last
 * reset, for roofline reporting: names[i] / avg_ms[i] / launches[i]. */
int s3g_kernel_times(s3g_ctx *c, int max, char (*names)[32], double *avg_ms, long *launches);
int s3g_reset_timers(s3g_ctx *c, int enable);
/* Tile words of the per-cell kernels (exmom x, exmom y, slayer_h, vel, exsal
 * fluxes, vel2, exsal right-hand side, turbulence interface terms, settrap,
 * save): layer-chunk size | grid order << 8, as chosen
 * by the autotuner at the first fd call (S3G_TUNE=0 in the environment keeps
 * the defaults; S3G_TILES="w0,w1,..." fixes them).  Returns the number
 * written. */
int s3g_get_tiles(s3g_ctx *c, int max, int *kc);
/* Streaming copy of `bytes` bytes (rounded down to 16) on the current device
 * with this library's own kernel (16-byte vector loads and stores,
 * grid-stride), `reps` timed passes after one untimed pass; *gbs receives the
 * read + write rate in GB/s.  No counterpart in the reference: it measures
 * the HBM rate a plain streaming kernel reaches on the device, the ceiling
 * against which the bench sets its kernels' measured traffic. */
int s3g_probe_stream(size_t bytes, int reps, double *gbs);
void s3g_finalize(s3g_ctx *c);
const char *s3g_last_error(void);

/* --- Open boundaries ------------------------------------------------------ */
/* Replaces openbc0 (si3d_boundaryconditions.f90:23-897) for boundary types 1
 * (water surface elevation) and 2 (free-surface flow); the library then runs
 * openbcUVH, MODexmom4openBCX/Y, MODqqddrr4openBC, MODvel4openBC,
 * MODexsal4openbc / MODexTracer4openbc and openbcSCA / openbcTracer inside
 * s3g_fd and shifts the boundary time levels in s3g_settrap / s3g_save.
 * Call once after the initial state is set and before the first s3g_fd.
 * The fields are the nopen block of si3d_inp.txt (si3d_utils.f90:123-144);
 * values holds the openbcNN.txt records: boundary nn, record r, value q at
 * (nn*npts + r)*nvar + q with q = 0 the elevation (m) or flow (m3/s,
 * positive toward +x / +y), 1 the active scalar, 2.. the tracers (nvar =
 * ntr + 2).  Types 3 (submerged flow) and 4 (nested grids) return
 * S3G_ERR_UNSUPPORTED; a cell that is not a wet edge cell of its side, or a
 * dry cell on an elevation boundary, S3G_ERR_ARG (the reference STOPs). */
typedef struct {
  int nopen;
  double dtsec;              /* dtsecOpenBC: seconds between records       */
  double dzmin;              /* dzmin of si3d_inp.txt                      */
  const int *iside;          /* [nopen] 1 W, 2 N, 3 E, 4 S                  */
  const int *itype;          /* [nopen] 1 elevation, 2 flow                 */
  const int *isbc, *jsbc, *iebc, *jebc; /* [nopen] first / last cell (i, j) */
  int npts, nvar;            /* records per boundary, values per record     */
  const double *values;      /* [nopen*npts*nvar]                           */
} s3g_openbc;
int s3g_set_open_boundaries(s3g_ctx *c, const s3g_openbc *bc);

/* --- Learned initial guess of the free-surface solve ---------------------- */
/* No counterpart in the reference (its solve is NSPCG, si3d_procedures.f90:
 * 1870-2065).  A U-Net maps the normalised residual of the plain initial
 * guess, r0/sigma (r0 = qq - A zeta_guess, sigma = rms r0), the coupling
 * field (rr-1)/rr and the wet mask to (zeta - zeta_guess)/sigma.  The CG
 * that follows still stops at the same FP64 true-residual tolerance. */

typedef struct {
  int pairs, capacity;   /* training pairs recorded / ring capacity          */
  long train_steps;      /* Adam steps taken                                  */
  double last_loss;      /* mean masked square error of the last step         */
  int height, width;     /* dense box of the network (multiples of 8)         */
  int mode;              /* 0 plain guess, 1 learned guess                    */
  long guess_solves;     /* solves with the learned guess since the mode was set */
  double guess_error_ratio; /* mean ||x - x_learned||_A / ||x - x_plain||_A over
                              those solves (x: the converged solution)        */
  long guess_rejected;   /* solves whose learned guess was not closer to the
                            solution in the energy norm than the plain guess
                            (the energy functional did not drop), which then
                            started from the plain guess                     */
  long corrections;      /* correction stages run since s3g_learn_flexible  */
  int flex_period;       /* the period set by s3g_learn_flexible (0: off)    */
} s3g_learn_stats;

/* Allocate the network (He-initialised from `seed`) and a device ring of
 * `capacity` training pairs; from now on every free-surface solve records
 * (r0/sigma, coupling, (zeta-guess)/sigma) while recording is on. */
int s3g_learn_setup(s3g_ctx *c, int capacity, unsigned seed);
int s3g_learn_record(s3g_ctx *c, int on);
/* `steps` Adam steps (learning rate lr) on minibatches of `batch` pairs drawn
 * from the ring, forward and backward on the GPU; *loss_out = mean loss of
 * the last step. */
int s3g_learn_train(s3g_ctx *c, int steps, int batch, double lr, double *loss_out);
/* 0: plain initial guess (zeta of the latest solve); 1: plus the network's
 * correction. */
int s3g_learn_mode(s3g_ctx *c, int mode);
/* Learned correction stage inside flexible CG (period > 0; 0 turns it off):
 * on iterations 0, period, 2 period, ... the preconditioned residual is
 * z = y + M^-1 (r - A y), where y is the network's estimate of A^-1 r (the
 * map it was trained on: normalised residual of an iterate -> its normalised
 * error) and M the chosen preconditioner; on the other iterations z = M^-1 r.
 * The search direction uses the Polak-Ribiere beta
 * ((z_i, r_i) - (z_i, r_{i-1})) / (z_{i-1}, r_{i-1}), which stays valid for a
 * preconditioner that changes between iterations, and the solve stops at the
 * same FP64 true-residual tolerance.  Needs a trained network
 * (s3g_learn_train); independent of s3g_learn_mode.  Replaces the CG
 * accelerator of NSPCG (nspcg.f:8328-8450 cgw, 8967-9092 itcg), which has no
 * such stage. */
int s3g_learn_flexible(s3g_ctx *c, int period);
long s3g_learn_num_weights(void);
int s3g_learn_get_weights(s3g_ctx *c, float *host, long n);
int s3g_learn_set_weights(s3g_ctx *c, const float *host, long n);
int s3g_learn_info(s3g_ctx *c, s3g_learn_stats *out);
/* Forward + backward of the network with the given weights on one dense
 * H x W input (x3: 3 planes of H*W), loss = sum mask (out-target)^2 / sum
 * mask; writes out (H*W), the weight gradients (same order as the weights)
 * and the loss.  For gradient checks. */
int s3g_unet_check(int H, int W, const float *weights, const float *x3, const float *target, const float *mask,
                   float *out, float *grads, double *loss);

/* Like s3g_unet_check, but allocates no device memory: no counterpart in the
 * reference; for callers that cannot allocate (code under HIP graph capture,
 * serving loops, harnesses that forbid hipMalloc).  The network, its
 * activations and gradients and the inputs live in a static device array of
 * 40 MiB that every GPU the library runs on holds until the process ends;
 * H * W may be at most 4096 cells (S3G_ERR_ARG otherwise, with the reason in
 * s3g_last_error).  Each call clears what it uses; calls that overlap on one
 * GPU (from several host threads) race on that array. */
int s3g_unet_check_static(int H, int W, const float *weights, const float *x3, const float *target,
                          const float *mask, float *out, float *grads, double *loss);

/* --- Standalone driver (no Fortran host) --------------------------------- */

/* Restates input/bathy/init on the host and calls s3g_init + s3g_set_field_*. */
int s3g_create_from_deck(const s3g_deck *deck, const s3g_params *params, s3g_ctx **out);

/* --- Ensemble of lakes ---------------------------------------------------- */
/* No counterpart in the reference (one lake per run).  nlakes independent
 * lakes of uneven size and layer count, each from its own deck and forcing
 * (params[i] may differ from params[0] only in ifsurfbc, cw, wa, phi,
 * dtsurfbc, n_surfbc, surfbc), advanced together: every phase is one launch
 * over the columns of all lakes, and the free-surface solve is one CG on the
 * block-diagonal system (its tolerance applies to the whole ensemble).
 * Fields use the same layout as for one lake with the lakes' columns one
 * after the other (lake i owns columns col_offset[i] .. col_offset[i+1]-1)
 * and km1 = the largest km1 of the lakes. */
int s3g_create_ensemble(int nlakes, const s3g_deck *decks, const s3g_params *params, s3g_ctx **out);
/* Number of lakes; with max >= that number also col_offset[0..nlakes] and
 * each lake's km1. */
int s3g_ensemble_info(s3g_ctx *c, int max, int *col_offset, int *km1);

#ifdef __cplusplus
}
#endif
#endif
