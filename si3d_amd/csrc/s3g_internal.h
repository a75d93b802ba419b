// s3g_internal.h -- device state and kernel entry points of libsi3dgpu.so.
//
// Device layout (HBM): every 3-D field is a stack of horizontal planes, one
// plane per layer k (k = 0 .. km1+2, reference numbering, planes 0 and km1+1,
// km1+2 are zero padding), each plane holding the wet columns in the
// reference's l order (i outer, j inner, si3d_utils.f90:781-793) at index
// c = l-1, followed by the dry sentinel column c = lm.  The plane stride LP is
// lm+1 rounded up to 256 columns.  With j running fastest, north/south
// neighbours are adjacent columns and east/west neighbours sit one row of wet
// cells away, so a wavefront of 64 consecutive columns reads every neighbour
// plane with coalesced runs.  One lane owns one column and walks it in k.
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>
#include <string>
#include <vector>

#include "../../include/si3d_gpu.h"

namespace s3g {

constexpr int kBlock = 256;

// Run parameters and derived constants (si3d_utils.f90:299-310).
struct Params {
  double dt, dx, dy, twodt, dtdx, dtdy, gdtdx, gdtdy, gdt2dx2, gdt2dy2, twodx, twody, fourdx,
      fourdy, dxdx, dydy, dxdy, beta2;
  int itrap, niter, ismooth, iturb, iadv, itrmom, ihd, ibc, isal, ifsurfbc;
  double beta, Av0, Dv0, Ax0, Ay0, f, theta, cd, cw, wa, phi, dtsurfbc;
  double t_1, t_2, t_3, t_4, t_5;  // turbulence closure constants (si3d_mixing.f90:641-645)
};

// Spatially uniform surface heat-budget forcing of one surfbc call (ifsurfbc
// 1-3, si3d_boundaryconditions.f90:4523-4598, 5160-5320): what does not depend
// on the column, evaluated on the host.
struct HeatForce {
  int mode;       // ifsurfbc
  double eta;     // light attenuation coefficient (1/m)
  double Qsw;     // shortwave flux at the surface (W/m2); < -100 means no data
  double Qn;      // net heat flux (mode 1)
  double Qbra;    // absorbed atmospheric longwave (modes 2, 3)
  double ea;      // air vapour pressure
  double Ta, Pa;  // air temperature (C), pressure (Pa)
  double ws;      // wind speed
  double coeffE, coeffH;  // bulk transfer coefficients (Kondo 1975)
};

// The met-station modes (ifsurfbc 10, 11) of one surfbc call: the stations'
// series values interpolated in time and, for the two-pass Barnes scheme,
// the first pass's residuals at the stations; per station q = 0 Qsw, 1 Ta,
// 2 Rh, 3 Cc|Qlw, 4 uair, 5 vair at [6*m + q].
constexpr int kMaxMet = 16;
struct MetArgs {
  int mode, nmet, iinterp, icw;
  double eta, Pa, gamma, cw;
  double st[6 * kMaxMet], res[6 * kMaxMet];
};

// Everything a kernel needs, passed by value.  R is the storage type of the
// 3-D fields, the turbulence state included (double, or float in the
// mixed-precision mode).  The free surface and the 2-D solver arrays are FP64
// in every mode; kernels compute in FP64 throughout.
template <typename R>
struct Dev {
  int lm, LP, km1, km, k1, KP;  // KP: planes per 3-D field
  Params P;
  double uair, vair, cdw;  // spatially uniform forcing (surfbc cases 0, 20)
  const int *lE, *lW, *lN, *lS, *kmz;
  // bottom layer of the u face east of each column (min(kmz, kmz east)) and of
  // the v face north of it; -1 where the face does not exist (dry neighbour)
  // and at the sentinel column
  const int *kmu, *kmv;
  int *k1z, *k1u, *k1v;
  const double *zlevel;  // [km1+3], zlevel[km1+1] = +huge; [nlakes][KPz] for an ensemble
  // ensemble of lakes (null for one lake): lake of each column (sentinel: 0),
  // each lake's km1, and its forcing (uair, vair, cdw) per lake
  const int *lake, *lake_km1;
  int KPz;
  const double *force;
  // per-column wind of the met-station modes: uair, vair, cdw at 3 c .. 3 c + 2
  const double *wind;
  const double *hhs, *hhu, *hhv;
  double *s, *sp, *spp, *sx, *sy, *dd, *qq, *rr, *eagx, *earx, *eagy, *eary;
  R *h, *hp, *hpp, *hu, *hup, *hupp, *hv, *hvp, *hvpp;
  R *uh, *uhp, *uhpp, *u, *up, *upp, *vh, *vhp, *vhpp, *v, *vp, *vpp, *wp;
  R *sal, *salp, *salpp, *rhop, *Av, *Dv, *kh, *ex, *ey, *agx, *arx, *agy, *ary;
  R *fx_ew, *fx_ns, *bex, *t0, *t1;
  // matmom's downward-pass factors c1 of the x and y systems (its d1, e1 go
  // to agx/arx, agy/ary); vel back-substitutes from them
  R *cfx, *cfy;
  R *tr, *trpp;  // ntr tracers, each a full plane stack, one after the other
  int ntr;
  R *heat;  // HeatSource (surface heat budget, ifsurfbc 1-3), null otherwise
  // turbulence state (q2, q2l and their time levels; REAL*8 in the reference)
  // and the closure's scratch planes, in the storage type like the other
  // column systems; d0, d1 hold 2 planes per level (the sweep's factor pairs)
  R *q2, *q2p, *q2pp, *q2l, *q2lp, *q2lpp, *d0, *d1, *d2;
  // basin energy budget of the closure (iturb >= 1, one lake): per-level
  // (dissipation rate x rho h, rho h) of the running call, in single
  // precision like the reference's rhoxh, per-workgroup partial sums, and the
  // budget state: the call's ShearProduction, BuoyancyProduction,
  // Dissipation, TKinE, then ShearCum, BuocyCum, DisspCum
  float2 *tkb;
  double *tke_part, *tke;
};

// One level of the aggregation multigrid preconditioner (levels >= 1; level 0
// is the lake's own wet-column grid).  Cells of level l+1 are 2x2 blocks of
// level-l cells in (i,j), numbered i outer / j inner like the fine grid.
struct MgLevel {
  int n = 0;                          // cells; index n is the sentinel
  int *lE = nullptr, *lW = nullptr, *lN = nullptr, *lS = nullptr;
  int *ch = nullptr;                  // [4n] children on the finer level (finer sentinel if absent)
  int *parent_of_fine = nullptr;      // [n_finer + 1] this level's cell of each finer cell
  void *rr = nullptr, *sx = nullptr, *sy = nullptr, *b = nullptr, *x = nullptr;  // T (double or float)
};

// Free-surface solver work space (FP64).
struct Pcg {
  int n, nblk;
  double *r, *r1, *z, *p0, *p1, *q, *diag, *x;  // r / r1: residual of even / odd iterations
  double *part;    // partial sums: [10][nblk]
  double *tol2 = nullptr;  // squared relative tolerance of the running CG
  double *scal;    // device scalars
  int *iters;      // device iteration counter
  float *rf, *rf1, *zf, *pf0, *pf1, *qf, *sxf, *syf, *rrf, *bf, *df;  // FP32 inner solve
  // line preconditioner: tridiagonal factors along N-S lines
  int nlines;
  int *line_start;  // [nlines+1] first column of each line (lines are contiguous in l)
  double *lc, *ld;  // modified super-diagonal and inverse pivots
  float *lcf, *ldf;
  std::vector<MgLevel> mg;  // mg[0] unused (the fine grid), mg[1..] coarse levels
  // ensembles: each lake's (r,r) and (b,b) of the true residual, so that every
  // lake meets the tolerance against its own right-hand side
  int *lake_off = nullptr;      // [nlakes+1] on the device
  double *lake_part = nullptr;  // [nlakes][kLakeSlices][2]
  double *h_lake = nullptr;     // pinned copy of lake_part
  // learned correction stage (flexible CG): the network's correction y and
  // r - A y, allocated on first use
  double *yc = nullptr, *ry = nullptr;
  float *ycf = nullptr, *ryf = nullptr;
};

// FP64 view of the free-surface system for the outer solver level (true
// residual, refinement, learned initial guess).
struct Outer {
  const int *lE, *lW, *lN, *lS;
  const double *rr, *sx, *sy, *b;
  double *x;
  int n, nblk;
  double *part;
};

struct Learner;  // s3g_unet.hip

struct Timer {
  std::string name;
  double total_ms = 0.0;
  long n = 0;
};

}  // namespace s3g

namespace s3g {
// Open boundaries (types 1 and 2).  The boundary transports and thicknesses
// keep the reference's storage (si3d_utils.f90:441-455): one row of km1+2
// layers per j for the west and east sides and per i for the north and south
// sides, at time levels n+1, n and n-1: array (side*3 + level) of bq / bh, row
// a at a*KB.
constexpr int kMaxOpen = 20;  // the nopen fields of si3d_inp.txt hold 20 values
struct OpenBc {
  int n = 0;
  std::vector<int> side, type;        // side 0 W, 1 N, 2 E, 3 S
  std::vector<int> first, count;      // cells of boundary nn: cells[first .. first+count)
  int np = 0, nv = 0;                 // records, values per record (2 + ntr)
  double dtsec = 0.0, dzmin = 0.0;
  std::vector<double> vals;           // [n][nv][np+1], 1-based series
  int KB = 0, NA = 0;                 // row length, rows per array
  int *cells = nullptr, *along = nullptr;  // column and row index of each boundary cell
  double *bq = nullptr, *bh = nullptr;     // [12][NA*KB]
  std::vector<double> hbc;                 // depth of the deepest cell of each section
  double *part = nullptr;                  // per-workgroup partial sums of a section's area
  bool any_flow = false;
};

// Everything an open-boundary kernel needs, by value; val and sc hold the
// boundary values of one call (elevation or flow, and the scalar).
struct ObcArgs {
  int n, KB, SZ;
  int side[kMaxOpen], type[kMaxOpen], first[kMaxOpen], count[kMaxOpen];
  double val[kMaxOpen], sc[kMaxOpen], hbc[kMaxOpen];
  const int *cells, *along;
  double *bq, *bh, *part;
  double dzmin;
};
}  // namespace s3g

namespace s3g {
// Per-cell kernels with a layer-chunk size; the first TILE_TUNED are autotuned.
enum TileId {
  TILE_EXMOM_X, TILE_EXMOM_Y, TILE_SLAYER_H, TILE_VEL, TILE_HFLUX, TILE_VEL2, TILE_RHS, TILE_TURB, TILE_TUNED,
  TILE_SETTRAP = TILE_TUNED, TILE_SAVE, TILE_NUM
};
}

struct s3g_ctx {
  int precision = S3G_FP64;
  int prec_kind = S3G_PREC_JACOBI;
  int im1 = 0, jm1 = 0, km1 = 0, km = 0, k1 = 2, lm = 0, LP = 0, KP = 0;
  long n_wet = 0;
  s3g::Params P{};
  std::vector<double> surfbc;  // 3*n records
  double pcg_tol = 1e-10;
  int pcg_maxit = 4000;
  double thrs = 0.0;
  double uair = 0.0, vair = 0.0, cdw = 0.0;
  // surface heat budget: set by surfbc when HeatSource must be recomputed
  bool heat_due = false;
  s3g::HeatForce heat_force{};
  // met stations (ifsurfbc 10, 11): positions, settings, the station-to-
  // station weights of the two-pass scheme (host) and the per-column
  // station weights [m][LP] and winds [3 LP] (device)
  std::vector<double> metxy, met_wn;
  int iinterp = 1, icw = 0;
  double gammaB = 0.0, delNfactor = 0.0;
  double *d_wmet = nullptr, *d_wind = nullptr;
  bool met_due = false;
  s3g::MetArgs met{};
  hipStream_t stream = nullptr;
  // device memory
  std::vector<void *> allocs;
  s3g::Dev<double> d64{};
  s3g::Dev<float> d32{};
  s3g::Pcg pcg{};
  std::vector<int> l2i, l2j;  // (i, j) of each wet column (host copy)
  // statistics
  s3g_stats stats{};
  // timers
  bool timing = false;
  bool debug = false;  // S3G_DEBUG set: per-solve diagnostics on stderr
  std::vector<s3g::Timer> timers;
  std::vector<std::pair<hipEvent_t, hipEvent_t>> ev_pool;
  std::vector<std::pair<int, int>> pending;  // (timer index, event pair index)
  int *h_iters = nullptr;                    // pinned
  double *h_scal = nullptr;                  // pinned
  // layer-chunk size of each per-cell kernel (autotuned at the first fd call
  // unless S3G_TUNE=0)
  int kc[s3g::TILE_NUM] = {1, 1, 1, 1, 1, 1, 1, 1, 1, 1};
  // exsal: right-hand side per cell + column solve (true), or fused into the
  // column sweep (false); chosen by the autotuner
  bool exsal_split = true;
  // multigrid: levels of at most mg_small cells run in one workgroup;
  // Jacobi sweeps on the coarsest level; coarse couplings scaled by mg_cs
  // (1 = Galerkin, 0.5 = the coarse grid's own discretisation; 0.55 took the
  // fewest iterations on the 1024x512 lake: 40.7 per step against 72.3 with
  // the Galerkin operator) (S3G_MG="small,sweeps[,cs]")
  int mg_small = 2048, mg_sweeps = 5;
  double mg_cs = 0.55;
  // initial guess of the leapfrog solve (istep 1): 0 = the latest zeta,
  // 1 = extrapolated from the two latest time levels (S3G_GUESS)
  int pcg_guess = 0;
  // CG iterations i >= 2 replayed in pairs from a captured graph of two
  // iterations (per storage type: [0] FP64, [1] FP32), rebuilt when the
  // kernel arguments change (S3G_GRAPH=0 launches every kernel instead)
  bool cg_graph = true;
  // iterations of the latest solve of each kind (0 leapfrog, 1.. trapezoidal
  // iteration): the first chunk of CG iterations is sized from the same kind
  int solve_slot = 0;
  int cg_hint[4] = {0, 0, 0, 0};
  hipGraphExec_t cg_exec[2] = {nullptr, nullptr};
  std::vector<char> cg_key[2];
  // The kernel sequences of a step around the free-surface solves -- fd before
  // and after the solve, settrap, settrap2, save -- captured once per kind
  // ([segment][istep - 1][lastiter]) and replayed as HIP graphs
  // (S3G_STEP_GRAPH=0 launches every kernel instead).  Not used with open
  // boundaries (their values are kernel arguments of each call) or while
  // kernel timers run.
  bool step_graph = true;
  hipGraphExec_t seg_exec[5][2][2] = {};
  long seg_replays = 0;
  bool tke_budget = true;   // accumulate the closure's basin energy budget (S3G_BUDGET=0: off)
  bool turb_split = false;  // turbulence closure: per-cell interface terms + sweep (else one fused sweep)
  bool tune = true, tuned = false;
  // windowed per-cell kernels (S3G_WINDOW=0 keeps the generic ones); every
  // wet neighbour points back (lW[lE[c]] = c, lS[lN[c]] = c, ...), which the
  // windowed exmom kernel relies on
  bool window = true, nb_symmetric = false;
  // ensemble of lakes (s3g_create_ensemble): column offset of each lake
  // ([nlakes+1]), each lake's km1 and forcing parameters
  int ntr = 0;                           // conservative tracers
  std::vector<std::string> tracer_names;  // "tracer1", "tracerpp1", ...
  int nlakes = 1;
  std::vector<int> lake_off, lake_km1;
  struct LakeForcing {
    int ifsurfbc = 0;
    double cw = 0, wa = 0, phi = 0, dtsurfbc = 0;
    std::vector<double> surfbc;  // 3 per record
  };
  std::vector<LakeForcing> lake_forcing;
  std::vector<double> lake_force;  // 3 per lake: uair, vair, cdw
  double *d_force = nullptr;
  // open boundaries (s3g_set_open_boundaries)
  s3g::OpenBc obc;
  // learned initial guess (s3g_learn_setup)
  s3g::Learner *learn = nullptr;
  std::vector<void *> learn_allocs;
};

namespace s3g {
// kernels.hip
template <typename R> int fd(s3g_ctx *c, int n, int istep, int lastiter);
template <typename R> int phase(s3g_ctx *c, int ph, int n, int istep);
template <typename R> int settrap(s3g_ctx *c);
template <typename R> int settrap2(s3g_ctx *c);
template <typename R> int save(s3g_ctx *c, int istep);
int init_density(s3g_ctx *c);
template <typename R> int autotune(s3g_ctx *c);
template <typename R> int momentum_solution(s3g_ctx *c, bool x, void *ag, void *ar);
// pcg.hip
template <typename R> int solve_free_surface(s3g_ctx *c, double tz);
// openbc.hip: the open-boundary steps of fd / settrap / save (no-ops without
// open boundaries); q selects the scalar series (1 active scalar, 2+ tracers)
template <typename R> void obc_uvh(s3g_ctx *c);
template <typename R> void obc_exmom(s3g_ctx *c, bool x, double twodt1);
template <typename R> void obc_qq(s3g_ctx *c, double tz);
template <typename R> void obc_vel(s3g_ctx *c);
template <typename R> void obc_exscalar(s3g_ctx *c, const R *cpp, int q);
template <typename R> void obc_scalar(s3g_ctx *c, R *cn, int q);
void obc_shift(s3g_ctx *c, int mode);
// unet.hip
int learn_before_solve(s3g_ctx *c, const void *outer, double *x);
int learn_after_solve(s3g_ctx *c, const void *outer);
// Learned correction stage of flexible CG (s3g_learn_flexible): the period of
// the stage for the current solve (0: off), and the network run on a CG
// residual r (double or float) whose (r,r) partials are rr_part[0..nblk-1].
struct LearnOut {
  const float *out;   // dense network output (normalised correction)
  const int *pix;     // wet column -> dense pixel
  const double *sig;  // device scalar: the residual's rms (the output's scale)
};
int learn_flex_period(const s3g_ctx *c);
int learn_correction(s3g_ctx *c, const void *r, bool fp32, const double *rr_part, int nblk, LearnOut *out);
void learn_destroy(s3g_ctx *c);
int pcg_setup(s3g_ctx *c, const int *lN0, const int *l2i, const int *l2j);
// host
void set_error(const std::string &msg);
int hip_check(hipError_t e, const char *what);
void timer_begin(s3g_ctx *c, const char *name, int *slot);
void timer_end(s3g_ctx *c, int slot);
void surfbc(s3g_ctx *c, int n, int istep);
// host.hip: context creation shared by s3g_init and s3g_create_ensemble;
// zl_lakes (nlakes x KP) and lake_of (lm+1) are null for one lake
int init_ctx(const s3g_grid *g, const s3g_params *p, int nlakes, const double *zl_lakes, const int *lake_of,
             const int *lake_km1, s3g_ctx **out);
int set_forcing(s3g_ctx *c);
}  // namespace s3g
