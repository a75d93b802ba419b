/*
 * si3d_oracle.h -- C interface of the CPU oracle for pSi3D's per-time-step
 * hydrodynamic core.
 *
 * TEST INFRASTRUCTURE ONLY.  The oracle is a plain C++ restatement of the
 * reference algorithm (si3d.f90 time loop, si3d_procedures.f90 fd/settrap/
 * settrap2/save and everything fd calls, si3d_mixing.f90, si3d_utils.f90
 * bathy/densty_s/parab).  It exists to check the HIP product library; the
 * product never links, loads or calls it.  Only tests, the smoke check and
 * the cpu_baseline leg of bench.py use it.
 *
 * Layout: the reference's own (k,l) layout, k fastest, 1-based indices,
 * column lm1 = lm+1 is the dry sentinel (si3d_utils.f90:757-806).
 */
#ifndef SI3D_ORACLE_H
#define SI3D_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* Input deck: the content of si3d_inp.txt, si3d_bathy, si3d_layer.txt,
 * si3d_init.txt and si3d_surfbc_W.txt / si3d_surfbc.txt for the options this
 * path supports. */
typedef struct {
  int imm, jmm;          /* interior cells in x / y (si3d_bathy header)        */
  int ibathyf;           /* >=0 uniform layers of idz, <0 si3d_layer.txt       */
  int km1_file;          /* number of zlevel records when ibathyf < 0          */
  double dx, dy, ddz, dzmin, datadj, zetainit, zl, dt;
  int itrap, niter, ismooth, iturb, iadv, itrmom, ihd, ibc, isal, ifsurfbc;
  double beta, Av0, Dv0, Ax0, Ay0, f, theta, cd, dtsurfbc, cw, wa, phi;
  const double *zlevel_file;  /* km1_file values (ibathyf < 0)                 */
  const int *bathy_dm;        /* imm*jmm depths in dm, index (i-2)*jmm+(j-2)   */
  const double *init_profile; /* km1 values: active scalar per layer k=1..km1  */
  int n_surfbc;               /* records of the surface series                 */
  const double *surfbc;       /* n_surfbc*nv: ifsurfbc 20: cdw uair vair;
                                 1: eta Qsw Qn cdw uair vair; 2: eta Qsw Ta Pa
                                 Rh Cc cdw uair vair; 3: eta Qsw Ta Pa Rh Qlw
                                 cdw uair vair                                 */
  double pcg_tol;             /* relative true-residual tolerance              */
  int pcg_maxit;
  int ntr;                    /* conservative tracers (ecomod = 0)             */
  const double *tracer_profiles; /* ntr*km1: profile of tracer nn at nn*km1+k-1 */
  int nopen;                  /* open boundaries (types 1 and 2)              */
  double dtsec_openbc;        /* seconds between boundary records             */
  const int *openbc;          /* nopen*6: iside itype isbc jsbc iebc jebc     */
  int openbc_npts;            /* records per boundary                         */
  const double *openbc_values;/* nopen*npts*(ntr+2): boundary nn, record r,
                                 value q at (nn*npts + r)*(ntr+2) + q; q = 0
                                 zeta (m) or flow (m3/s), 1 active scalar,
                                 2.. tracers (openbcNN.txt)                   */
  int icw;                    /* drag of the met-station modes: 0 cw, 1
                                 Amorocho & DeVries, 2 Andreas et al.         */
  int nmetstat;               /* met stations (ifsurfbc 10, 11); the series
                                 then holds 6*nmetstat+2 values per record:
                                 eta Pa, then per station Qsw Ta Rh Cc|Qlw
                                 uair vair                                    */
  const double *metxy;        /* 2*nmetstat: station x, y in grid units       */
  int iinterp;                /* 1 Barnes, 2 two-pass Barnes, 3 inverse
                                 distance                                     */
  double gammaB, delNfactor;  /* iinterp 2                                    */
} OrcDeck;

typedef struct OrcLake OrcLake;

OrcLake *orc_create(const OrcDeck *deck);
void orc_destroy(OrcLake *o);

/* grid sizes: out[0..7] = im, jm, km1, lm, k1, nts_unused, im1, jm1 */
void orc_dims(const OrcLake *o, int *out);

/* One full time step n (si3d.f90:324-393), or the individual calls. */
int orc_step(OrcLake *o, int n);
int orc_fd(OrcLake *o, int n, int istep, int lastiter, double thrs);
void orc_settrap(OrcLake *o);
void orc_settrap2(OrcLake *o);
void orc_save(OrcLake *o, int istep);

/* Individual phases of fd for phase-level parity; tz = 1/istep. */
int orc_phase(OrcLake *o, int phase, int n, int istep, double thrs);

/* Field access by name.  3-D fields are (km1, lm1) k-fastest (1-based k,l
 * mapped to 0-based storage), 2-D fields are (lm1). Integer fields are
 * returned as doubles.  Returns the element count or -1 if unknown. */
long orc_get(const OrcLake *o, const char *name, double *out);
long orc_set(OrcLake *o, const char *name, const double *in);

/* Statistics of the last free-surface solve: iterations, relative true
 * residual; and cumulative iteration count. */
void orc_solver_stats(const OrcLake *o, double *out3);

/* Unit-level entry points pinned by tests against closed forms. */
double orc_densty(double temperature, double salinity, double elevation);
double orc_tvd_flux(double vel, double s1, double s2, double s3, double s4, double cfdef, double courant);
void orc_trid(int n, const double *a, const double *b, const double *c,
              const double *g, const double *r, double *ag, double *ar);
void orc_trid1(int n, const double *a, const double *b, const double *c,
               const double *d, double *x);
double orc_parab(double frstpt, double x, const double *fx, int nfx, double dxs);

#ifdef __cplusplus
}
#endif
#endif
