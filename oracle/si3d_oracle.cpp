// si3d_oracle.cpp -- CPU oracle (FP64) for pSi3D's per-time-step hydrodynamic core.
//
// TEST INFRASTRUCTURE ONLY: a plain restatement of the reference algorithm used
// to check the HIP product library.  Nothing in the product links or calls it.
//
// Every routine cites the reference lines it restates.  The restatement keeps the
// reference's (k,l) layout (k fastest, 1-based, dry sentinel column lm1) and its
// global semantics: the OpenMP thread slabs and thread-private B* buffers of the
// reference are replaced by one sweep over all columns per phase, which is what
// those slabs compute.  Arrays the reference keeps across calls and only
// partially overwrites (eagx/earx/agx/arx for dried faces, the horizontal TVD
// fluxes of exsal) are kept persistent here too.
//
// Arithmetic is FP64 throughout, except the grid set-up (bathy, zlevel, initial
// layer thicknesses and top-layer indices), which is done in FP32 exactly as the
// reference computes it so that masks and indices are bit-exact.
#include "si3d_oracle.h"

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

namespace {

// si3d_types.f90:33-35, 216-226; si3d_types.f90:420 (sed_h)
const double G = 9.806, RHOAIR = 1.3, KAPPAS = 0.41, PI = 3.1415926535897932;
const double HMIN = 0.0, ZERO = 0.0, SED_H = 0.5;
const double AV_MOL = 1.3e-6, DV_MOL = 1.4e-7;
const double Q2_MIN = 2.e-6, Q2L_MIN = 6.80236723501e-9;
const double GH_MAX = 2.9e-2, GH_MIN = -2.8e-1;
const double A_1 = 0.92, A_2 = 0.74, B_1 = 16.6, B_2 = 10.1, E_1 = 1.80, E_2 = 1.33,
             E_3 = 0.25, C_2 = 0.70, C_3 = 0.20, SQ = 0.20;

struct F3 {
  int ks = 0;  // k stride = km1 + 3 (k = 0 .. km1+2)
  std::vector<double> v;
  void init(int km1, int lm1) {
    ks = km1 + 3;
    v.assign((size_t)ks * (lm1 + 1), 0.0);
  }
  double &operator()(int k, int l) { return v[(size_t)l * ks + k]; }
  double operator()(int k, int l) const { return v[(size_t)l * ks + k]; }
};

}  // namespace

// ---------------------------------------------------------------------------
// UNESCO/Gill EOS with the fixed-point pressure iteration, si3d_utils.f90:3865-3949.
// salinity**(3/2) is integer division (= salinity**1) in the reference (3902,3925,3935).
static double densty_s(double t, double s, double elev) {
  double t2 = t * t, t3 = t2 * t, t4 = t3 * t, t5 = t4 * t;
  double densw = 999.842594 + 6.793952e-2 * t - 9.095290e-3 * t2 + 1.001685e-4 * t3 -
                 1.120083e-6 * t4 + 6.536332e-9 * t5;
  double s32 = s;  // salinity**(3/2) with integer 3/2 = 1
  double densws = densw + s * (0.824493 - 4.0899e-3 * t + 7.6438e-5 * t2 - 8.2467e-7 * t3 +
                               5.3875e-9 * t4) +
                  s32 * (-5.72466e-3 + 1.0227e-4 * t - 1.6546e-6 * t2) + 4.8314e-4 * s * s;
  double rhoguess = densws, delta = 1.0, rho = densws;
  int iter = 0;
  const int maxiter = 10000;
  double kw = 19652.21 + 148.4206 * t - 2.327105 * t2 + 1.360477e-2 * t3 - 5.155288e-5 * t4;
  double ks = kw + s * (54.6746 - 0.603459 * t + 1.09987e-2 * t2 - 6.1670e-5 * t3) +
              s32 * (7.944e-2 + 1.6483e-2 * t - 5.3009e-4 * t2);
  while (delta > 1e-6 && iter < maxiter) {
    double p = 1e-5 * (rhoguess * 9.806 * elev);
    double k = ks + p * (3.239908 + 1.43713e-3 * t + 1.16092e-4 * t2 - 5.77905e-7 * t3) +
               p * s * (2.2838e-3 - 1.0981e-5 * t - 1.6078e-6 * t2) + 1.91075e-4 * p * s32 +
               p * p * (8.50935e-5 - 6.12293e-6 * t + 5.2787e-8 * t2) +
               p * p * s * (-9.9348e-7 + 2.0816e-8 * t + 9.1697e-10 * t2);
    rho = densws / (1.0 - p / k);
    delta = std::fabs(rho - rhoguess);
    rhoguess = rho;
    ++iter;
  }
  return rho;
}

// Double-sweep solver with two right-hand sides, si3d_procedures.f90:3210-3289 (trid).
static void trid2(int n, const double *a, const double *b, const double *c, const double *d,
                  const double *e, double *x, double *y) {
  std::vector<double> c1(n + 1), d1(n + 1), e1(n + 1);
  c1[0] = c[0] / b[0];
  d1[0] = d[0] / b[0];
  e1[0] = e[0] / b[0];
  for (int k = 1; k < n; ++k) {
    double den = b[k] - a[k] * c1[k - 1];
    c1[k] = c[k] / den;
    d1[k] = (d[k] - a[k] * d1[k - 1]) / den;
    e1[k] = (e[k] - a[k] * e1[k - 1]) / den;
  }
  x[n - 1] = d1[n - 1];
  y[n - 1] = e1[n - 1];
  for (int k = n - 2; k >= 0; --k) {
    x[k] = d1[k] - c1[k] * x[k + 1];
    y[k] = e1[k] - c1[k] * y[k + 1];
  }
}

// e/f double sweep, si3d_procedures.f90:3292-3352 (trid1) and si3d_mixing.f90:554-612 (tridT).
static void trid_ef(int n, const double *a, const double *b, const double *c, const double *d,
                    double *x) {
  std::vector<double> e(n + 1), f(n + 1);
  e[0] = -c[0] / b[0];
  f[0] = d[0] / b[0];
  for (int k = 1; k < n - 1; ++k) {
    double den = b[k] + a[k] * e[k - 1];
    e[k] = -c[k] / den;
    f[k] = (d[k] - a[k] * f[k - 1]) / den;
  }
  x[n - 1] = (d[n - 1] - a[n - 1] * f[n - 2]) / (b[n - 1] + a[n - 1] * e[n - 2]);
  for (int k = n - 2; k >= 0; --k) x[k] = e[k] * x[k + 1] + f[k];
}

// Parabolic interpolation in a uniformly spaced series, si3d_utils.f90:895-920.
static double parab(double frstpt, double x, const double *fx /*1-based*/, double dx) {
  int m = (int)((x - frstpt) / dx);  // INTEGER assignment truncates
  double om = m;
  double theta = (x - frstpt - om * dx) / dx;
  if (m == 0) {
    m = 2;
    theta = theta - 1.0;
  } else {
    m = m + 1;
  }
  return fx[m] + 0.5 * theta * (fx[m + 1] - fx[m - 1] + theta * (fx[m + 1] + fx[m - 1] - 2.0 * fx[m]));
}

struct OrcLake {
  // basin energy budget of SolveTKEandLengthScale (si3d_mixing.f90:465-474,
  // 536-548): tke[0..3] = the per-call sums ShearProduction,
  // BuoyancyProduction, Dissipation, TKinE (one thread's private values, i.e.
  // the whole basin for a single-threaded run); tke[4..6] = ShearCum,
  // BuocyCum, DisspCum, cleared by the writer of ScalarBalance.txt
  double tke[7] = {0, 0, 0, 0, 0, 0, 0};
  // ---- parameters (si3d_types.f90:67-90, derived in si3d_utils.f90:299-310)
  OrcDeck d{};
  std::vector<double> zfile, profile, sbc;
  int im, im1, jm, jm1, km, km1, k1, lm, lm1;
  double dt, dx, dy, twodt, dtdx, dtdy, gdtdx, gdtdy, gdt2dx2, gdt2dy2, twodx, twody, fourdx,
      fourdy, dxdx, dydy, dxdy, beta2;
  double t_1, t_2, t_3, t_4, t_5;
  double tz = 1.0, thrs = 0.0;
  // ---- grid
  std::vector<char> mask2d;  // (im1+2)*(jm1+2), 1-based
  std::vector<int> ij2l, l2i, l2j, lEC, lWC, lNC, lSC, kmz, k1z, k1u, k1v;
  std::vector<double> zlevel, hhs, hhu, hhv;
  // ---- 2-D state
  std::vector<double> s, sp, spp, sx, sy, dd, qq, rr, eagx, earx, eagy, eary, uair, vair, cdw;
  // ---- 3-D state
  F3 h, hp, hpp, hu, hup, hupp, hv, hvp, hvpp;
  F3 uh, uhp, uhpp, u, up, upp, vh, vhp, vhpp, v, vp, vpp, wp;
  F3 sal, salp, salpp, rhop, Av, Dv, kh;
  F3 ex, agx, arx, agy, ary, haxpp, haypp, th, th1;
  F3 fx_ew, fx_ns, fx_ud, bex;  // exsal fluxes (Bth2, Bth4, Bth3) and Bex
  F3 q2, q2p, q2pp, q2l, q2lp, q2lpp;
  F3 heat, qswfr;  // HeatSource and QswFr (surface heat budget, ifsurfbc 1-3)
  // ---- conservative tracers (ecomod = 0): tracer and tracerpp per tracer
  int ntr = 0;
  std::vector<double> tprof;  // ntr*km1 initial profiles
  std::vector<F3> tracer, tracerpp;
  // ---- open boundaries, types 1 (water surface elevation) and 2 (free-surface
  // flow) (si3d_boundaryconditions.f90).  The boundary transports and
  // thicknesses keep the reference's storage (si3d_utils.f90:441-455):
  // uhWB/huWB, uhEB/huEB over (k, j) and vhNB/hvNB, vhSB/hvSB over (k, i),
  // each at n+1, n and n-1; here bq/bh[side][level] with side = iside-1
  // (0 W, 1 N, 2 E, 3 S) and level 0 = n+1, 1 = n, 2 = n-1.
  struct Obc {
    int side, type, is, js, ie, je;
  };
  std::vector<Obc> obc;
  int obc_np = 0, obc_nv = 0;
  double obc_dtsec = 0.0;
  std::vector<double> obc_vals;  // [nopen][nv] series of obc_np values, 1-based
  std::vector<double> bq[4][3], bh[4][3];
  double &BQ(int sd, int lev, int k, int a) { return bq[sd][lev][(size_t)a * (km1 + 2) + k]; }
  double &BH(int sd, int lev, int k, int a) { return bh[sd][lev][(size_t)a * (km1 + 2) + k]; }
  // ---- solver statistics
  int last_iters = 0;
  double last_relres = 0.0;
  long total_iters = 0;

  int &IJ(int i, int j) { return ij2l[(size_t)i * (jm1 + 2) + j]; }
  bool M2(int i, int j) const { return mask2d[(size_t)i * (jm1 + 2) + j] != 0; }
  bool maskE(int l) const { return M2(l2i[l] + 1, l2j[l]); }
  bool maskN(int l) const { return M2(l2i[l], l2j[l] + 1); }

  bool build(const OrcDeck *deck);
  void init_state();
  double zdens(int k, int l) const {  // depth used for densty_s, si3d_procedures.f90:3477-3481
    return (zlevel[k] == -100.0) ? 0.5 * hp(k, l) : zlevel[k] + 0.5 * hp(k, l);
  }

  // phases of fd
  void surfbc(int n, int istep);
  int sbc_nvar() const {
    return d.ifsurfbc == 20 ? 3 : d.ifsurfbc == 1 ? 6 : (d.ifsurfbc == 2 || d.ifsurfbc == 3) ? 9
           : (d.ifsurfbc == 10 || d.ifsurfbc == 11) ? 6 * d.nmetstat + 2 : 0;
  }
  // met stations (ifsurfbc 10, 11): positions, per-column weights w[m][l]
  // and the station-to-station weights of the two-pass Barnes scheme
  std::vector<double> metxy, wst, wn;
  void met_weights();
  void met_distribute(const double *v, bool heat_too);
  void surface_heat(int l, double Qsw, double Qbra, double ea, double Ta, double Pa, double ws);
  double solar_fr(double eta, double depth) const;
  void distribute_qsw(double eta);
  void distribute_heat(const double *v);
  void hmix();
  void exmom(int ieq);
  void matmom(int ieq);
  void bclnc_km(int l, int kb, int ieq, double &huvpdrho);
  void matcon();
  int solver();
  void slayer_h();
  void layer_h();
  void layer_hp2();
  void toplayer_indexp2();
  void vel();
  void continuity(int ist);
  void exsal();
  void exscalar(const F3 &cpp, bool is_tracer);
  void imsal();
  void imtracer(int nt);
  void vel2();
  void smooth();
  void turbulence(int istep);
  int fd(int n, int istep, int lastiter);
  // open boundaries
  double obc_value(int nn, int q) const;
  bool openbc0();
  void openbc_uvh();
  void flow_section(const Obc &b, double qbc, bool first);
  void mod_qq();
  void mod_exmom(int ieq);
  void mod_vel();
  void openbc_scalar(F3 &c, int q);
  void mod_exscalar(const F3 &cpp, int q);
  void obc_shift(int mode);
  void settrap();
  void settrap2();
  void save(int istep);
};

// ---------------------------------------------------------------------------
// input (si3d_utils.f90:299-310), ZGridDimensions (318-382), bathy (595-892).
bool OrcLake::build(const OrcDeck *deck) {
  d = *deck;
  if (deck->ibathyf < 0 && deck->zlevel_file)
    zfile.assign(deck->zlevel_file, deck->zlevel_file + deck->km1_file);
  d.zlevel_file = nullptr;
  if (deck->n_surfbc > 0 && deck->surfbc) sbc.assign(deck->surfbc, deck->surfbc + (size_t)sbc_nvar() * deck->n_surfbc);
  d.surfbc = nullptr;
  if ((d.ifsurfbc == 10 || d.ifsurfbc == 11) && deck->metxy) metxy.assign(deck->metxy, deck->metxy + 2 * (size_t)d.nmetstat);
  d.metxy = nullptr;

  dt = d.dt;
  dx = d.dx;
  dy = d.dy;
  twodt = 2.0 * dt;
  dtdx = dt / dx;
  dtdy = dt / dy;
  gdtdx = G * dtdx;
  gdtdy = G * dtdy;
  gdt2dx2 = gdtdx * dtdx;
  gdt2dy2 = gdtdy * dtdy;
  twodx = 2.0 * dx;
  twody = 2.0 * dy;
  fourdx = 2.0 * twodx;
  fourdy = 2.0 * twody;
  dxdx = dx * dx;
  dydy = dy * dy;
  dxdy = dx * dy;
  beta2 = d.beta / 2.0;
  im = d.imm + 1;
  im1 = im + 1;
  jm = d.jmm + 1;
  jm1 = jm + 1;
  k1 = 2;

  // ZGridDimensions: zlevel is REAL (FP32) in the reference.
  std::vector<float> zl32;
  if (d.ibathyf < 0) {
    km1 = d.km1_file;
    km = km1 - 1;
    for (int k = 0; k < km1; ++k) zl32.push_back((float)zfile[k]);
  } else {
    float zlf = (float)d.zl, ddzf = (float)d.ddz;
    km = (int)std::ceil(zlf / ddzf) + 1;
    km1 = km + 1;
    zl32.assign(km1, 0.0f);
    zl32[k1 - 1] = 0.0f;
    for (int k = k1 + 1; k <= km1; ++k) zl32[k - 1] = zl32[k - 2] + ddzf;
    zl32[k1 - 1] = -100.0f;
    zl32[0] = -100.0f;
  }
  zlevel.assign(km1 + 3, 1e30);
  for (int k = 1; k <= km1; ++k) zlevel[k] = zl32[k - 1];

  // bathy: h4 = dm*0.1 + datadj in FP32, boundary ring set to 0 (667-682)
  std::vector<float> h4((size_t)(im1 + 2) * (jm1 + 2), 0.0f);
  auto H4 = [&](int i, int j) -> float & { return h4[(size_t)i * (jm1 + 2) + j]; };
  for (int i = 2; i <= im; ++i)
    for (int j = 2; j <= jm; ++j)
      H4(i, j) = (float)deck->bathy_dm[(size_t)(i - 2) * d.jmm + (j - 2)] * 0.1f + (float)d.datadj;
  for (int j = 1; j <= jm1; ++j) H4(1, j) = 0.0f, H4(im1, j) = 0.0f;
  for (int i = 1; i <= im1; ++i) H4(i, 1) = 0.0f, H4(i, jm1) = 0.0f;
  mask2d.assign((size_t)(im1 + 2) * (jm1 + 2), 0);
  for (int i = 1; i <= im1; ++i)
    for (int j = 1; j <= jm1; ++j) mask2d[(size_t)i * (jm1 + 2) + j] = H4(i, j) > 0.0f;
  // l numbering: i outer, j inner (757-793)
  lm = 0;
  for (int i = 1; i <= im1; ++i)
    for (int j = 1; j <= jm1; ++j) lm += M2(i, j);
  lm1 = lm + 1;
  ij2l.assign((size_t)(im1 + 2) * (jm1 + 2), lm1);
  l2i.assign(lm1 + 1, 0);
  l2j.assign(lm1 + 1, 0);
  int l = 0;
  for (int i = 1; i <= im1; ++i)
    for (int j = 1; j <= jm1; ++j)
      if (M2(i, j)) {
        ++l;
        l2i[l] = i;
        l2j[l] = j;
        IJ(i, j) = l;
      }
  // neighbours (796-806); the sentinel's own neighbours are the sentinel.
  lEC.assign(lm1 + 1, lm1);
  lWC.assign(lm1 + 1, lm1);
  lNC.assign(lm1 + 1, lm1);
  lSC.assign(lm1 + 1, lm1);
  for (l = 1; l <= lm; ++l) {
    int i = l2i[l], j = l2j[l];
    lEC[l] = IJ(i + 1, j);
    lWC[l] = IJ(i - 1, j);
    lNC[l] = IJ(i, j + 1);
    lSC[l] = IJ(i, j - 1);
  }
  // kmz, hhs (812-867), FP32 arithmetic
  kmz.assign(lm1 + 1, 0);
  std::vector<float> hhs32(lm1 + 1, 0.0f);
  float ddzf = (float)d.ddz, dzminf = (float)d.dzmin;
  for (l = 1; l <= lm; ++l) {
    float hs1 = H4(l2i[l], l2j[l]);
    if (d.ibathyf >= 0) {
      int kz = (int)std::floor(hs1 / ddzf);
      if ((hs1 - (float)kz * ddzf) > dzminf) {
        kz = kz + 1;
        hhs32[l] = hs1;
      } else {
        hhs32[l] = (float)kz * ddzf;
      }
      kmz[l] = kz + 1;
    } else {
      for (int k = k1; k <= km; ++k) {
        if (zl32[k] >= hs1) {  // zlevel(k+1)
          hhs32[l] = zl32[k - 1] + std::max(dzminf, hs1 - zl32[k - 1]);
          kmz[l] = k;
          break;
        }
      }
      if (kmz[l] == 0) {
        std::fprintf(stderr, "oracle: column %d deeper than the deepest layer\n", l);
        return false;
      }
    }
  }
  kmz[lm1] = km1;
  std::vector<float> hhu32(lm1 + 1, 0.0f), hhv32(lm1 + 1, 0.0f);
  for (int j = 1; j <= jm; ++j)
    for (int i = 1; i <= im; ++i) {
      if (M2(i + 1, j) && M2(i, j)) {
        int ll = IJ(i, j);
        hhu32[ll] = std::min(hhs32[lEC[ll]], hhs32[ll]);
      }
      if (M2(i, j + 1) && M2(i, j)) {
        int ll = IJ(i, j);
        hhv32[ll] = std::min(hhs32[lNC[ll]], hhs32[ll]);
      }
    }
  hhs.assign(lm1 + 1, 0.0);
  hhu.assign(lm1 + 1, 0.0);
  hhv.assign(lm1 + 1, 0.0);
  for (l = 0; l <= lm1; ++l) hhs[l] = hhs32[l], hhu[l] = hhu32[l], hhv[l] = hhv32[l];

  // ---- init (si3d_procedures.f90:23-181): FP32 layer thicknesses and top indices
  F3 *all3[] = {&h,   &hp,   &hpp,  &hu,   &hup,  &hupp, &hv,   &hvp,  &hvpp,   &uh,   &uhp,
                &uhpp, &u,   &up,   &upp,  &vh,   &vhp,  &vhpp, &v,    &vp,     &vpp,  &wp,
                &sal,  &salp, &salpp, &rhop, &Av,  &Dv,   &kh,   &ex,   &agx,    &arx,  &agy,
                &ary,  &haxpp, &haypp, &th, &th1, &fx_ew, &fx_ns, &fx_ud, &bex, &q2, &q2p,
                &q2pp, &q2l,  &q2lp, &q2lpp, &heat, &qswfr};
  for (F3 *f : all3) f->init(km1, lm1);
  std::vector<double> *all2[] = {&s,    &sp,   &spp,  &sx,   &sy,   &dd,   &qq,  &rr,
                                 &eagx, &earx, &eagy, &eary, &uair, &vair, &cdw};
  for (auto *a : all2) a->assign(lm1 + 1, 0.0);
  float zinit = (float)d.zetainit;
  std::vector<float> s32(lm1 + 1, zinit);
  k1z.assign(lm1 + 1, km1);
  k1u.assign(lm1 + 1, km1);
  k1v.assign(lm1 + 1, km1);
  for (l = 1; l <= lm; ++l) {
    int i = l2i[l], j = l2j[l], kms = kmz[l], nw = 0;
    for (int k = k1; k <= kms; ++k) {
      float hv32 = std::min(zl32[k], hhs32[l]) - std::max(zl32[k - 1], -s32[l]);
      if (hv32 > 0.0f) {
        if (++nw == 1) k1z[l] = k;
        hp(k, l) = hv32;
      } else {
        hp(k, l) = 0.0;
      }
    }
    hp(kms + 1, l) = SED_H;
    if (k1z[l] == km1) s32[l] = -hhs32[l] + 0.0f;
    if (M2(i + 1, j)) {
      int kmx = std::min(kmz[l], kmz[lEC[l]]);
      nw = 0;
      for (int k = k1; k <= kmx; ++k) {
        float hv32 = std::min(zl32[k], hhu32[l]) - std::max(zl32[k - 1], -(s32[l] + s32[lEC[l]]) / 2.0f);
        if (hv32 > 0.0f) {
          if (++nw == 1) k1u[l] = k;
          hup(k, l) = hv32;
        } else
          hup(k, l) = 0.0;
      }
    }
    if (M2(i, j + 1)) {
      int kmy = std::min(kmz[l], kmz[lNC[l]]);
      nw = 0;
      for (int k = k1; k <= kmy; ++k) {
        float hv32 = std::min(zl32[k], hhv32[l]) - std::max(zl32[k - 1], -(s32[l] + s32[lNC[l]]) / 2.0f);
        if (hv32 > 0.0f) {
          if (++nw == 1) k1v[l] = k;
          hvp(k, l) = hv32;
        } else
          hvp(k, l) = 0.0;
      }
    }
  }
  for (l = 0; l <= lm1; ++l) s[l] = sp[l] = spp[l] = s32[l];
  hupp = hup;
  hu = hup;
  hvpp = hvp;
  hv = hvp;
  hpp = hp;
  h = hp;
  for (l = 0; l <= lm1; ++l) rr[l] = 1.0;
  init_state();
  return openbc0();
}

// InitializeScalarFields (si3d_procedures.f90:184-379, InitProc=4) and
// InitializeTurbulenceModel (si3d_mixing.f90:615-724) and surfbc0 (si3d_boundaryconditions.f90:4140-4487).
void OrcLake::init_state() {
  profile.assign(km1 + 1, 0.0);
  for (int k = 1; k <= km1; ++k) profile[k] = d.init_profile ? d.init_profile[k - 1] : 0.0;
  d.init_profile = nullptr;
  d.bathy_dm = nullptr;
  for (int l = 1; l <= lm1; ++l)
    for (int k = 1; k <= km1; ++k) salp(k, l) = profile[k];
  sal = salp;
  salpp = salp;
  // tracers (si3d_procedures.f90:330-356, ecomod = 0): profile column nn+1 of si3d_init.txt
  ntr = d.tracer_profiles ? std::max(0, d.ntr) : 0;
  tprof.assign(d.tracer_profiles, d.tracer_profiles + (size_t)ntr * km1);
  d.tracer_profiles = nullptr;
  tracer.assign(ntr, F3());
  for (int nt = 0; nt < ntr; ++nt) {
    tracer[nt].init(km1, lm1);
    for (int l = 1; l <= lm1; ++l)
      for (int k = 1; k <= km1; ++k) tracer[nt](k, l) = tprof[(size_t)nt * km1 + (k - 1)];
  }
  tracerpp = tracer;
  for (int l = 1; l <= lm1; ++l)
    for (int k = k1; k <= km1; ++k) rhop(k, l) = densty_s(salp(k, l), 0.00004, zdens(k, l)) - 1000.0;

  if (d.iturb == 0) {
    for (int l = 1; l <= lm; ++l) {
      int kms = kmz[l], k1s = k1z[l];
      Av(k1s, l) = 0.0;
      Dv(k1s, l) = 0.0;
      for (int k = k1s + 1; k <= kms; ++k) Av(k, l) = d.Av0, Dv(k, l) = d.Dv0;
      Av(kms + 1, l) = 0.0;
      Dv(kms + 1, l) = 0.0;
    }
  } else {
    t_1 = A_2 * (1. - 6. * A_1 / B_1);
    t_2 = 3. * A_2 * (6. * A_1 + B_2 * (1. - C_3));
    t_3 = std::pow(B_1, -1. / 3.);
    t_4 = 9. * A_1 * (2. * A_1 + A_2 * (1. - C_2));
    t_5 = 9. * A_1 * A_2;
    for (auto &x : q2.v) x = Q2_MIN;
    q2p = q2;
    q2pp = q2;
    std::vector<double> zfromt(km1 + 2, 0.0);
    for (int m = 1; m <= lm; ++m) {
      int kms = kmz[m], k1s = k1z[m];
      double htot = 0.0;
      for (int k = 1; k <= km1; ++k) htot += hp(k, m);  // SUM(hp(:,m)) includes the sediment cell
      if (kms > k1s) {
        double z0 = 0.0;
        for (int k = k1s + 1; k <= kms; ++k) {
          z0 += hp(k - 1, m);
          zfromt[k] = z0;
        }
        for (int k = k1s + 1; k <= kms; ++k) {
          double mixl = KAPPAS * zfromt[k] * std::sqrt(std::max(1. - (zfromt[k] / htot), 0.0));
          q2l(k, m) = q2(k, m) * mixl;
        }
        q2l(k1s, m) = 0.0;
        q2l(kms + 1, m) = 0.0;
      }
    }
    q2lp = q2l;
    q2lpp = q2l;
  }
  const int nv = sbc_nvar();
  if (d.ifsurfbc == 10 || d.ifsurfbc == 11) {  // surfbc0 cases 10, 11: weights, then the first record's wind (4301-4450)
    met_weights();
    if ((int)sbc.size() >= nv) met_distribute(sbc.data(), false);
  } else if (nv > 0 && (int)sbc.size() >= nv) {  // momentum flux terms at t = 0 (surfbc0)
    for (int l = 0; l <= lm1; ++l) cdw[l] = sbc[nv - 3], uair[l] = sbc[nv - 2], vair[l] = sbc[nv - 1];
  }
}

static double sat_vapor_pressure(double T);
static double latent_heat_vaporization(double T);

// InitializeInterpolationMethods, si3d_boundaryconditions.f90:5510-5646: the
// weight of each met station at each wet column (i, j) -- 1: Barnes,
// exp(-ln(100) d^2 / da^2) normalised, da the column's mean distance to the
// stations; 2: two-pass Barnes (Koch et al. 1983), exp(-d^2 / kappa) with
// kappa from the mean nearest-station distance, gammaB and delNfactor, not
// normalised, plus the same weights between stations (at the cell nearest to
// each station); 3: inverse squared distance (1 at the station's own cell),
// normalised.  Distances in grid units.
void OrcLake::met_weights() {
  const int nm = d.nmetstat;
  wst.assign((size_t)nm * (lm1 + 1), 0.0);
  wn.assign((size_t)nm * nm, 0.0);
  auto dist2 = [&](int m, double x, double y) {
    const double xi = metxy[2 * m] - x, yi = metxy[2 * m + 1] - y;
    return xi * xi + yi * yi;
  };
  double kappa1 = 0.0;
  if (d.iinterp == 2) {
    double da = 0.0;
    for (int m = 0; m < nm; ++m) {
      double mn = HUGE_VAL;  // MINVAL of no other station: the largest value
      for (int q = 0; q < nm; ++q)
        if (q != m) mn = std::min(mn, std::sqrt(dist2(q, metxy[2 * m], metxy[2 * m + 1])));
      da += mn;
    }
    da /= nm;
    const double g = d.gammaB;
    const double f = g <= 0.2 ? 5.0515 : g <= 0.3 ? 3.5132 : g <= 0.5 ? 2.3838 : g <= 0.6 ? 2.1153 : 1.7826;
    kappa1 = f * 4.0 / (PI * PI) * (d.delNfactor * da) * (d.delNfactor * da);
    for (int m = 0; m < nm; ++m) {
      const double x = std::nearbyint(metxy[2 * m]), y = std::nearbyint(metxy[2 * m + 1]);
      for (int q = 0; q < nm; ++q) wn[(size_t)m * nm + q] = std::exp(-dist2(q, x, y) / kappa1);
    }
  }
  for (int l = 1; l <= lm; ++l) {
    const double x = l2i[l], y = l2j[l];
    auto w = [&](int m) -> double & { return wst[(size_t)m * (lm1 + 1) + l]; };
    if (d.iinterp == 1) {
      double da = 0.0;
      for (int m = 0; m < nm; ++m) da += std::sqrt(dist2(m, x, y));
      da /= nm;
      double sum = 0.0;
      for (int m = 0; m < nm; ++m) sum += (w(m) = std::exp(-4.60517018598809 * dist2(m, x, y) / (da * da)));
      for (int m = 0; m < nm; ++m) w(m) /= sum;
    } else if (d.iinterp == 2) {
      for (int m = 0; m < nm; ++m) w(m) = std::exp(-dist2(m, x, y) / kappa1);
    } else {
      double sum = 0.0;
      for (int m = 0; m < nm; ++m) {
        const double di = dist2(m, x, y);
        sum += (w(m) = di > 0.0 ? 1.0 / di : 1.0);
      }
      for (int m = 0; m < nm; ++m) w(m) /= sum;
    }
  }
}

// The met-station modes on one record of interpolated series values v (eta,
// Pa, then per station Qsw Ta Rh Cc|Qlw uair vair): CalculateMetResiduals
// (5650-5700), InterpMetData (5703-5748) and the cases 10 / 11 of
// DistributeMomentumHeatSourcesH (5331-5483) -- per wet column the station
// values interpolated (the two-pass scheme adds the first pass's residuals at
// the stations, spread with the weights to the power 1/gammaB), the column's
// wind, its drag (icw) and, with heat_too, its heat sources.  Mode 10 reads
// the cloud cover of the module variable Cc, which nothing sets in that mode
// (0), and interpolates no longwave (the reference sums the station longwave
// array there, which mode 10 never allocates, into a value it does not use).
void OrcLake::met_distribute(const double *v, bool heat_too) {
  const int nm = d.nmetstat, mode = d.ifsurfbc;
  const double eta = v[0], Pa = v[1];
  auto st = [&](int m, int q) { return v[2 + 6 * m + q]; };  // q: 0 Qsw 1 Ta 2 Rh 3 Cc|Qlw 4 uair 5 vair
  std::vector<double> resid((size_t)6 * nm, 0.0);
  if (d.iinterp == 2) {
    for (int k = 0; k < nm; ++k) {
      double sumw = 0.0;
      for (int m = 0; m < nm; ++m) sumw += wn[(size_t)k * nm + m];
      for (int q = 0; q < 6; ++q) {
        if (q == 3 && mode == 10) continue;
        double dum = 0.0;
        for (int m = 0; m < nm; ++m) dum += wn[(size_t)k * nm + m] / sumw * st(m, q);
        resid[(size_t)6 * k + q] = st(k, q) - dum;
      }
    }
  }
  if (heat_too) {
    distribute_qsw(eta);
    for (int l = 1; l <= lm; ++l)
      for (int k = 1; k <= km1; ++k) heat(k, l) = 0.0;
  }
  const double EmAirC = 0.642, StephanBoltzman = 5.6697e-8, Al = 0.03;
  for (int l = 1; l <= lm; ++l) {
    double x[6] = {0, 0, 0, 0, 0, 0};
    auto w = [&](int m) { return wst[(size_t)m * (lm1 + 1) + l]; };
    for (int m = 0; m < nm; ++m)
      for (int q = 0; q < 6; ++q) x[q] += w(m) * st(m, q);
    if (d.iinterp == 2) {
      double sumw = 0.0, sumg = 0.0;
      for (int m = 0; m < nm; ++m) sumw += w(m), sumg += std::pow(w(m), 1.0 / d.gammaB);
      for (int q = 0; q < 6; ++q) x[q] /= sumw;
      for (int m = 0; m < nm; ++m)
        for (int q = 0; q < 6; ++q) x[q] += std::pow(w(m), 1.0 / d.gammaB) / sumg * resid[(size_t)6 * m + q];
    }
    uair[l] = x[4];
    vair[l] = x[5];
    const double ws = std::sqrt(std::pow(uair[l], 2.) + std::pow(vair[l], 2.));
    if (d.icw == 0) {
      cdw[l] = d.cw;
    } else if (d.icw == 1) {
      cdw[l] = 0.0015 * 1. / (1.0 + std::exp((12.5 - ws) / 1.56)) + 0.00104;
    } else {  // (1/2) is the integer 0 there: the square root term is 1
      cdw[l] = ws == 0 ? d.cw : std::pow((0.239 + 0.0433 * ((ws - 8.271) + 1.0)) / ws, 2);
    }
    if (!heat_too) continue;
    const double Qsw = x[0], Ta = x[1], Rh = x[2];
    const double ea = sat_vapor_pressure(Ta + 273.) * Rh;
    double Qbra;
    if (mode == 10) {
      const double Cc = 0.0;
      double EmAir = EmAirC * std::pow(ea / (Ta + 273.), 0.14285714);
      EmAir = EmAir * (1. + 0.17 * std::pow(Cc, 2.));
      Qbra = EmAir * StephanBoltzman * std::pow(Ta + 273., 4.) * (1. - Al);
    } else {
      Qbra = x[3] * (1. - Al);
    }
    surface_heat(l, Qsw, Qbra, ea, Ta, Pa, ws);
  }
}

// surfbc, si3d_boundaryconditions.f90:4490-4669: cases 0 (constant wind),
// 1-3 (spatially uniform heat budget) and 20 (wind series).  Series values are
// interpolated in time with parab; the series cases return early on
// trapezoidal steps after step 1.
void OrcLake::surfbc(int n, int istep) {
  if (d.ifsurfbc == 0) {
    double ua = -d.wa * std::sin(PI * d.phi / 180.), va = -d.wa * std::cos(PI * d.phi / 180.);
    for (int l = 1; l <= lm; ++l) uair[l] = ua, vair[l] = va, cdw[l] = d.cw;
    return;
  }
  const int nv = sbc_nvar();
  if (nv == 0) return;
  if (n > 1 && istep == 2) return;
  int np = (int)(sbc.size() / nv);
  double dth = d.dtsurfbc / 3600.;
  std::vector<double> f(np + 2), v(nv);
  for (int q = 0; q < nv; ++q) {
    for (int j = 0; j < np; ++j) f[j + 1] = sbc[(size_t)nv * j + q];
    v[q] = parab(0., thrs, f.data(), dth);
  }
  if (d.ifsurfbc == 10 || d.ifsurfbc == 11) {
    met_distribute(v.data(), true);
    return;
  }
  for (int l = 0; l <= lm1; ++l) cdw[l] = v[nv - 3], uair[l] = v[nv - 2], vair[l] = v[nv - 1];
  if (d.ifsurfbc >= 1 && d.ifsurfbc <= 3) {
    distribute_qsw(v[0]);
    distribute_heat(v.data());
  }
}

// SolarFr, si3d_boundaryconditions.f90:4786-4816 (Henderson-Sellers eq. 2.25).
double OrcLake::solar_fr(double eta, double depth) const {
  const double zA = 0.60;
  if (eta >= 0.1) {
    double beta = 0.265 * std::log(eta) + 0.614;
    return (depth < zA) ? (1. - beta * depth / zA) : (1. - beta) * std::exp(-eta * (depth - zA));
  }
  return std::exp(-eta * depth);
}

// DistributeQswH, si3d_boundaryconditions.f90:4672-4727.
void OrcLake::distribute_qsw(double eta) {
  std::vector<double> zfromt(km1 + 3);
  for (int l = 1; l <= lm; ++l) {
    int kms = kmz[l], k1s = k1z[l], nwl = kms - k1s + 1;
    if (nwl == 1) {
      qswfr(k1s, l) = 1.;
    } else if (nwl >= 2) {
      zfromt[k1s] = 0.0;
      for (int k = k1s + 1; k <= kms + 1; ++k) {
        zfromt[k] = zfromt[k - 1] + hp(k - 1, l);
        qswfr(k - 1, l) = solar_fr(eta, zfromt[k - 1]) - solar_fr(eta, zfromt[k]);
      }
      double sum = 0.0;
      for (int k = k1s; k <= kms; ++k) sum += qswfr(k, l);
      double rem = 1. - sum;
      if (rem > 1.e-6)
        for (int k = k1s; k <= kms; ++k) qswfr(k, l) += rem / nwl;
    }
  }
}

static double sat_vapor_pressure(double T) { return 2.1718e10 * std::exp(-4157. / (T - 33.91)); }
static double latent_heat_vaporization(double T) { return 1.91846e6 * std::pow(T / (T - 33.91), 2.); }

// DistributeMomentumHeatSourcesH, si3d_boundaryconditions.f90:5144-5485,
// cases 1-3.  v holds the interpolated series values of one record.
void OrcLake::distribute_heat(const double *v) {
  const double StephanBoltzman = 5.6697e-8, Al = 0.03, Cp = 4181.6;
  const int mode = d.ifsurfbc;
  const double Qsw = v[1];
  for (int l = 1; l <= lm; ++l)
    for (int k = 1; k <= km1; ++k) heat(k, l) = 0.0;
  double Qn = 0, Ta = 0, Pa = 0, Rh = 0, Cc = 0, Qlw = 0, ea = 0, Qbra = 0;
  if (mode == 1) {
    Qn = v[2];
  } else {
    Ta = v[2], Pa = v[3], Rh = v[4];
    if (mode == 2) Cc = v[5]; else Qlw = v[5];
    ea = sat_vapor_pressure(Ta + 273.) * Rh;
    if (mode == 2) {
      double EmAir = 0.642 * std::pow(ea / (Ta + 273.), 0.14285714);
      EmAir = EmAir * (1. + 0.17 * std::pow(Cc, 2.));
      double Qbri = EmAir * StephanBoltzman * std::pow(Ta + 273., 4.);
      Qbra = Qbri * (1. - Al);
    } else {
      Qbra = Qlw * (1. - Al);
    }
  }
  for (int l = 1; l <= lm; ++l) {
    if (mode != 1) {
      surface_heat(l, Qsw, Qbra, ea, Ta, Pa, std::sqrt(std::pow(uair[l], 2.) + std::pow(vair[l], 2.)));
      continue;
    }
    const int k1s = k1z[l], kms = kmz[l];
    heat(k1s, l) = heat(k1s, l) + (Qn - Qsw);
    for (int k = k1s; k <= kms; ++k) {
      heat(k, l) = (heat(k, l) + Qsw * qswfr(k, l)) / ((rhop(k, l) + 1000.) * Cp);
      if (Qsw < -100.) heat(k, l) = 0.0;
    }
  }
}

// The column part of DistributeMomentumHeatSourcesH for the budget modes
// (cases 2, 3, 10, 11): Kondo (1975) neutral transfer coefficients for the
// column's wind speed, longwave, latent and sensible fluxes into the surface
// layer, shortwave over the column, all in temperature units.
void OrcLake::surface_heat(int l, double Qsw, double Qbra, double ea, double Ta, double Pa, double ws) {
  const double range[6] = {0.000, 2.2000, 5.0000, 8.0000, 25.00, 50.0};
  const double a_h[5] = {0.000, 0.9270, 1.1500, 1.1700, 1.652}, b_h[5] = {1.185, 0.0521, 0.0100, 0.0075, -0.017};
  const double c_h[5] = {0.000, 0.0000, 0.0000, -4.5E-4, 0.000}, p_h[5] = {-0.157, 1.0000, 1.0000, 1.0000, 1.000};
  const double a_e[5] = {0.000, 0.9690, 1.1800, 1.1960, 1.680}, b_e[5] = {1.230, 0.0521, 0.0100, 0.0080, -0.016};
  const double c_e[5] = {0.000, 0.0000, 0.0000, -4.0E-4, 0.000}, p_e[5] = {-0.160, 1.0000, 1.0000, 1.0000, 1.000};
  const double EmWater = 0.97, StephanBoltzman = 5.6697e-8, Cp = 4181.6, SpecificHeatAir = 1012.0;
  int irg;
  if (ws > range[0] && ws <= range[1]) irg = 0;
  else if (ws > range[1] && ws <= range[2]) irg = 1;
  else if (ws > range[2] && ws <= range[3]) irg = 2;
  else if (ws > range[3] && ws <= range[4]) irg = 3;
  else irg = 4;
  const int k1s = k1z[l], kms = kmz[l];
  double coeffH = (a_h[irg] + b_h[irg] * std::pow(ws, p_h[irg]) + c_h[irg] * std::pow(ws - 8., 2.)) * 1.e-3;
  double coeffE = (a_e[irg] + b_e[irg] * std::pow(ws, p_e[irg]) + c_e[irg] * std::pow(ws - 8., 2.)) * 1.e-3;
  double esw = sat_vapor_pressure(salp(k1s, l) + 273.);
  double Lv = latent_heat_vaporization(salp(k1s, l) + 273.);
  heat(k1s, l) = Qbra - EmWater * StephanBoltzman * std::pow(salp(k1s, l) + 273., 4.);
  heat(k1s, l) = heat(k1s, l) - RHOAIR * Lv * coeffE * (esw - ea) * 0.622 / Pa * ws -
                 RHOAIR * SpecificHeatAir * coeffH * (salp(k1s, l) - Ta) * ws;
  for (int k = k1s; k <= kms; ++k) {
    heat(k, l) = (heat(k, l) + Qsw * qswfr(k, l)) / ((rhop(k, l) + 1000.) * Cp);
    if (Qsw < -100.) heat(k, l) = 0.0;
  }
}

// UpdateHorizontalMixingCoefficients, si3d_mixing.f90:894-966 (Smagorinsky, ihd>1).
void OrcLake::hmix() {
  if (d.ihd <= 1) return;
  const double horcon = 0.03;
  for (int l = 1; l <= lm; ++l) {
    int kms = kmz[l], k1s = k1z[l];
    for (int k = k1s; k <= kms; ++k) {
      if (hp(k, lWC[l]) <= ZERO || hp(k, lEC[l]) <= ZERO || hp(k, lSC[l]) <= ZERO ||
          hp(k, lNC[l]) <= ZERO) {
        kh(k, l) = d.Ax0;
      } else {
        double dudx = (upp(k, l) - upp(k, lWC[l])) / dx;
        double dvdy = (vpp(k, l) - vpp(k, lSC[l])) / dy;
        double dudy = ((upp(k, lNC[l]) + upp(k, lNC[lWC[l]])) - (upp(k, lSC[l]) + upp(k, lSC[lWC[l]]))) / fourdy;
        double dvdx = ((vpp(k, lEC[l]) + vpp(k, lEC[lSC[l]])) - (vpp(k, lWC[l]) + vpp(k, lWC[lSC[l]]))) / fourdx;
        kh(k, l) = std::max(d.Ax0, horcon * dx * dy *
                                       std::sqrt(dudx * dudx + dvdy * dvdy + .5 * (dudy + dvdx) * (dudy + dvdx)));
      }
    }
  }
}

// exmom, si3d_procedures.f90:578-1173.
void OrcLake::exmom(int ieq) {
  const double twodt1 = twodt * tz;
  for (int k = 0; k <= km1 + 1; ++k) haxpp(k, lm1) = 0.0, haypp(k, lm1) = 0.0;
  const bool x = (ieq == 1);
  F3 &hq = x ? hup : hvp;     // thickness at the momentum point (for th)
  F3 &hqpp = x ? hupp : hvpp;
  // coefficient arrays (632-690 / 966-1026)
  for (int l = 1; l <= lm; ++l) {
    int ln = x ? lEC[l] : lNC[l];
    int kmq = std::min(kmz[ln], kmz[l]);
    int k1q = x ? k1u[l] : k1v[l];
    for (int k = 1; k <= k1q - 1; ++k) haypp(k, l) = haxpp(k, l) = th(k, l) = th1(k, l) = 0.0;
    for (int k = kmq + 1; k <= km1; ++k) haypp(k, l) = haxpp(k, l) = th(k, l) = th1(k, l) = 0.0;
    if (!(x ? maskE(l) : maskN(l))) {
      for (int k = k1q; k <= kmq; ++k) haypp(k, l) = haxpp(k, l) = th(k, l) = th1(k, l) = 0.0;
      continue;
    }
    for (int k = k1q; k <= kmq; ++k) {
      if (d.ihd == 1) {
        if (x) {
          haxpp(k, l) = d.Ax0 * std::min(hupp(k, lEC[l]), hupp(k, l));
          haypp(k, l) = d.Ay0 * std::min(hupp(k, lNC[l]), hupp(k, l));
        } else {
          haypp(k, l) = d.Ay0 * std::min(hvpp(k, lNC[l]), hvpp(k, l));
          haxpp(k, l) = d.Ax0 * std::min(hvpp(k, lEC[l]), hvpp(k, l));
        }
      } else if (d.ihd > 1) {
        double k4 = (kh(k, lEC[l]) + kh(k, l) + kh(k, lNC[l]) + kh(k, lEC[lNC[l]])) * 0.25;
        if (x) {
          haxpp(k, l) = kh(k, lEC[l]) * std::min(hupp(k, lEC[l]), hupp(k, l));
          haypp(k, l) = k4 * std::min(hupp(k, lNC[l]), hupp(k, l));
        } else {
          haypp(k, l) = kh(k, lNC[l]) * std::min(hvpp(k, lNC[l]), hvpp(k, l));
          haxpp(k, l) = k4 * std::min(hvpp(k, lEC[l]), hvpp(k, l));
        }
      } else {
        haxpp(k, l) = haypp(k, l) = 0.0;
      }
    }
    for (int k = k1q; k <= kmq; ++k) {
      th(k, l) = hq(k - 1, l) / (hq(k - 1, l) + hq(k, l));
      th1(k, l) = 1. - th(k, l);
    }
    th(k1q, l) = 0.0;
    th1(k1q, l) = 0.0;
    th(kmq + 1, l) = 0.5;
    th1(kmq + 1, l) = 0.5;
  }
  (void)hqpp;
  // explicit terms (795-950 / 1031-1160)
  for (int l = 1; l <= lm; ++l) {
    int ln = x ? lEC[l] : lNC[l];
    int kmq = std::min(kmz[ln], kmz[l]);
    int k1q = x ? k1u[l] : k1v[l];
    if (!x) {
      for (int k = 0; k <= km1 + 1; ++k) ex(k, l) = 0.0;
      if (!maskN(l)) continue;
    } else {
      for (int k = 1; k <= k1q - 1; ++k) ex(k, l) = 0.0;
      for (int k = kmq + 1; k <= km1; ++k) ex(k, l) = 0.0;
      if (!maskE(l)) {
        for (int k = k1q; k <= kmq; ++k) ex(k, l) = 0.0;
        continue;
      }
    }
    const int lE = lEC[l], lW = lWC[l], lN = lNC[l], lS = lSC[l];
    for (int k = k1q; k <= kmq; ++k) {
      if (hp(k, l) <= ZERO || hp(k, ln) <= ZERO) {
        ex(k, l) = x ? uhpp(k, l) : vhpp(k, l);
        continue;
      }
      double adv = 0.0, cor, hd = 0.0, uE, uW, vN, vS, wU, wD;
      F3 &q = x ? up : vp;     // advected velocity at n (centered)
      F3 &qpp = x ? upp : vpp; // advected velocity at n-1 (upwind, diffusion)
      if (x) {
        cor = 0.25 * d.f * (vhp(k, lE) + vhp(k, l) + vhp(k, lSC[lE]) + vhp(k, lS));
        uE = uhp(k, lE) + uhp(k, l);
        uW = uhp(k, l) + uhp(k, lW);
        vN = vhp(k, lE) + vhp(k, l);
        vS = vhp(k, lSC[lE]) + vhp(k, lS);
        wU = wp(k, lE) + wp(k, l);
        wD = wp(k + 1, lE) + wp(k + 1, l);
      } else {
        cor = 0.25 * d.f * (uhp(k, lN) + uhp(k, l) + uhp(k, lWC[lN]) + uhp(k, lW));
        uE = uhp(k, lN) + uhp(k, l);
        uW = uhp(k, lWC[lN]) + uhp(k, lW);
        vN = vhp(k, lN) + vhp(k, l);
        vS = vhp(k, l) + vhp(k, lS);
        wU = wp(k, lN) + wp(k, l);
        wD = wp(k + 1, lN) + wp(k + 1, l);
      }
      if (k == k1q) wU = 0.0;
      if (k == kmq) wD = 0.0;
      auto centered_h = [&]() {
        double scE = q(k, lE) + q(k, l), scW = q(k, lW) + q(k, l), scN = q(k, lN) + q(k, l),
               scS = q(k, lS) + q(k, l);
        return (uE * scE - uW * scW) / fourdx + (vN * scN - vS * scS) / fourdy;
      };
      auto upwind_h = [&]() {
        return ((uE + std::fabs(uE)) * qpp(k, l) + (uE - std::fabs(uE)) * qpp(k, lE) -
                (uW + std::fabs(uW)) * qpp(k, lW) - (uW - std::fabs(uW)) * qpp(k, l)) / fourdx +
               ((vN + std::fabs(vN)) * qpp(k, l) + (vN - std::fabs(vN)) * qpp(k, lN) -
                (vS + std::fabs(vS)) * qpp(k, lS) - (vS - std::fabs(vS)) * qpp(k, l)) / fourdy;
      };
      auto th_v = [&]() {
        return (wU * (th(k, l) * q(k, l) + th1(k, l) * q(k - 1, l)) -
                wD * (th(k + 1, l) * q(k + 1, l) + th1(k + 1, l) * q(k, l))) / 2.;
      };
      switch (d.itrmom) {
        case 1: adv = centered_h() + th_v(); break;
        case 2:
          adv = upwind_h() + ((wU + std::fabs(wU)) * qpp(k, l) + (wU - std::fabs(wU)) * qpp(k - 1, l)) / 4. -
                ((wD + std::fabs(wD)) * qpp(k + 1, l) + (wD - std::fabs(wD)) * qpp(k, l)) / 4.;
          break;
        case 3: {
          double scU = (q(k, l) * hq(k, l) + q(k - 1, l) * hq(k - 1, l)) / (hq(k, l) + hq(k - 1, l));
          double scD = (q(k, l) * hq(k, l) + q(k + 1, l) * hq(k + 1, l)) / (hq(k, l) + hq(k + 1, l));
          adv = centered_h() + (wU * scU - wD * scD) / 2.;
          break;
        }
        case 4: adv = upwind_h() + th_v(); break;
      }
      if (d.ihd == 1) {
        hd = (haypp(k, l) * (qpp(k, lN) - qpp(k, l)) - haypp(k, lS) * (qpp(k, l) - qpp(k, lS))) / dydy +
             (haxpp(k, l) * (qpp(k, lE) - qpp(k, l)) - haxpp(k, lW) * (qpp(k, l) - qpp(k, lW))) / dxdx;
      } else if (d.ihd > 1) {
        if (x) {
          hd = 2. * (haxpp(k, l) * (upp(k, lE) - upp(k, l)) - haxpp(k, lW) * (upp(k, l) - upp(k, lW))) / dxdx +
               (haypp(k, l) * (upp(k, lN) - upp(k, l)) - haypp(k, lS) * (upp(k, l) - upp(k, lS))) / dydy +
               (haypp(k, l) * (vpp(k, lE) - vpp(k, l)) - haypp(k, lS) * (vpp(k, lEC[lS]) - vpp(k, lS))) / dxdy;
        } else {
          hd = 2. * (haypp(k, l) * (vpp(k, lN) - vpp(k, l)) - haypp(k, lS) * (vpp(k, l) - vpp(k, lS))) / dydy +
               (haxpp(k, l) * (vpp(k, lE) - vpp(k, l)) - haxpp(k, lW) * (vpp(k, l) - vpp(k, lW))) / dxdx +
               (haxpp(k, l) * (upp(k, lN) - upp(k, l)) - haxpp(k, lW) * (upp(k, lNC[lW]) - upp(k, lW))) / dxdy;
        }
      }
      if (x)
        ex(k, l) = uhpp(k, l) - twodt1 * (adv * d.iadv - cor - hd);
      else
        ex(k, l) = vhpp(k, l) - twodt1 * (adv * d.iadv + cor - hd);
    }
  }
  if (!obc.empty()) mod_exmom(ieq);  // MODexmom4openBCX / Y, 957 / 1165
}

// bclnc_km, si3d_procedures.f90:1633-1746.
void OrcLake::bclnc_km(int l, int kb, int ieq, double &huvpdrho) {
  const double eps = FLT_EPSILON;  // EPSILON(0.0) of a default REAL
  const bool x = (ieq == 1);
  int ln = x ? lEC[l] : lNC[l];
  double hq = x ? hup(kb, l) : hvp(kb, l);
  double hplus = hq + eps, hminus = hq - eps;
  bool ce, cw;
  double rn, rw;
  if (hp(kb, ln) > hplus || hp(kb, ln) < hminus) {
    ce = true;
    double dz1 = 0.5 * (hp(kb - 1, ln) + hp(kb, ln)), dz2 = 0.5 * (hp(kb - 1, ln) + hq);
    rn = rhop(kb - 1, ln) + (dz2 / dz1) * (rhop(kb, ln) - rhop(kb - 1, ln));
  } else {
    ce = false;
    rn = rhop(kb, ln);
  }
  if (hp(kb, l) > hplus || hp(kb, l) < hminus) {
    cw = true;
    double dz1 = 0.5 * (hp(kb - 1, l) + hp(kb, l)), dz2 = 0.5 * (hp(kb - 1, l) + hq);
    rw = rhop(kb - 1, l) + (dz2 / dz1) * (rhop(kb, l) - rhop(kb - 1, l));
  } else {
    cw = false;
    rw = rhop(kb, l);
  }
  if (ce || cw) huvpdrho = (x ? gdtdx : gdtdy) * hq * (rn - rw);
}

// matmom, si3d_procedures.f90:1176-1630 (no point sources).
void OrcLake::matmom(int ieq) {
  const double twodt1 = twodt * tz;
  const bool x = (ieq == 1);
  std::vector<double> Avq(km1 + 3), hupdrho(km1 + 3), bcl(km1 + 3), dudz(km1 + 3), vdiff(km1 + 3),
      hh(km1 + 3), gg(km1 + 3), a1(km1 + 3), a2(km1 + 3), a3(km1 + 3), ag(km1 + 3), ar(km1 + 3);
  F3 &hq = x ? hup : hvp;
  F3 &qpp = x ? upp : vpp;
  F3 &ag3 = x ? agx : agy;
  F3 &ar3 = x ? arx : ary;
  std::vector<double> &eag = x ? eagx : eagy;
  std::vector<double> &ear = x ? earx : eary;
  for (int l = 1; l <= lm; ++l) {
    if (!(x ? maskE(l) : maskN(l))) continue;  // CE / CN column lists (4663-4743)
    int ln = x ? lEC[l] : lNC[l];
    int kmq = std::min(kmz[ln], kmz[l]);
    int k1q = x ? k1u[l] : k1v[l];
    int nwl = kmq - k1q + 1;
    if (nwl < 1) continue;
    for (int k = 0; k <= km1 + 2; ++k) Avq[k] = 0.0;
    for (int k = (x ? k1q + 1 : k1q); k <= kmq; ++k) Avq[k] = 0.5 * (Av(k, ln) + Av(k, l));
    double rsum = 0.0;
    for (int k = k1q; k <= kmq; ++k) rsum += 1000. + rhop(k, l);
    double rhoq = rsum / nwl;
    double ws0 = rhoq * (x ? gdtdx : gdtdy) * (spp[ln] - spp[l]);
    if (d.ibc == 0) {
      for (int k = k1q; k <= kmq; ++k) bcl[k] = 0.0;
    } else {
      for (int k = k1q; k <= kmq; ++k) hupdrho[k] = (x ? gdtdx : gdtdy) * hq(k, l) * (rhop(k, ln) - rhop(k, l));
      bcl[k1q] = hupdrho[k1q];
      if (kmq > k1q) {
        bclnc_km(l, kmq, ieq, hupdrho[kmq]);
        for (int k = k1q + 1; k <= kmq; ++k) bcl[k] = bcl[k - 1] + (hupdrho[k - 1] + hupdrho[k]);
      }
    }
    if (nwl == 1) {
      vdiff[k1q] = 0.0;
    } else {
      for (int k = k1q + 1; k <= kmq; ++k) {
        double dz = hq(k - 1, l) + hq(k, l);
        dudz[k] = Avq[k] * (qpp(k - 1, l) - qpp(k, l)) / dz;
      }
      dudz[k1q] = 0.0;
      dudz[kmq + 1] = 0.0;
      for (int k = k1q; k <= kmq; ++k) vdiff[k] = twodt * (dudz[k] - dudz[k + 1]) * 2. * (1. - d.theta);
    }
    for (int k = k1q; k <= kmq; ++k) {
      hh[k] = hq(k, l) / rhoq;
      gg[k] = ex(k, l) - hh[k] * (bcl[k] + ws0) * tz + vdiff[k] * tz;
    }
    if (nwl == 1) {
      a2[k1q] = 1.0;
    } else {
      for (int k = k1q; k <= kmq - 1; ++k)
        a3[k] = -twodt1 * Avq[k + 1] / (hq(k + 1, l) * (hq(k, l) + hq(k + 1, l))) * 2. * d.theta;
      a3[kmq] = 0.0;
      for (int k = k1q + 1; k <= kmq; ++k)
        a1[k] = -twodt1 * Avq[k] / (hq(k - 1, l) * (hq(k - 1, l) + hq(k, l))) * 2. * d.theta;
      a1[k1q] = 0.0;
      for (int k = k1q; k <= kmq; ++k)
        a2[k] = 1.0 - (hq(k - 1, l) / hq(k, l)) * a1[k] - (hq(k + 1, l) / hq(k, l)) * a3[k];
    }
    // wind stress (1348-1360 / 1541-1553)
    double usurf, vsurf, cwq;
    if (x) {
      usurf = up(k1q, l);
      int lS = lSC[l], lE = lEC[l], lSE = lSC[lEC[l]];
      vsurf = (vp(k1v[l], l) + vp(k1v[lS], lS) + vp(k1v[lE], lE) + vp(k1v[lSE], lSE)) / 4.;
      cwq = cdw[l] * RHOAIR * std::sqrt((vair[l] - vsurf) * (vair[l] - vsurf) + (uair[l] - usurf) * (uair[l] - usurf));
      gg[k1q] += cwq / rhoq * uair[l] * twodt1;
    } else {
      vsurf = vp(k1q, l);
      int lW = lWC[l], lN = lNC[l], lNW = lWC[lNC[l]];
      usurf = (up(k1u[l], l) + up(k1u[lW], lW) + up(k1u[lN], lN) + up(k1u[lNW], lNW)) / 4.;
      cwq = cdw[l] * RHOAIR * std::sqrt((uair[l] - usurf) * (uair[l] - usurf) + (vair[l] - vsurf) * (vair[l] - vsurf));
      gg[k1q] += cwq / rhoq * vair[l] * twodt1;
    }
    a2[k1q] += cwq / rhoq / hq(k1q, l) * twodt1;
    // bottom stress (1364-1369 / 1557-1562)
    double taub;
    if (x) {
      double vb = 0.25 * (vhpp(kmq, lEC[l]) + vhpp(kmq, l) + vhpp(kmq, lSC[l]) + vhpp(kmq, lSC[lEC[l]]));
      taub = d.cd * std::sqrt(uhpp(kmq, l) * uhpp(kmq, l) + vb * vb) / (hup(kmq, l) * hup(kmq, l));
    } else {
      double ub = 0.25 * (uhpp(kmq, lNC[l]) + uhpp(kmq, lWC[lNC[l]]) + uhpp(kmq, lWC[l]) + uhpp(kmq, l));
      taub = d.cd * std::sqrt(vhpp(kmq, l) * vhpp(kmq, l) + ub * ub) / (hvp(kmq, l) * hvp(kmq, l));
    }
    a2[kmq] += taub * twodt1;
    if (nwl == 1) {
      ag[k1q] = gg[k1q] / a2[k1q];
      ar[k1q] = hh[k1q] / a2[k1q];
    } else {
      trid2(nwl, &a1[k1q], &a2[k1q], &a3[k1q], &gg[k1q], &hh[k1q], &ag[k1q], &ar[k1q]);
    }
    double sg = 0.0, sr = 0.0;
    for (int k = k1q; k <= kmq; ++k) {
      ag3(k, l) = ag[k];
      ar3(k, l) = ar[k];
      sg += ag[k];
      sr += ar[k];
    }
    eag[l] = sg;
    ear[l] = sr;
  }
}

// matcon, si3d_procedures.f90:1749-1867 (no point sources, no open boundaries).
void OrcLake::matcon() {
  const double cx = gdt2dx2 * tz * tz, cy = gdt2dy2 * tz * tz, dtdx1 = dtdx * tz, dtdy1 = dtdy * tz,
               dt1 = dt * tz;
  sx[lm1] = 0.0;
  sy[lm1] = 0.0;
  for (int l = 1; l <= lm; ++l) {
    sx[l] = maskE(l) ? cx * 1000. * earx[l] : 0.0;
    sy[l] = maskN(l) ? cy * 1000. * eary[l] : 0.0;
  }
  for (int l = 1; l <= lm; ++l) {
    int lW = lWC[l], lS = lSC[l];
    double sum = 0.0;
    for (int k = k1; k <= kmz[l]; ++k)
      sum += (uhpp(k, l) - uhpp(k, lW)) / dx + (vhpp(k, l) - vhpp(k, lS)) / dy;
    dd[l] = dt1 * sum;
    double eW = (lW == lm1) ? 0.0 : eagx[lW], eS = (lS == lm1) ? 0.0 : eagy[lS];
    qq[l] = spp[l] - dtdx1 * (eagx[l] - eW) - dtdy1 * (eagy[l] - eS) - dd[l];
    rr[l] = 1 + sx[l] + sx[lW] + sy[l] + sy[lS];
  }
  if (!obc.empty()) mod_qq();  // MODqqddrr4openBC, 1859
}

// SolverSparse (si3d_procedures.f90:1870-2065): A zeta = qq with ELLPACK rows
// {rr, -sx_W, -sy_S, -sy, -sx} (1954-1967) and initial guess sp (1970).  The
// reference calls NSPCG MIC(0)+CG (nspcg.f mic1/cg).  Restated here: MIC(0) with
// omega = 1 in natural l order (the modified-diagonal recurrence of nspcg.f
// icfp 20789-20806 for a matrix with property A) and conjugate gradients, stopped
// on the FP64 true residual ||qq - A zeta|| <= tol ||qq|| instead of NSPCG's
// emin-scaled test 2.
int OrcLake::solver() {
  const int N = lm;
  std::vector<double> x(N + 1), b(N + 1), r(N + 1), z(N + 1), p(N + 1), q(N + 1), dg(N + 1);
  auto A = [&](const std::vector<double> &v, std::vector<double> &o) {
    for (int m = 1; m <= N; ++m) {
      double acc = rr[m] * v[m];
      int lW = lWC[m], lS = lSC[m], lN = lNC[m], lE = lEC[m];
      if (lW != lm1) acc -= sx[lW] * v[lW];
      if (lS != lm1) acc -= sy[lS] * v[lS];
      if (lN != lm1) acc -= sy[m] * v[lN];
      if (lE != lm1) acc -= sx[m] * v[lE];
      o[m] = acc;
    }
  };
  // MIC(0) factorization
  for (int m = 1; m <= N; ++m) {
    double dm = rr[m];
    int lW = lWC[m], lS = lSC[m];
    if (lW != lm1) dm -= sx[lW] * (sx[lW] + sy[lW]) / dg[lW];
    if (lS != lm1) dm -= sy[lS] * (sy[lS] + sx[lS]) / dg[lS];
    dg[m] = dm;
  }
  auto prec = [&](const std::vector<double> &rin, std::vector<double> &zo) {
    for (int m = 1; m <= N; ++m) {  // (D + L) y = r
      double acc = rin[m];
      int lW = lWC[m], lS = lSC[m];
      if (lW != lm1) acc += sx[lW] * zo[lW];
      if (lS != lm1) acc += sy[lS] * zo[lS];
      zo[m] = acc / dg[m];
    }
    for (int m = N; m >= 1; --m) {  // (I + D^-1 U) z = y
      double acc = 0.0;
      int lN = lNC[m], lE = lEC[m];
      if (lN != lm1) acc += sy[m] * zo[lN];
      if (lE != lm1) acc += sx[m] * zo[lE];
      zo[m] += acc / dg[m];
    }
  };
  double bn = 0.0;
  for (int m = 1; m <= N; ++m) {
    b[m] = qq[m];
    x[m] = sp[m];
    bn += b[m] * b[m];
  }
  bn = std::sqrt(bn);
  if (bn == 0.0) bn = 1.0;
  const double tol = d.pcg_tol > 0 ? d.pcg_tol : 1e-12;
  const int maxit = d.pcg_maxit > 0 ? d.pcg_maxit : 2000;
  int it = 0;
  double relres = 0.0;
  for (int outer = 0; outer < 3; ++outer) {  // restart from the true residual if needed
    A(x, q);
    double rn = 0.0;
    for (int m = 1; m <= N; ++m) {
      r[m] = b[m] - q[m];
      rn += r[m] * r[m];
    }
    relres = std::sqrt(rn) / bn;
    if (relres <= tol) break;
    prec(r, z);
    double rz = 0.0;
    for (int m = 1; m <= N; ++m) p[m] = z[m], rz += r[m] * z[m];
    while (it < maxit) {
      A(p, q);
      double pq = 0.0;
      for (int m = 1; m <= N; ++m) pq += p[m] * q[m];
      double alpha = rz / pq;
      rn = 0.0;
      for (int m = 1; m <= N; ++m) {
        x[m] += alpha * p[m];
        r[m] -= alpha * q[m];
        rn += r[m] * r[m];
      }
      ++it;
      if (std::sqrt(rn) / bn <= 0.5 * tol) break;
      prec(r, z);
      double rz1 = 0.0;
      for (int m = 1; m <= N; ++m) rz1 += r[m] * z[m];
      double beta = rz1 / rz;
      rz = rz1;
      for (int m = 1; m <= N; ++m) p[m] = z[m] + beta * p[m];
    }
  }
  A(x, q);
  double rn = 0.0;
  for (int m = 1; m <= N; ++m) rn += (b[m] - q[m]) * (b[m] - q[m]);
  relres = std::sqrt(rn) / bn;
  last_iters = it;
  last_relres = relres;
  total_iters += it;
  for (int m = 1; m <= N; ++m) s[m] = x[m];
  slayer_h();
  return relres <= tol ? 0 : 1;
}

// slayer_h, si3d_procedures.f90:2070-2175 (including the k=k+1; k=k1u(l) re-test at 2136).
void OrcLake::slayer_h() {
  for (int k = 0; k <= km1 + 1; ++k) h(k, lm1) = hp(k, lm1), hu(k, lm1) = hup(k, lm1), hv(k, lm1) = hvp(k, lm1);
  for (int l = 1; l <= lm; ++l) {
    for (int k = 1; k <= k1z[l] - 1; ++k) h(k, l) = hp(k, l);
    for (int k = 1; k <= k1u[l] - 1; ++k) hu(k, l) = hup(k, l);
    for (int k = 1; k <= k1v[l] - 1; ++k) hv(k, l) = hvp(k, l);
    for (int k = k1z[l] + 2; k <= km1; ++k) h(k, l) = hp(k, l);
    for (int k = k1u[l] + 1; k <= km1; ++k) hu(k, l) = hup(k, l);
    for (int k = k1v[l] + 2; k <= km1; ++k) hv(k, l) = hvp(k, l);
    int k = k1z[l];
    double haux = std::min(zlevel[k + 1], hhs[l]) + s[l];
    if (haux <= HMIN) {
      h(k, l) = ZERO;
      k = k + 1;
      haux = std::min(zlevel[k + 1], hhs[l]) + s[l];
      h(k, l) = (haux <= HMIN) ? ZERO : haux;
    } else {
      h(k, l) = haux;
    }
    if (maskE(l)) {
      k = k1u[l];
      haux = std::min(zlevel[k + 1], hhu[l]) + std::max(s[l], s[lEC[l]]);
      if (haux <= HMIN) {
        hu(k, l) = ZERO;
        k = k1u[l];  // the reference re-tests the same layer
        haux = std::min(zlevel[k + 1], hhu[l]) + std::max(s[l], s[lEC[l]]);
        hu(k, l) = (haux <= HMIN) ? ZERO : haux;
      } else {
        hu(k, l) = haux;
      }
    } else {
      for (int kk = k1u[l]; kk <= std::min(k1u[l] + 1, km1); ++kk) hu(kk, l) = hup(kk, l);
    }
    if (maskN(l)) {
      k = k1v[l];
      haux = std::min(zlevel[k + 1], hhv[l]) + std::max(s[l], s[lNC[l]]);
      if (haux <= HMIN) {
        hv(k, l) = ZERO;
        k = k + 1;
        haux = std::min(zlevel[k + 1], hhv[l]) + std::max(s[l], s[lNC[l]]);
        hv(k, l) = (haux <= HMIN) ? ZERO : haux;
      } else {
        hv(k, l) = haux;
      }
    } else {
      for (int kk = k1v[l]; kk <= std::min(k1v[l] + 1, km1); ++kk) hv(kk, l) = hvp(kk, l);
    }
  }
}

// layer_h, si3d_procedures.f90:2181-2231.
void OrcLake::layer_h() {
  for (int l = 1; l <= lm; ++l) {
    for (int k = k1; k <= kmz[l]; ++k) {
      double v = std::min(zlevel[k + 1], hhs[l]) - std::max(zlevel[k], -s[l]);
      h(k, l) = v <= HMIN ? ZERO : v;
    }
    if (maskE(l)) {
      int kmx = std::min(kmz[l], kmz[lEC[l]]);
      for (int k = k1; k <= kmx; ++k) {
        double v = std::min(zlevel[k + 1], hhu[l]) - std::max(zlevel[k], -std::max(s[l], s[lEC[l]]));
        hu(k, l) = v <= HMIN ? ZERO : v;
      }
    }
    if (maskN(l)) {
      int kmy = std::min(kmz[l], kmz[lNC[l]]);
      for (int k = k1; k <= kmy; ++k) {
        double v = std::min(zlevel[k + 1], hhv[l]) - std::max(zlevel[k], -std::max(s[l], s[lNC[l]]));
        hv(k, l) = v <= HMIN ? ZERO : v;
      }
    }
  }
}

// layer_hp2 (si3d_procedures.f90:2238-2293); layer_hp3 (2296-2364) computes the same.
void OrcLake::layer_hp2() {
  for (int l = 1; l <= lm; ++l) {
    for (int k = k1; k <= kmz[l]; ++k) {
      double v = std::min(zlevel[k + 1], hhs[l]) - std::max(zlevel[k], -sp[l]);
      hp(k, l) = v <= HMIN ? ZERO : v;
    }
    if (maskE(l)) {
      int kmx = std::min(kmz[l], kmz[lEC[l]]);
      for (int k = k1; k <= kmx; ++k) {
        double v = std::min(zlevel[k + 1], hhu[l]) - std::max(zlevel[k], -std::max(sp[l], sp[lEC[l]]));
        hup(k, l) = v <= HMIN ? ZERO : v;
      }
    }
    if (maskN(l)) {
      int kmy = std::min(kmz[l], kmz[lNC[l]]);
      for (int k = k1; k <= kmy; ++k) {
        double v = std::min(zlevel[k + 1], hhv[l]) - std::max(zlevel[k], -std::max(sp[l], sp[lNC[l]]));
        hvp(k, l) = v <= HMIN ? ZERO : v;
      }
    }
  }
}

// TopLayerIndexp2, si3d_procedures.f90:2368-2427.
void OrcLake::toplayer_indexp2() {
  for (int l = 1; l <= lm; ++l) {
    k1z[l] = k1u[l] = k1v[l] = km1;
    for (int k = k1; k <= kmz[l]; ++k)
      if (hp(k, l) > ZERO) {
        k1z[l] = k;
        break;
      }
    if (maskE(l)) {
      int kmx = std::min(kmz[l], kmz[lEC[l]]);
      for (int k = k1; k <= kmx; ++k)
        if (hup(k, l) > ZERO) {
          k1u[l] = k;
          break;
        }
    }
    if (maskN(l)) {
      int kmy = std::min(kmz[l], kmz[lNC[l]]);
      for (int k = k1; k <= kmy; ++k)
        if (hvp(k, l) > ZERO) {
          k1v[l] = k;
          break;
        }
    }
  }
}

// vel, si3d_procedures.f90:2430-2567.  A face whose top index already lies below
// its bottom (kmx < k1u) is left alone; the reference would write one element past
// the column there (uh(km1+1,l)).
void OrcLake::vel() {
  const double gthx1 = gdtdx * tz, gthy1 = gdtdy * tz;
  for (int l = 1; l <= lm; ++l) {
    if (maskE(l)) {
      int kmx = std::min(kmz[lEC[l]], kmz[l]), k0 = k1u[l];
      double cxx = gthx1 * 1000. * (s[lEC[l]] - s[l]);
      for (int k = k0; k <= kmx; ++k) uh(k, l) = agx(k, l) - cxx * arx(k, l);
      if (k0 <= kmx && hu(k0, l) <= ZERO) {
        uh(k0 + 1, l) = uh(k0, l) + uh(k0 + 1, l);
        uh(k0, l) = 0.0;
        k1u[l] = k0 + 1;
      }
    }
    if (maskN(l)) {
      int kmy = std::min(kmz[lNC[l]], kmz[l]), k0 = k1v[l];
      double cyy = gthy1 * 1000. * (s[lNC[l]] - s[l]);
      for (int k = k0; k <= kmy; ++k) vh(k, l) = agy(k, l) - cyy * ary(k, l);
      if (k0 <= kmy && hv(k0, l) <= ZERO) {
        vh(k0 + 1, l) = vh(k0, l) + vh(k0 + 1, l);
        vh(k0, l) = 0.0;
        k1v[l] = k0 + 1;
      }
    }
  }
  continuity(1);
  for (int l = 1; l <= lm; ++l) {
    int k0 = k1z[l];
    if (h(k0, l) <= ZERO) k1z[l] = std::min(k0 + 1, km1);
  }
}

// continuity cases 1 and 2, si3d_procedures.f90:2635-2745.
void OrcLake::continuity(int ist) {
  for (int l = 1; l <= lm; ++l) {
    int kms = kmz[l], k1s = k1z[l], lW = lWC[l], lS = lSC[l];
    if (k1s > kms) continue;
    for (int k = 0; k <= km1 + 1; ++k) wp(k, l) = 0.0;
    auto div = [&](int k) {
      if (ist == 1)
        return (uh(k, l) - uh(k, lW) + uhpp(k, l) - uhpp(k, lW)) / twodx +
               (vh(k, l) - vh(k, lS) + vhpp(k, l) - vhpp(k, lS)) / twody;
      return (uhp(k, l) - uhp(k, lW)) / dx + (vhp(k, l) - vhp(k, lS)) / dy;
    };
    for (int k = kms; k >= k1s; --k) wp(k, l) = wp(k + 1, l) - div(k);
    for (int k = k1; k <= k1s - 1; ++k) wp(k1s, l) = wp(k1s, l) - div(k);
  }
  if (!obc.empty()) mod_vel();  // MODvel4openBC, 2821
}

// exsal, si3d_procedures.f90:2827-3076 (superbee TVD; no open boundaries).
void OrcLake::exsal() { exscalar(salpp, false); }

// Face flux of exsal's TVD scheme (si3d_procedures.f90:2900-2930, 2948-2975,
// 2995-3030): s1..s4 are the scalar at the four cells along the flow axis
// (face between s2 and s3, positive velocity from s2 to s3), the superbee
// limiter C_f of the upwind gradient ratio (2928) blends first-order upwind
// (C_f = 0) and Lax-Wendroff (C_f = 1); cfdef is C_f where s3 = s2 (0 for the
// horizontal faces, 1 for the vertical ones) and courant = dt'/dx.
static double tvd_flux(double vel, double s1, double s2, double s3, double s4, double cfdef, double courant) {
  const double ss[5] = {0.0, s1, s2, s3, s4};
  double cf = cfdef;
  const int g1 = (vel >= 0.0) ? -1 : 1;  // -SIGN(1.,vel)
  if (ss[3] - ss[2] != 0) {
    const double ratio = (ss[3 + g1] - ss[2 + g1]) / (ss[3] - ss[2]);
    cf = std::max(0.0, std::max(std::min(1.0, 2. * ratio), std::min(2.0, ratio)));
  }
  return vel / 2. * (ss[3] + ss[2]) - ((1. - cf) * std::fabs(vel) + vel * vel * courant * cf) * (ss[3] - ss[2]) / 2.;
}

// The explicit transport terms of a scalar: exsal (salinity/temperature, cpp =
// salpp) and exTracer (si3d_procedures.f90:4090-4357, cpp = tracerpp of one
// tracer, which also sets Bex of the sediment layer, 4342).  Both write the
// same flux buffers (the reference's thread buffers Bth2/Bth4/Bth3).
void OrcLake::exscalar(const F3 &salpp, bool is_tracer) {
  const double twodt1 = twodt * tz;
  // EW fluxes at u-faces (2892-2930)
  for (int l = 1; l <= lm; ++l) {
    int kms = kmz[l], k1s = k1z[l], lE = lEC[l], lW = lWC[l];
    for (int k = k1s; k <= kms; ++k) {
      if (hup(k, l) > ZERO) {
        const double vel = (uhpp(k, l) + uh(k, l)) / 2.;
        const double s2 = salpp(k, l), s3 = salpp(k, lE);
        const double s1 = (hpp(k, lW) <= ZERO) ? s2 : salpp(k, lW);
        const double s4 = (hpp(k, lEC[lE]) <= ZERO) ? s3 : salpp(k, lEC[lE]);
        fx_ew(k, l) = tvd_flux(vel, s1, s2, s3, s4, 0.0, twodt1 / dx);
      } else {
        fx_ew(k, l) = 0.0;
      }
    }
  }
  // NS and vertical fluxes (2932-3035)
  for (int l = 1; l <= lm; ++l) {
    int kms = kmz[l], k1s = k1z[l], lN = lNC[l], lS = lSC[l];
    for (int k = 0; k <= km1 + 1; ++k) fx_ud(k, l) = 0.0;
    for (int k = k1s; k <= kms; ++k) {
      if (hvp(k, l) > ZERO) {
        const double vel = (vhpp(k, l) + vh(k, l)) / 2.;
        const double s2 = salpp(k, l), s3 = salpp(k, lN);
        const double s1 = (hpp(k, lS) <= ZERO) ? s2 : salpp(k, lS);
        const double s4 = (hpp(k, lNC[lN]) <= ZERO) ? s3 : salpp(k, lNC[lN]);
        fx_ns(k, l) = tvd_flux(vel, s1, s2, s3, s4, 0.0, twodt1 / dy);
      } else {
        fx_ns(k, l) = 0.0;
      }
      if (hp(k - 1, l) > ZERO) {
        const double vel = (k == k1s) ? 0.0 : wp(k, l);
        const double s2 = salpp(k, l), s3 = salpp(k - 1, l);
        const double s4 = (hpp(k - 2, l) <= ZERO) ? s3 : salpp(k - 2, l);
        const double s1 = (hpp(k + 1, l) <= ZERO) ? s2 : salpp(k + 1, l);
        const double delz = (hp(k, l) + hp(k - 1, l)) / 2.;
        fx_ud(k, l) = tvd_flux(vel, s1, s2, s3, s4, 1.0, twodt1 / delz);
      } else {
        fx_ud(k, l) = 0.0;
      }
    }
  }
  // divergence + horizontal diffusion (3038-3067); Bhaxpp/Bhaypp of 2857-2880
  for (int l = 1; l <= lm; ++l) {
    int kms = kmz[l], k1s = k1z[l], lW = lWC[l], lS = lSC[l], lE = lEC[l], lN = lNC[l];
    for (int k = 0; k <= km1 + 1; ++k) bex(k, l) = 0.0;
    for (int k = k1s; k <= kms; ++k) {
      auto bhx = [&](int c) { return (c == lm1 || k > kmz[c] || k < k1) ? 0.0 : d.Ax0 * hupp(k, c); };
      auto bhy = [&](int c) { return (c == lm1 || k > kmz[c] || k < k1) ? 0.0 : d.Ay0 * hvpp(k, c); };
      double hd = (bhx(l) * (salpp(k, lE) - salpp(k, l)) - bhx(lW) * (salpp(k, l) - salpp(k, lW))) / dxdx +
                  (bhy(l) * (salpp(k, lN) - salpp(k, l)) - bhy(lS) * (salpp(k, l) - salpp(k, lS))) / dydy;
      double fW = (lW == lm1) ? 0.0 : fx_ew(k, lW), fS = (lS == lm1) ? 0.0 : fx_ns(k, lS);
      double b = hpp(k, l) * salpp(k, l) / twodt1 - (fx_ew(k, l) - fW) / dx - (fx_ns(k, l) - fS) / dy -
                 (fx_ud(k, l) - fx_ud(k + 1, l));
      if (d.ihd > 0) b += hd;
      bex(k, l) = b;
    }
    if (is_tracer) bex(kms + 1, l) = hpp(kms + 1, l) * salpp(kms + 1, l) / twodt1;
  }
  // MODexsal4openbc (3070) / MODexTracer4openbc (4350)
  if (!obc.empty()) {
    int q = 1;
    for (int nt = 0; nt < ntr; ++nt)
      if (&salpp == &tracerpp[nt]) q = 2 + nt;
    mod_exscalar(salpp, q);
  }
}

// imsal, si3d_procedures.f90:3079-3207 (no point sources); HeatSource is
// zero unless the surface heat budget (ifsurfbc 1-3) sets it.
void OrcLake::imsal() {
  const double twodt1 = twodt * tz;
  std::vector<double> hn(km1 + 3), a1(km1 + 3), a2(km1 + 3), a3(km1 + 3), ds(km1 + 3), x(km1 + 3);
  for (int l = 1; l <= lm; ++l) {
    int kms = kmz[l], k1s = k1z[l], nwl = kms - k1s + 1;
    if (nwl < 1) continue;
    for (int k = k1s + 1; k <= kms; ++k) hn[k] = h(k, l);
    hn[k1s] = twodt1 * wp(k1s, l) + hpp(k1s, l);
    if (hpp(k1s, l) <= ZERO) hn[k1s + 1] = hpp(k1s + 1, l);
    if (nwl == 1) {
      double a = hn[k1s] / twodt1;
      sal(k1s, l) = (bex(k1s, l) + heat(k1s, l)) / a;
      if (h(k1s, l) < 1.E-2) sal(k1s, l) = salpp(k1s, l);
    } else {
      for (int k = k1s; k <= kms - 1; ++k) a3[k] = -Dv(k + 1, l) / (hn[k] + hn[k + 1]) * 2.;
      a3[kms] = 0.0;
      for (int k = k1s + 1; k <= kms; ++k) a1[k] = -Dv(k, l) / (hn[k - 1] + hn[k]) * 2.;
      a1[k1s] = 0.0;
      for (int k = k1s; k <= kms; ++k) {
        a2[k] = hn[k] / twodt1 - a1[k] - a3[k];
        ds[k] = bex(k, l) + heat(k, l);
      }
      trid_ef(nwl, &a1[k1s], &a2[k1s], &a3[k1s], &ds[k1s], &x[k1s]);
      for (int k = k1s; k <= kms; ++k) sal(k, l) = x[k];
      for (int k = k1; k <= k1s - 1; ++k) sal(k, l) = x[k1s];
      sal(kms + 1, l) = sal(kms, l);
    }
  }
}

// ImTracer, si3d_procedures.f90:4360-4484 (sourcesink = 0 as srcsnk00,
// no point sources): the imsal matrix without the thin-layer clamp, the
// sediment layer from its own Bex, negative values clipped to zero.
void OrcLake::imtracer(int nt) {
  F3 &tr = tracer[nt];
  const double twodt1 = twodt * tz;
  std::vector<double> hn(km1 + 3), a1(km1 + 3), a2(km1 + 3), a3(km1 + 3), ds(km1 + 3), x(km1 + 3);
  for (int l = 1; l <= lm; ++l) {
    int kms = kmz[l], k1s = k1z[l], nwl = kms - k1s + 1;
    for (int k = k1s + 1; k <= kms; ++k) hn[k] = h(k, l);
    hn[k1s] = twodt1 * wp(k1s, l) + hpp(k1s, l);
    if (hpp(k1s, l) <= ZERO) hn[k1s + 1] = hpp(k1s + 1, l);
    if (nwl == 1) {
      tr(k1s, l) = bex(k1s, l) / (hn[k1s] / twodt1);
    } else if (nwl >= 2) {
      for (int k = k1s; k <= kms - 1; ++k) a3[k] = -Dv(k + 1, l) / (hn[k] + hn[k + 1]) * 2.;
      a3[kms] = 0.0;
      for (int k = k1s + 1; k <= kms; ++k) a1[k] = -Dv(k, l) / (hn[k - 1] + hn[k]) * 2.;
      a1[k1s] = 0.0;
      for (int k = k1s; k <= kms; ++k) {
        a2[k] = hn[k] / twodt1 - a1[k] - a3[k];
        ds[k] = bex(k, l);
      }
      trid_ef(nwl, &a1[k1s], &a2[k1s], &a3[k1s], &ds[k1s], &x[k1s]);
      for (int k = k1s; k <= kms; ++k) tr(k, l) = x[k];
      for (int k = k1; k <= k1s - 1; ++k) tr(k, l) = x[k1s];
      tr(k1 - 1, l) = x[k1s];
    }
    // sediment layer (4449-4452)
    tr(kms + 1, l) = bex(kms + 1, l) / (hpp(kms + 1, l) / twodt1);
    for (int k = k1 - 1; k <= kms + 1; ++k)
      if (tr(k, l) < 0.0) tr(k, l) = 0.0;
  }
}

// vel2, si3d_procedures.f90:2570-2631.
void OrcLake::vel2() {
  layer_h();
  for (int l = 1; l <= lm; ++l) {
    if (maskE(l)) {
      int k = k1u[l];
      if (hu(k - 1, l) > ZERO) {
        double t = uh(k, l) / (hu(k, l) + hu(k - 1, l));
        uh(k, l) = t * hu(k, l);
        uh(k - 1, l) = t * hu(k - 1, l);
      }
    }
    if (maskN(l)) {
      int k = k1v[l];
      if (hv(k - 1, l) > ZERO) {
        double t = vh(k, l) / (hv(k, l) + hv(k - 1, l));
        vh(k, l) = t * hv(k, l);
        vh(k - 1, l) = t * hv(k - 1, l);
      }
    }
  }
}

// smooth (Asselin filter), si3d_procedures.f90:3356-3415.
void OrcLake::smooth() {
  for (int l = 1; l <= lm; ++l) sp[l] = sp[l] + beta2 * (s[l] - 2. * sp[l] + spp[l]);
  layer_hp2();
  for (int l = 1; l <= lm; ++l) {
    if (maskE(l)) {
      int kmx = std::min(kmz[l], kmz[lEC[l]]);
      for (int k = k1; k <= kmx; ++k) {
        if (hup(k, l) <= 0.0) continue;
        uhp(k, l) = uhp(k, l) + beta2 * (uh(k, l) - 2. * uhp(k, l) + uhpp(k, l));
        up(k, l) = uhp(k, l) / hup(k, l);
      }
    }
    if (maskN(l)) {
      int kmy = std::min(kmz[l], kmz[lNC[l]]);
      for (int k = k1; k <= kmy; ++k) {
        if (hvp(k, l) <= 0.0) continue;
        vhp(k, l) = vhp(k, l) + beta2 * (vh(k, l) - 2. * vhp(k, l) + vhpp(k, l));
        vp(k, l) = vhp(k, l) / hvp(k, l);
      }
    }
  }
}

// UpdateMixingCoefficients (si3d_mixing.f90:33-298, case >=1) with
// SolveTKEandLengthScale (301-551).  The basin energy diagnostics are outputs only.
void OrcLake::turbulence(int istep) {
  if (d.iturb <= 0) return;
  const double twodt1 = twodt / istep;
  const int NK = km1 + 3;
  std::vector<double> zfromb(NK), zfromt(NK), xl(NK), N2(NK), ShearP(NK), BuoyP(NK), dco(NK), a1(NK),
      a2(NK), a3(NK), ds(NK), x(NK), dz(NK);
  const double b1_23 = std::pow(B_1, 2. / 3.);
  if (istep > 1)
    for (int i = 0; i < 4; ++i) tke[i] = 0.0;  // 333-338
  for (int m = 1; m <= lm; ++m) {
    int kms = kmz[m], k1s = k1z[m], nwl = kms - k1s + 1;
    if (nwl <= 1) continue;
    double htot = 0.0;
    for (int k = k1s; k <= kms; ++k) htot += hp(k, m);
    double z0 = 0.0;
    for (int k = kms; k >= k1s + 1; --k) {
      z0 += hp(k, m);
      zfromb[k] = z0;
    }
    zfromb[k1s] = htot;
    zfromb[kms + 1] = 0.0;
    z0 = 0.0;
    for (int k = k1s + 1; k <= kms; ++k) {
      z0 += hp(k - 1, m);
      zfromt[k] = z0;
    }
    zfromt[k1s] = 0.0;
    zfromt[kms + 1] = htot;
    for (int k = k1s + 1; k <= kms; ++k) xl[k] = q2lp(k, m) / q2p(k, m);
    xl[k1s] = 0.0;
    xl[kms + 1] = 0.0;
    const int lW = lWC[m], lS = lSC[m];
    for (int k = k1s + 1; k <= kms; ++k) {
      dz[k] = 0.5 * (hp(k - 1, m) + hp(k, m));
      double rhoavg = 0.5 * (rhop(k - 1, m) + rhop(k, m)) + 1000.;
      double drhodz = (rhop(k - 1, m) - rhop(k, m)) / dz[k];
      N2[k] = -G / rhoavg * drhodz;
      double dudz = 0.5 * (up(k - 1, m) + up(k - 1, lW) - up(k, m) - up(k, lW)) / dz[k];
      double dvdz = 0.5 * (vp(k - 1, m) + vp(k - 1, lS) - vp(k, m) - vp(k, lS)) / dz[k];
      if (N2[k] > 0.0) xl[k] = std::min(xl[k], 0.53 * std::sqrt(q2p(k, m) / N2[k]));
      ShearP[k] = Av(k, m) * (dudz * dudz + dvdz * dvdz);
      BuoyP[k] = -Dv(k, m) * N2[k];
    }
    for (int k = k1s; k <= kms; ++k)
      dco[k] = SQ * 0.50 * (xl[k] + xl[k + 1]) * std::sqrt(0.50 * (q2p(k, m) + q2p(k + 1, m)));
    a3[k1s] = 0.0;
    for (int k = k1s + 1; k <= kms; ++k) a3[k] = -dco[k] / (hp(k, m) * 0.50 * (hp(k, m) + hp(k - 1, m)));
    a3[kms + 1] = 0.0;
    a1[kms + 1] = 0.0;
    for (int k = k1s + 1; k <= kms; ++k) a1[k] = -dco[k - 1] / (hp(k - 1, m) * 0.50 * (hp(k, m) + hp(k - 1, m)));
    a1[k1s] = 0.0;
    a2[k1s] = 3. / twodt1;
    for (int k = k1s + 1; k <= kms; ++k)
      a2[k] = 3. / twodt1 - a1[k] - a3[k] + 2. * (std::sqrt(q2p(k, m)) / (B_1 * xl[k]));
    a2[kms + 1] = 3. / twodt1;
    for (int k = k1s + 1; k <= kms; ++k) {
      ds[k] = 4. * q2p(k, m) / twodt1 - q2pp(k, m) / twodt1;
      if (BuoyP[k] >= 0.0) {
        ds[k] += 2. * (ShearP[k] + BuoyP[k]);
      } else {
        ds[k] += 2. * ShearP[k];
        a2[k] = a2[k] - BuoyP[k] * 2. / q2p(k, m);
      }
    }
    {
      double ub = 0.5 * (u(kms, m) + u(kms, lW)), vb = 0.5 * (v(kms, m) + v(kms, lS));
      double ustar = d.cd * (ub * ub + vb * vb);
      ds[kms + 1] = 3. * b1_23 * ustar / twodt1;
      double usurf = 0.5 * (up(k1u[m], m) + up(k1u[lW], lW)), vsurf = 0.5 * (vp(k1v[m], m) + vp(k1v[lS], lS));
      ustar = cdw[m] * RHOAIR * ((uair[m] - usurf) * (uair[m] - usurf) + (vair[m] - vsurf) * (vair[m] - vsurf)) / 1000.;
      ds[k1s] = 3. * b1_23 * ustar / twodt1;
    }
    trid_ef(nwl + 1, &a1[k1s], &a2[k1s], &a3[k1s], &ds[k1s], &x[k1s]);
    for (int k = k1s; k <= kms + 1; ++k) q2(k, m) = std::max(Q2_MIN, x[k]);
    for (int k = k1; k <= k1s - 1; ++k) q2(k, m) = q2(k1s, m);
    if (istep > 1)  // 466-475
      for (int k = k1s + 1; k <= kms; ++k) {
        const double rhoxh = (rhop(k, m) + 1000.) * hp(k, m);
        tke[0] += ShearP[k] * rhoxh;
        tke[1] += BuoyP[k] * rhoxh;
        tke[2] += q2(k, m) * std::sqrt(q2p(k, m)) / (B_1 * xl[k]) * rhoxh;
        tke[3] += q2(k, m) * rhoxh;
      }
    // length scale (488-537)
    for (int k = k1s + 1; k <= kms; ++k) {
      double r1 = xl[k] / (KAPPAS * zfromb[k]), r2 = xl[k] / (KAPPAS * zfromt[k]);
      a2[k] = 3. / twodt1 - a1[k] - a3[k] + std::sqrt(q2p(k, m)) / (B_1 * xl[k]) * (1. + E_2 * r1 * r1 + E_3 * r2 * r2);
    }
    a2[k1s] = 3. / twodt1;
    a2[kms + 1] = 3. / twodt1;
    for (int k = k1s + 1; k <= kms; ++k) {
      if (d.iturb == 1) {
        if (BuoyP[k] >= 0) {
          ds[k] = xl[k] * (E_1 * ShearP[k] + E_1 * BuoyP[k]);
        } else {
          ds[k] = xl[k] * (E_1 * ShearP[k]);
          a2[k] = a2[k] - E_1 * BuoyP[k] / q2p(k, m);
        }
      } else {
        if (BuoyP[k] >= 0) {
          ds[k] = xl[k] * (E_1 * ShearP[k] + 2.00 * BuoyP[k]);
        } else {
          ds[k] = xl[k] * (E_1 * ShearP[k]);
          a2[k] = a2[k] - 5.06 * BuoyP[k] / q2p(k, m);
        }
      }
      ds[k] = ds[k] + 4 * q2lp(k, m) / twodt1 - q2lpp(k, m) / twodt1;
    }
    ds[kms + 1] = 0.0;
    ds[k1s] = 0.0;
    trid_ef(nwl + 1, &a1[k1s], &a2[k1s], &a3[k1s], &ds[k1s], &x[k1s]);
    for (int k = k1s; k <= kms + 1; ++k) q2l(k, m) = std::max(Q2L_MIN, x[k]);
    for (int k = k1; k <= k1s - 1; ++k) q2l(k, m) = q2l(k1s, m);
  }
  if (istep == 1) {  // 539-548
    const double a = dx * dy * dt;
    for (int i = 0; i < 3; ++i) tke[i] *= a, tke[4 + i] += tke[i];
    tke[3] = 0.5 * tke[3] * a;
  }
  // eddy coefficients from q2p, q2lp (UpdateMixingCoefficients 245-290)
  for (int ll = 1; ll <= lm; ++ll) {
    int kms = kmz[ll], k1s = k1z[ll], nwl = kms - k1s + 1;
    if (nwl <= 1) continue;
    for (int k = k1s + 1; k <= kms; ++k) {
      double dz = 0.5 * (hp(k - 1, ll) + hp(k, ll));
      double rhoavg = 0.5 * (rhop(k - 1, ll) + rhop(k, ll)) + 1000.;
      double drhodz = (rhop(k - 1, ll) - rhop(k, ll)) / dz;
      double N2k = -G * drhodz / rhoavg;
      double xlk = q2lp(k, ll) / q2p(k, ll);
      if (N2k > 0.0) xlk = std::min(xlk, 0.53 * std::sqrt(q2p(k, ll) / N2k));
      double Gh = -N2k * xlk * xlk / q2p(k, ll);
      Gh = std::max(GH_MIN, std::min(GH_MAX, Gh));
      double Sh = t_1 / (1. - Gh * t_2);
      double Sm = (t_3 + Sh * Gh * t_4) / (1. - t_5 * Gh);
      double Rt = std::sqrt(q2p(k, ll)) * xlk;
      Av(k, ll) = Rt * Sm + AV_MOL;
      Dv(k, ll) = Rt * Sh + DV_MOL;
    }
    if (hp(k1s, ll) < 1.E-2) {
      Av(k1s + 1, ll) = std::max(Av(k1s + 1, ll), 1.E-4);
      Dv(k1s + 1, ll) = std::max(Dv(k1s + 1, ll), 1.E-4);
    }
    for (int k = 1; k <= k1s; ++k) Av(k, ll) = 0.0, Dv(k, ll) = 0.0;
    for (int k = kms + 1; k <= km1; ++k) Av(k, ll) = 0.0, Dv(k, ll) = 0.0;
  }
}

// fd supervisor, si3d_procedures.f90:382-575 (no open boundaries or point
// sources; conservative tracers only, ecomod = 0).
int OrcLake::fd(int n, int istep, int lastiter) {
  tz = 1.0 / istep;
  openbc_uvh();  // 423
  surfbc(n, istep);
  hmix();
  exmom(1);
  matmom(1);
  exmom(2);
  matmom(2);
  matcon();
  int rc = solver();
  openbc_uvh();  // 482
  vel();
  if (d.isal != 0) {
    exsal();
    imsal();
    openbc_scalar(sal, 1);  // openbcSCA, 515
  }
  // non-active scalars on the last iteration only (521-547)
  if (ntr > 0 && d.niter > 0 && lastiter == 1)
    for (int nt = 0; nt < ntr; ++nt) {
      exscalar(tracerpp[nt], true);
      imtracer(nt);
      openbc_scalar(tracer[nt], 2 + nt);  // openbcTracer, 544
    }
  vel2();
  if (d.ismooth >= 1 && istep == 1) smooth();
  turbulence(istep);
  return rc;
}

// settrap, si3d_procedures.f90:3418-3618 (+ settrap_2EqTVars, si3d_mixing.f90:727-771).
void OrcLake::settrap() {
  spp[lm1] = sp[lm1];
  for (int l = 1; l <= lm; ++l) {
    spp[l] = sp[l];
    sp[l] = 0.5 * (s[l] + spp[l]);
  }
  hpp = hp;
  hupp = hup;
  hvpp = hvp;
  layer_hp2();
  toplayer_indexp2();
  for (int l = 1; l <= lm; ++l) {
    for (int k = k1; k <= kmz[l]; ++k) {
      salpp(k, l) = salp(k, l);
      salp(k, l) = (sal(k, l) + salpp(k, l)) / 2.;
      rhop(k, l) = densty_s(salp(k, l), 0.00004, zdens(k, l)) - 1000.;
    }
    if (maskE(l)) {
      int kmx = std::min(kmz[l], kmz[lEC[l]]);
      for (int k = k1; k <= kmx; ++k) {
        uhpp(k, l) = uhp(k, l);
        upp(k, l) = up(k, l);
        if (hup(k, l) > ZERO) {
          uhp(k, l) = 0.5 * (uh(k, l) + uhpp(k, l));
          up(k, l) = uhp(k, l) / hup(k, l);
        } else {
          uhp(k, l) = 0.0;
          up(k, l) = 0.0;
        }
      }
      int k = k1u[l];
      if (hu(k - 1, l) > ZERO) {
        uhp(k, l) = uhp(k, l) + uh(k - 1, l) / 2.;
        up(k, l) = uhp(k, l) / hup(k, l);
      }
      if (hupp(k - 1, l) > ZERO) {
        uhp(k, l) = uhp(k, l) + uhpp(k - 1, l) / 2.;
        up(k, l) = uhp(k, l) / hup(k, l);
      }
    }
    if (maskN(l)) {
      int kmy = std::min(kmz[l], kmz[lNC[l]]);
      for (int k = k1; k <= kmy; ++k) {
        vhpp(k, l) = vhp(k, l);
        vpp(k, l) = vp(k, l);
        if (hvp(k, l) > ZERO) {
          vhp(k, l) = 0.5 * (vh(k, l) + vhpp(k, l));
          vp(k, l) = vhp(k, l) / hvp(k, l);
        } else {
          vhp(k, l) = 0.0;
          vp(k, l) = 0.0;
        }
      }
      int k = k1v[l];
      if (hv(k - 1, l) > ZERO) {
        vhp(k, l) = vhp(k, l) + vh(k - 1, l) / 2.;
        vp(k, l) = vhp(k, l) / hvp(k, l);
      }
      if (hvpp(k - 1, l) > ZERO) {
        vhp(k, l) = vhp(k, l) + vhpp(k - 1, l) / 2.;
        vp(k, l) = vhp(k, l) / hvp(k, l);
      }
    }
  }
  continuity(2);
  obc_shift(0);  // 3558-3604
  if (d.iturb > 0) {
    for (int m = 1; m <= lm; ++m) {
      for (int k = k1; k <= kmz[m]; ++k) {
        q2pp(k, m) = q2p(k, m);
        q2lpp(k, m) = q2lp(k, m);
        q2p(k, m) = 0.5 * (q2(k, m) + q2pp(k, m));
        q2lp(k, m) = 0.5 * (q2l(k, m) + q2lpp(k, m));
      }
      int k = k1z[m];
      if (hpp(k, m) <= ZERO) {
        q2lp(k + 1, m) = q2lp(k + 2, m);
        q2p(k + 1, m) = q2p(k + 2, m);
      }
    }
  }
}

// settrap2, si3d_procedures.f90:3944-4087 (+ settrap2_2EqTVars, si3d_mixing.f90:775-816).
void OrcLake::settrap2() {
  sp[lm1] = 0.5 * (s[lm1] + spp[lm1]);
  for (int l = 1; l <= lm; ++l) sp[l] = 0.5 * (s[l] + spp[l]);
  layer_hp2();
  toplayer_indexp2();
  for (int l = 1; l <= lm; ++l) {
    for (int k = k1; k <= kmz[l]; ++k) {
      salp(k, l) = (sal(k, l) + salpp(k, l)) / 2.;
      rhop(k, l) = densty_s(salp(k, l), 0.00004, zdens(k, l)) - 1000.;
    }
    if (maskE(l)) {
      int kmx = std::min(kmz[l], kmz[lEC[l]]);
      for (int k = k1; k <= kmx; ++k) {
        if (hup(k, l) > ZERO) {
          uhp(k, l) = 0.5 * (uh(k, l) + uhpp(k, l));
          up(k, l) = uhp(k, l) / hup(k, l);
        } else {
          uhp(k, l) = 0.0;
          up(k, l) = 0.0;
        }
      }
      int k = k1u[l];
      if (hu(k - 1, l) > ZERO) {
        uhp(k, l) = uhp(k, l) + uh(k - 1, l) / 2.;
        up(k, l) = uhp(k, l) / hup(k, l);
      }
      if (hupp(k - 1, l) > ZERO) {
        uhp(k, l) = uhp(k, l) + uhpp(k - 1, l) / 2.;
        up(k, l) = uhp(k, l) / hup(k, l);
      }
    }
    if (maskN(l)) {
      int kmy = std::min(kmz[l], kmz[lNC[l]]);
      for (int k = k1; k <= kmy; ++k) {
        if (hvp(k, l) > ZERO) {
          vhp(k, l) = 0.5 * (vh(k, l) + vhpp(k, l));
          vp(k, l) = vhp(k, l) / hvp(k, l);
        } else {
          vhp(k, l) = 0.0;
          vp(k, l) = 0.0;
        }
      }
      int k = k1v[l];
      if (hv(k - 1, l) > ZERO) {
        vhp(k, l) = vhp(k, l) + vh(k - 1, l) / 2.;
        vp(k, l) = vhp(k, l) / hvp(k, l);
      }
      if (hvpp(k - 1, l) > ZERO) {
        vhp(k, l) = vhp(k, l) + vhpp(k - 1, l) / 2.;
        vp(k, l) = vhp(k, l) / hvp(k, l);
      }
    }
  }
  continuity(2);
  if (d.iturb > 0) {
    for (int m = 1; m <= lm; ++m) {
      for (int k = k1; k <= kmz[m]; ++k) {
        q2p(k, m) = 0.5 * (q2(k, m) + q2pp(k, m));
        q2lp(k, m) = 0.5 * (q2l(k, m) + q2lpp(k, m));
      }
      int k = k1z[m];
      if (hpp(k, m) <= ZERO) {
        q2lp(k + 1, m) = q2lp(k + 2, m);
        q2p(k + 1, m) = q2p(k + 2, m);
      }
    }
  }
}

// save, si3d_procedures.f90:3621-3941 (+ save_2EqTvars, si3d_mixing.f90:819-889).
void OrcLake::save(int istep) {
  if (ntr > 0) tracerpp = tracer;  // 3772-3774, 3927-3929
  if (istep == 1) {
    spp[lm1] = spp[lm1];
    for (int l = 1; l <= lm; ++l) {
      spp[l] = sp[l];
      sp[l] = s[l];
    }
    sp[lm1] = s[lm1];
    hpp = hp;
    hupp = hup;
    hvpp = hvp;
  } else {
    for (int l = 1; l <= lm1; ++l) sp[l] = s[l];
  }
  hp = h;
  hup = hu;
  hvp = hv;
  toplayer_indexp2();
  for (int l = 1; l <= lm; ++l) {
    for (int k = k1; k <= km; ++k) {
      if (istep == 1) salpp(k, l) = salp(k, l);
      salp(k, l) = sal(k, l);
      rhop(k, l) = densty_s(salp(k, l), 0.00004, zdens(k, l)) - 1000.;
    }
    if (istep == 1)
      for (int k = k1; k <= km; ++k) {
        uhpp(k, l) = uhp(k, l);
        vhpp(k, l) = vhp(k, l);
        upp(k, l) = up(k, l);
        vpp(k, l) = vp(k, l);
      }
    for (int k = k1; k <= km; ++k) {
      if (hup(k, l) > ZERO) {
        uhp(k, l) = uh(k, l);
        up(k, l) = uhp(k, l) / hup(k, l);
        u(k, l) = up(k, l);
      } else {
        uhp(k, l) = up(k, l) = u(k, l) = 0.0;
      }
      if (hvp(k, l) > ZERO) {
        vhp(k, l) = vh(k, l);
        vp(k, l) = vhp(k, l) / hvp(k, l);
        v(k, l) = vp(k, l);
      } else {
        vhp(k, l) = vp(k, l) = v(k, l) = 0.0;
      }
    }
  }
  continuity(2);
  obc_shift(istep == 1 ? 1 : 2);  // 3722-3768 (istep 2), 3872-3918 (istep 1)
  if (d.iturb > 0) {
    for (int m = 1; m <= lm1; ++m) {
      for (int k = 0; k <= km1 + 1; ++k) {
        if (istep == 1) {
          q2pp(k, m) = q2p(k, m);
          q2lpp(k, m) = q2lp(k, m);
        }
        q2p(k, m) = q2(k, m);
        q2lp(k, m) = q2l(k, m);
      }
      if (m == lm1) continue;
      int k1s = k1z[m];
      bool guard = (istep == 1) ? true : (km1 - k1s > 1);
      if (guard && hpp(k1s, m) <= ZERO) {
        q2lp(k1s + 1, m) = q2lp(k1s + 2, m);
        q2p(k1s + 1, m) = q2p(k1s + 2, m);
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Open boundaries (si3d_boundaryconditions.f90), types 1 and 2 on the four
// sides; one sweep over the boundaries in input order, as the reference's
// single-thread loops do (si3d.f90:196-246 give each boundary its full range
// with one thread).

// Boundary value q of boundary nn at thrs: parab over records dtsecOpenBC
// apart (938-939, 3187-3188).  Past the last record the end parabola is
// extended (the reference reads beyond the series there).
double OrcLake::obc_value(int nn, int q) const {
  const double *fx = &obc_vals[((size_t)nn * obc_nv + q) * (obc_np + 1)];
  const double dxs = obc_dtsec / 3600.;
  int m = (int)(thrs / dxs);
  double theta = (thrs - (double)m * dxs) / dxs;
  if (m == 0) {
    m = 2;
    theta -= 1.0;
  } else {
    m = m + 1;
  }
  if (m + 1 > obc_np) m = obc_np - 1;
  return fx[m] + 0.5 * theta * (fx[m + 1] - fx[m - 1] + theta * (fx[m + 1] + fx[m - 1] - 2.0 * fx[m]));
}

// Thickness and transport of the boundary cells of a flow boundary for flow
// qbc spread uniformly over the section (openbc0 420-647, openbcUVH 976-1176).
// The free surface of the section is taken from the column one past its end
// (the Fortran loop index after the loop); the south side sums hvSB where the
// reference sums hvNB (1170), which divides by zero without a north boundary.
void OrcLake::flow_section(const Obc &b, double qbc, bool first) {
  const int sd = b.side;
  const bool we = (sd == 0 || sd == 2);
  double hbc = -1.e6;
  for (int a = (we ? b.js : b.is); a <= (we ? b.je : b.ie); ++a) {
    int l = we ? IJ(b.is, a) : IJ(a, b.js);
    if (hhs[l] > hbc) hbc = hhs[l];
  }
  int lpast = we ? IJ(b.is, b.je + 1) : IJ(b.ie + 1, b.js);
  double sbc = std::max(s[lpast], -hbc + d.dzmin);
  double areatot = 0.0;
  const int a0 = we ? b.js : b.is, a1 = we ? b.je : b.ie;
  // W/E zero the thickness first; N/S do so only in openbc0 (1089, 1146)
  if (we || first)
    for (int a = a0; a <= a1; ++a)
      for (int k = 1; k <= km1; ++k) BH(sd, 0, k, a) = ZERO;
  for (int a = a0; a <= a1; ++a) {
    int l = we ? IJ(b.is, a) : IJ(a, b.js);
    for (int k = k1; k <= kmz[l]; ++k) {
      double hh = std::min(zlevel[k + 1], hhs[l]) - std::max(zlevel[k], -sbc);
      if (hh <= HMIN) hh = ZERO;
      BH(sd, 0, k, a) = hh;
      areatot = areatot + hh * dy;  // dy on every side (440-647, 990-1170)
    }
  }
  double uavg = qbc / areatot;
  for (int a = a0; a <= a1; ++a)
    for (int k = 1; k <= km1; ++k) BQ(sd, 0, k, a) = 0.0;
  for (int a = a0; a <= a1; ++a) {
    int l = we ? IJ(b.is, a) : IJ(a, b.js);
    int kf = (first && sd == 1) ? k1z[l] : k1;  // openbc0 north: from k1z (600)
    for (int k = kf; k <= kmz[l]; ++k) BQ(sd, 0, k, a) = uavg * BH(sd, 0, k, a);
  }
}

// openbc0, si3d_boundaryconditions.f90:23-897 for types 1 and 2: checks that
// every boundary cell is wet with a dry outer neighbour, then the t = 0
// values (sp of water-level boundaries, the flow section of flow boundaries).
bool OrcLake::openbc0() {
  const int nopen = std::max(0, d.nopen);
  obc.clear();
  if (nopen == 0 || !d.openbc) return true;
  obc_np = d.openbc_npts;
  obc_nv = ntr + 2;
  obc_dtsec = d.dtsec_openbc;
  if (obc_np < 3 || !d.openbc_values || obc_dtsec <= 0.0) return false;
  obc_vals.assign((size_t)nopen * obc_nv * (obc_np + 1), 0.0);
  for (int nn = 0; nn < nopen; ++nn)
    for (int r = 0; r < obc_np; ++r)
      for (int q = 0; q < obc_nv; ++q)
        obc_vals[((size_t)nn * obc_nv + q) * (obc_np + 1) + r + 1] =
            d.openbc_values[((size_t)nn * obc_np + r) * obc_nv + q];
  const size_t na = (size_t)(std::max(im1, jm1) + 2) * (km1 + 2);
  for (int sd = 0; sd < 4; ++sd)
    for (int lev = 0; lev < 3; ++lev) bq[sd][lev].assign(na, 0.0), bh[sd][lev].assign(na, 0.0);
  for (int nn = 0; nn < nopen; ++nn) {
    const int *o6 = d.openbc + 6 * nn;
    Obc b{o6[0] - 1, o6[1], o6[2], o6[3], o6[4], o6[5]};
    if (b.side < 0 || b.side > 3 || (b.type != 1 && b.type != 2)) return false;
    const bool we = (b.side == 0 || b.side == 2);
    // one column (W/E) or one row (N/S); the range runs along it
    if (we ? (b.ie < b.is || b.je < b.js) : (b.je < b.js || b.ie < b.is)) return false;
    if (we ? (b.is < 1 || b.is > im1 || b.js < 1 || b.je > jm1) : (b.js < 1 || b.js > jm1 || b.is < 1 || b.ie > im1))
      return false;
    for (int a = (we ? b.js : b.is); a <= (we ? b.je : b.ie); ++a) {
      int i = we ? b.is : a, j = we ? a : b.js;
      int io = i + (b.side == 0 ? -1 : b.side == 2 ? 1 : 0), jo = j + (b.side == 1 ? 1 : b.side == 3 ? -1 : 0);
      if (!M2(i, j) || M2(io, jo)) return false;  // "not a bdry. point"
      int l = IJ(i, j);
      if (b.type == 1 && kmz[l] - k1z[l] + 1 < 1) return false;  // dry point on boundary
    }
    obc.push_back(b);
  }
  // t = 0 values (337-647); the n+1 arrays start at zero (334-339)
  for (auto &b : obc) {
    const bool we = (b.side == 0 || b.side == 2);
    const double v0 = obc_vals[((size_t)(&b - &obc[0]) * obc_nv) * (obc_np + 1) + 1];
    if (b.type == 1) {
      for (int a = (we ? b.js : b.is); a <= (we ? b.je : b.ie); ++a) sp[we ? IJ(b.is, a) : IJ(a, b.js)] = v0;
    } else {
      flow_section(b, v0, true);
    }
  }
  return true;
}

// openbcUVH, si3d_boundaryconditions.f90:900-1398 (types 1 and 2).
void OrcLake::openbc_uvh() {
  for (size_t nn = 0; nn < obc.size(); ++nn) {
    const Obc &b = obc[nn];
    const bool we = (b.side == 0 || b.side == 2);
    double v = obc_value((int)nn, 0);
    if (b.type == 1) {
      for (int a = (we ? b.js : b.is); a <= (we ? b.je : b.ie); ++a) s[we ? IJ(b.is, a) : IJ(a, b.js)] = v;
    } else {
      flow_section(b, v, false);
    }
  }
}

// MODqqddrr4openBC, si3d_boundaryconditions.f90:1508-1690 (types 1 and 2):
// a water-level boundary moves its coupling to the known zeta into qq of the
// inner neighbour and drops it from the matrix; a flow boundary adds the
// boundary inflow at n+1 and n-1 to qq of the boundary cell.
void OrcLake::mod_qq() {
  const double dtdx1 = dtdx * tz, dtdy1 = dtdy * tz;
  for (const Obc &b : obc) {
    const bool we = (b.side == 0 || b.side == 2);
    const int a0 = we ? b.js : b.is, a1 = we ? b.je : b.ie;
    if (b.type == 1) {
      for (int a = a0; a <= a1; ++a) {
        int l = we ? IJ(b.is, a) : IJ(a, b.js);
        int li = b.side == 0 ? lEC[l] : b.side == 1 ? lSC[l] : b.side == 2 ? lWC[l] : lNC[l];  // inner neighbour
        double cpl = b.side == 0 ? sx[l] : b.side == 1 ? sy[li] : b.side == 2 ? sx[li] : sy[l];
        if (li != lm1) qq[li] = qq[li] + cpl * s[l];
      }
      for (int a = a0; a <= a1; ++a) {
        int l = we ? IJ(b.is, a) : IJ(a, b.js);
        if (b.side == 0) sx[l] = 0.0;
        else if (b.side == 1) { if (lSC[l] != lm1) sy[lSC[l]] = 0.0; }
        else if (b.side == 2) { if (lWC[l] != lm1) sx[lWC[l]] = 0.0; }
        else sy[l] = 0.0;
      }
    } else {
      for (int a = a0; a <= a1; ++a) {
        int l = we ? IJ(b.is, a) : IJ(a, b.js);
        double s0 = 0.0, s2 = 0.0;
        for (int k = k1; k <= kmz[l]; ++k) s0 += BQ(b.side, 0, k, a);
        for (int k = k1; k <= kmz[l]; ++k) s2 += BQ(b.side, 2, k, a);
        switch (b.side) {
          case 0: qq[l] = qq[l] + dtdx1 * s0 + dtdx1 * s2; break;
          case 1: qq[l] = qq[l] - dtdy1 * s0 - dtdy1 * s2; break;
          case 2: qq[l] = qq[l] - dtdx1 * s0 - dtdx1 * s2; break;
          default: qq[l] = qq[l] + dtdy1 * s0 + dtdy1 * s2; break;
        }
      }
    }
  }
}

// MODexmom4openBCX / MODexmom4openBCY, si3d_boundaryconditions.f90:1801-2936
// (types 1 and 2): the explicit term of the first momentum face inside the
// boundary is recomputed from upwind advection only, with the boundary
// transport as the outer neighbour of a flow boundary.  West and east faces
// in x, north and south faces in y; the east water-level face keeps ex = 0
// (its assignment is commented out in the reference, 2070).
void OrcLake::mod_exmom(int ieq) {
  const double twodt1 = twodt * tz;
  const bool x = (ieq == 1);
  auto P = [](double a) { return a + std::fabs(a); };
  auto M = [](double a) { return a - std::fabs(a); };
  for (const Obc &b : obc) {
    if (x != (b.side == 0 || b.side == 2)) continue;
    const int a0 = x ? b.js : b.is, a1 = x ? b.je : b.ie;
    for (int a = a0; a <= a1; ++a) {
      int lb = x ? IJ(b.is, a) : IJ(a, b.js);
      // the face column: the boundary cell (W, S) or its inner neighbour (E, N)
      int l = (b.side == 0 || b.side == 3) ? lb : (b.side == 2 ? IJ(b.is - 1, a) : IJ(a, b.js - 1));
      if (l == lm1) continue;  // no wet cell inside (the reference would write the sentinel column)
      if (b.type == 2) {
        if (x && !M2(l2i[l] + 1, l2j[l])) continue;
        if (!x && !M2(l2i[l], l2j[l] + 1)) continue;
      }
      int kmq = x ? std::min(kmz[lEC[l]], kmz[l]) : std::min(kmz[lNC[l]], kmz[l]);
      int k1q = x ? k1u[l] : k1v[l];
      if (kmq - k1q + 1 < 1) continue;
      for (int k = 1; k <= km1; ++k) ex(k, l) = 0.0;
      if (b.type == 1 && b.side == 2) continue;
      const int lE = lEC[l], lW = lWC[l], lN = lNC[l], lS = lSC[l];
      for (int k = k1q; k <= kmq; ++k) {
        double bq0 = 0.0, bdry = 0.0;
        if (b.type == 2) {
          bq0 = (BQ(b.side, 0, k, a) + BQ(b.side, 2, k, a)) / 2.;
          bdry = (BH(b.side, 2, k, a) > ZERO) ? BQ(b.side, 2, k, a) / BH(b.side, 2, k, a) : 0.0;
        }
        double adv;
        if (x) {
          double uE, uW;
          double vN = vhp(k, lE) + vhp(k, l), vS = vhp(k, lSC[lE]) + vhp(k, lS);
          double wU = wp(k, lE) + wp(k, l), wD = wp(k + 1, lE) + wp(k + 1, l);
          double hx;
          if (b.side == 0) {
            uE = uhp(k, lE) + uhp(k, l);
            uW = uhp(k, l) + (b.type == 2 ? bq0 : uhp(k, l));
            double uout = (b.type == 2) ? bdry : upp(k, l);
            hx = P(uE) * upp(k, l) + M(uE) * upp(k, lE) - P(uW) * uout - M(uW) * upp(k, l);
          } else {  // east flow boundary
            uE = uhp(k, l) + bq0;
            uW = uhp(k, lW) + uhp(k, l);
            hx = P(uE) * upp(k, l) + M(uE) * bdry - P(uW) * upp(k, lW) - M(uW) * upp(k, l);
          }
          adv = hx / fourdx + (P(vN) * upp(k, l) + M(vN) * upp(k, lN) - P(vS) * upp(k, lS) - M(vS) * upp(k, l)) / fourdy;
          adv = adv + (P(wU) * upp(k, l) + M(wU) * upp(k - 1, l)) / 4. - (P(wD) * upp(k + 1, l) + M(wD) * upp(k, l)) / 4.;
          ex(k, l) = uhpp(k, l) - twodt1 * (adv * d.iadv);
        } else {
          double uE = uhp(k, lN) + uhp(k, l), uW = uhp(k, lWC[lN]) + uhp(k, lW);
          double wU = wp(k, lN) + wp(k, l), wD = wp(k + 1, lN) + wp(k + 1, l);
          double vN, vS, hy;
          if (b.side == 1) {  // face south of a north boundary
            vN = vhp(k, l) + (b.type == 2 ? bq0 : vhp(k, l));
            vS = vhp(k, l) + vhp(k, lS);
            double vout = (b.type == 2) ? bdry : vpp(k, l);
            hy = P(vN) * vpp(k, l) + M(vN) * vout - P(vS) * vpp(k, lS) - M(vS) * vpp(k, l);
          } else {  // face north of a south boundary
            vN = vhp(k, lN) + vhp(k, l);
            vS = vhp(k, l) + (b.type == 2 ? bq0 : vhp(k, l));
            double vout = (b.type == 2) ? bdry : vpp(k, l);
            hy = P(vN) * vpp(k, l) + M(vN) * vpp(k, lN) - P(vS) * vout - M(vS) * vpp(k, l);
          }
          adv = (P(uE) * vpp(k, l) + M(uE) * vpp(k, lE) - P(uW) * vpp(k, lW) - M(uW) * vpp(k, l)) / fourdx + hy / fourdy;
          adv = adv + (P(wU) * vpp(k, l) + M(wU) * vpp(k - 1, l)) / 4. - (P(wD) * vpp(k + 1, l) + M(wD) * vpp(k, l)) / 4.;
          ex(k, l) = vhpp(k, l) - twodt1 * (adv * d.iadv);
        }
      }
    }
  }
}

// MODvel4openBC, si3d_boundaryconditions.f90:2941-3154 (types 1 and 2), at
// the end of every continuity call.
void OrcLake::mod_vel() {
  for (const Obc &b : obc) {
    const bool we = (b.side == 0 || b.side == 2);
    for (int a = (we ? b.js : b.is); a <= (we ? b.je : b.ie); ++a) {
      int l = we ? IJ(b.is, a) : IJ(a, b.js);
      if (b.type == 1) {
        F3 &t = we ? vh : uh;
        int ln = b.side == 0 ? lEC[l] : b.side == 1 ? lSC[l] : b.side == 2 ? lWC[l] : lNC[l];
        for (int k = 1; k <= km1; ++k) t(k, l) = 0.0;
        for (int k = 1; k <= km1; ++k) wp(k, l) = wp(k, ln);
      } else {
        int kms = kmz[l];
        const double den = we ? twodx : twody;
        wp(kms + 1, l) = 0.0;
        double dru = 0.0;
        for (int k = kms; k >= k1; --k) {
          double q0 = BQ(b.side, 0, k, a), q2 = BQ(b.side, 2, k, a);
          if (b.side == 0 || b.side == 3)
            dru = dru - (-q0 - q2) / den;
          else
            dru = dru - (q0 + q2) / den;
          wp(k, l) = wp(k, l) + dru;
        }
      }
    }
  }
}

// openbcSCA / openbcTracer, si3d_boundaryconditions.f90:3157-3311 and
// 3661-3806: on a water-level boundary the scalar takes the boundary value
// where the flow enters and the inner neighbour's value elsewhere.
void OrcLake::openbc_scalar(F3 &c, int q) {
  for (size_t nn = 0; nn < obc.size(); ++nn) {
    const Obc &b = obc[nn];
    if (b.type != 1) continue;
    const double sbc = obc_value((int)nn, q);
    const bool we = (b.side == 0 || b.side == 2);
    for (int a = (we ? b.js : b.is); a <= (we ? b.je : b.ie); ++a) {
      int l = we ? IJ(b.is, a) : IJ(a, b.js);
      for (int k = k1z[l]; k <= kmz[l]; ++k) {
        bool in;
        int ln;
        switch (b.side) {
          case 0: in = uh(k, l) > 0.0; ln = lEC[l]; break;
          case 1: in = vh(k, lSC[l]) < 0.0; ln = lSC[l]; break;
          case 2: in = uh(k, lWC[l]) < 0.0; ln = lWC[l]; break;
          default: in = vh(k, l) > 0.0; ln = lNC[l]; break;
        }
        c(k, l) = (in && sbc > 0.0) ? sbc : c(k, ln);
      }
    }
  }
}

// MODexsal4openbc / MODexTracer4openbc, si3d_boundaryconditions.f90:3314-3658
// and 3809-4137 (type 2): the advective flux through the boundary face, with
// the boundary value on inflow and the cell's own value on outflow.
void OrcLake::mod_exscalar(const F3 &cpp, int q) {
  for (size_t nn = 0; nn < obc.size(); ++nn) {
    const Obc &b = obc[nn];
    if (b.type != 2) continue;
    const double sbc = obc_value((int)nn, q);
    const bool we = (b.side == 0 || b.side == 2);
    for (int a = (we ? b.js : b.is); a <= (we ? b.je : b.ie); ++a) {
      int l = we ? IJ(b.is, a) : IJ(a, b.js);
      for (int k = k1z[l]; k <= kmz[l]; ++k) {
        double f = BQ(b.side, 0, k, a) + BQ(b.side, 2, k, a);
        switch (b.side) {
          case 0: bex(k, l) = bex(k, l) + f * ((f >= 0.0 && sbc > 0.0) ? sbc : cpp(k, l)) / twodx; break;
          case 1: bex(k, l) = bex(k, l) - f * ((f < 0.0 && sbc > 0.0) ? sbc : cpp(k, l)) / twody; break;
          case 2: bex(k, l) = bex(k, l) - f * ((f < 0.0 && sbc > 0.0) ? sbc : cpp(k, l)) / twodx; break;
          default: bex(k, l) = bex(k, l) + f * ((f > 0.0 && sbc > 0.0) ? sbc : cpp(k, l)) / twody; break;
        }
      }
    }
  }
}

// Time levels of the boundary arrays: settrap (mode 0, 3558-3604), save with
// istep 1 (mode 1, 3872-3918) and istep 2 (mode 2, 3722-3768).  Each boundary
// shifts both arrays of its orientation (W and E, or N and S) over its range,
// so overlapping ranges shift twice, as in the reference.
void OrcLake::obc_shift(int mode) {
  for (const Obc &b : obc) {
    const bool we = (b.side == 0 || b.side == 2);
    const int sides[2] = {we ? 2 : 1, we ? 0 : 3};  // E and W, or N and S (3569-3572, 3581-3584)
    for (int a = (we ? b.js : b.is); a <= (we ? b.je : b.ie); ++a)
      for (int t = 0; t < 2; ++t) {
        const int sd = sides[t];
        for (int k = 1; k <= km1; ++k) {
          if (mode <= 1) {
            BQ(sd, 2, k, a) = BQ(sd, 1, k, a);
            BH(sd, 2, k, a) = BH(sd, 1, k, a);
          }
          if (mode == 0) {
            BQ(sd, 1, k, a) = (BQ(sd, 2, k, a) + BQ(sd, 0, k, a)) / 2.;
            BH(sd, 1, k, a) = (BH(sd, 1, k, a) + BH(sd, 0, k, a)) / 2.;
          } else {
            BQ(sd, 1, k, a) = BQ(sd, 0, k, a);
            BH(sd, 1, k, a) = BH(sd, 0, k, a);
          }
        }
      }
  }
}

// ---------------------------------------------------------------------------
extern "C" {

OrcLake *orc_create(const OrcDeck *deck) {
  OrcLake *o = new OrcLake();
  if (!o->build(deck)) {
    delete o;
    return nullptr;
  }
  return o;
}

void orc_destroy(OrcLake *o) { delete o; }

void orc_dims(const OrcLake *o, int *out) {
  out[0] = o->im;
  out[1] = o->jm;
  out[2] = o->km1;
  out[3] = o->lm;
  out[4] = o->k1;
  out[5] = 0;
  out[6] = o->im1;
  out[7] = o->jm1;
}

// One time step of the si3d.f90 loop (324-393).
int orc_step(OrcLake *o, int n) {
  o->thrs = (double)n * o->dt / 3600.;
  int niter1 = o->d.niter, lastiter = 0, rc = 0, istep;
  const int itrap = o->d.itrap;
  bool trap = true;
  if (n == 1) {
    niter1 = std::max(o->d.niter, 2);
  } else {
    istep = 1;
    bool lf_only = (itrap == 0) || (n % std::max(itrap, 1) != 0);
    if (lf_only) lastiter = 1;
    rc |= o->fd(n, istep, lastiter);
    if (lf_only) {
      o->save(1);
      return rc;
    }
    o->settrap();
  }
  if (trap) {
    istep = 2;
    for (int iter = 1; iter <= niter1; ++iter) {
      if (iter == niter1) lastiter = 1;
      rc |= o->fd(n, istep, lastiter);
      if (iter < niter1) o->settrap2();
    }
    o->save(2);
  }
  return rc;
}

int orc_fd(OrcLake *o, int n, int istep, int lastiter, double thrs) {
  o->thrs = thrs;
  return o->fd(n, istep, lastiter);
}
void orc_settrap(OrcLake *o) { o->settrap(); }
void orc_settrap2(OrcLake *o) { o->settrap2(); }
void orc_save(OrcLake *o, int istep) { o->save(istep); }

int orc_phase(OrcLake *o, int phase, int n, int istep, double thrs) {
  o->thrs = thrs;
  o->tz = 1.0 / istep;
  switch (phase) {
    case 0: o->openbc_uvh(); o->surfbc(n, istep); break;
    case 1: o->hmix(); break;
    case 2: o->exmom(1); break;
    case 3: o->matmom(1); break;
    case 4: o->exmom(2); break;
    case 5: o->matmom(2); break;
    case 6: o->matcon(); break;
    case 7: {
      int rc = o->solver();
      o->openbc_uvh();
      return rc;
    }
    case 8: o->vel(); break;
    case 9: o->exsal(); break;
    case 10: o->imsal(); o->openbc_scalar(o->sal, 1); break;
    case 11: o->vel2(); break;
    case 12: o->smooth(); break;
    case 13: o->turbulence(istep); break;
    default: return -1;
  }
  return 0;
}

static const F3 *field3(const OrcLake *o, const std::string &n) {
  OrcLake *m = const_cast<OrcLake *>(o);
  for (int nt = 0; nt < m->ntr; ++nt) {
    if (n == "tracer" + std::to_string(nt + 1)) return &m->tracer[nt];
    if (n == "tracerpp" + std::to_string(nt + 1)) return &m->tracerpp[nt];
  }
#define F(x) if (n == #x) return &m->x;
  F(h) F(hp) F(hpp) F(hu) F(hup) F(hupp) F(hv) F(hvp) F(hvpp) F(uh) F(uhp) F(uhpp) F(u) F(up) F(upp)
  F(vh) F(vhp) F(vhpp) F(v) F(vp) F(vpp) F(wp) F(sal) F(salp) F(salpp) F(rhop) F(Av) F(Dv) F(kh)
  F(ex) F(agx) F(arx) F(agy) F(ary) F(q2) F(q2p) F(q2pp) F(q2l) F(q2lp) F(q2lpp) F(bex)
  if (n == "HeatSource") return &m->heat;
  if (n == "QswFr") return &m->qswfr;
#undef F
  return nullptr;
}

static std::vector<double> *field2(const OrcLake *o, const std::string &n) {
  OrcLake *m = const_cast<OrcLake *>(o);
#define F(x) if (n == #x) return &m->x;
  F(s) F(sp) F(spp) F(sx) F(sy) F(dd) F(qq) F(rr) F(eagx) F(earx) F(eagy) F(eary) F(uair) F(vair) F(cdw)
  F(hhs) F(hhu) F(hhv)
#undef F
  return nullptr;
}

static std::vector<int> *fieldi(const OrcLake *o, const std::string &n) {
  OrcLake *m = const_cast<OrcLake *>(o);
#define F(x) if (n == #x) return &m->x;
  F(kmz) F(k1z) F(k1u) F(k1v) F(l2i) F(l2j) F(lEC) F(lWC) F(lNC) F(lSC)
#undef F
  return nullptr;
}

long orc_get(const OrcLake *o, const char *name, double *out) {
  std::string n(name);
  if (const F3 *f = field3(o, n)) {
    for (int l = 1; l <= o->lm1; ++l)
      for (int k = 1; k <= o->km1; ++k) out[(size_t)(l - 1) * o->km1 + (k - 1)] = (*f)(k, l);
    return (long)o->km1 * o->lm1;
  }
  if (auto *a = field2(o, n)) {
    for (int l = 1; l <= o->lm1; ++l) out[l - 1] = (*a)[l];
    return o->lm1;
  }
  if (auto *a = fieldi(o, n)) {
    for (int l = 1; l <= o->lm1; ++l) out[l - 1] = (*a)[l];
    return o->lm1;
  }
  if (n == "zlevel") {
    for (int k = 1; k <= o->km1; ++k) out[k - 1] = o->zlevel[k];
    return o->km1;
  }
  if (n == "mask2d") {
    for (int i = 1; i <= o->im1; ++i)
      for (int j = 1; j <= o->jm1; ++j) out[(size_t)(i - 1) * o->jm1 + (j - 1)] = o->M2(i, j);
    return (long)o->im1 * o->jm1;
  }
  if (n == "thrs") {
    out[0] = o->thrs;
    return 1;
  }
  if (n == "energy_budget") {
    for (int i = 0; i < 7; ++i) out[i] = o->tke[i];
    return 7;
  }
  return -1;
}

long orc_set(OrcLake *o, const char *name, const double *in) {
  std::string n(name);
  if (const F3 *fc = field3(o, n)) {
    F3 *f = const_cast<F3 *>(fc);
    for (int l = 1; l <= o->lm1; ++l)
      for (int k = 1; k <= o->km1; ++k) (*f)(k, l) = in[(size_t)(l - 1) * o->km1 + (k - 1)];
    return (long)o->km1 * o->lm1;
  }
  if (auto *a = field2(o, n)) {
    for (int l = 1; l <= o->lm1; ++l) (*a)[l] = in[l - 1];
    return o->lm1;
  }
  if (auto *a = fieldi(o, n)) {
    for (int l = 1; l <= o->lm1; ++l) (*a)[l] = (int)in[l - 1];
    return o->lm1;
  }
  if (n == "energy_budget") {
    for (int i = 0; i < 7; ++i) o->tke[i] = in[i];
    return 7;
  }
  return -1;
}

void orc_solver_stats(const OrcLake *o, double *out3) {
  out3[0] = o->last_iters;
  out3[1] = o->last_relres;
  out3[2] = (double)o->total_iters;
}

double orc_densty(double t, double s, double e) { return densty_s(t, s, e); }

double orc_tvd_flux(double vel, double s1, double s2, double s3, double s4, double cfdef, double courant) {
  return tvd_flux(vel, s1, s2, s3, s4, cfdef, courant);
}

void orc_trid(int n, const double *a, const double *b, const double *c, const double *g, const double *r,
              double *ag, double *ar) {
  trid2(n, a, b, c, g, r, ag, ar);
}

void orc_trid1(int n, const double *a, const double *b, const double *c, const double *d, double *x) {
  trid_ef(n, a, b, c, d, x);
}

double orc_parab(double frstpt, double x, const double *fx, int nfx, double dxs) {
  std::vector<double> f(nfx + 2, 0.0);
  for (int i = 0; i < nfx; ++i) f[i + 1] = fx[i];
  return parab(frstpt, x, f.data(), dxs);
}

}  // extern "C"
