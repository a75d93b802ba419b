// s3g_host.hip -- host side of libsi3dgpu.so: the C-ABI, device allocation and
// layout marshalling, the time-loop driver (si3d.f90:324-393), wind forcing,
// timers, and the standalone deck path that restates input/bathy/init.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "s3g_internal.h"

namespace s3g {

static thread_local std::string g_err;
void set_error(const std::string &m) { g_err = m; }

int hip_check(hipError_t e, const char *what) {
  if (e == hipSuccess) return 0;
  set_error(std::string(what) + ": " + hipGetErrorString(e));
  return S3G_ERR_HIP;
}

template <typename T>
static T *dalloc(s3g_ctx *c, size_t n) {
  void *p = nullptr;
  if (hipMalloc(&p, std::max<size_t>(n, 1) * sizeof(T)) != hipSuccess) return nullptr;
  hipMemset(p, 0, std::max<size_t>(n, 1) * sizeof(T));
  c->allocs.push_back(p);
  return (T *)p;
}

// ---------------------------------------------------------------------------
// Timers: an event pair around each timed launch, resolved lazily.
void timer_begin(s3g_ctx *c, const char *name, int *slot) {
  int ti = -1;
  for (size_t i = 0; i < c->timers.size(); ++i)
    if (c->timers[i].name == name) ti = (int)i;
  if (ti < 0) {
    c->timers.push_back(Timer{name, 0.0, 0});
    ti = (int)c->timers.size() - 1;
  }
  if (c->pending.size() >= 4096) {
    hipStreamSynchronize(c->stream);
    for (auto &pr : c->pending) {
      float ms = 0.f;
      hipEventElapsedTime(&ms, c->ev_pool[pr.second].first, c->ev_pool[pr.second].second);
      c->timers[pr.first].total_ms += ms;
      c->timers[pr.first].n += 1;
    }
    c->pending.clear();
  }
  const int ei = (int)c->pending.size();
  while ((int)c->ev_pool.size() <= ei) {
    hipEvent_t a, b;
    hipEventCreate(&a);
    hipEventCreate(&b);
    c->ev_pool.push_back({a, b});
  }
  hipEventRecord(c->ev_pool[ei].first, c->stream);
  c->pending.push_back({ti, ei});
  *slot = ei;
}

void timer_end(s3g_ctx *c, int slot) {
  if (slot >= 0) hipEventRecord(c->ev_pool[slot].second, c->stream);
}

static void timers_flush(s3g_ctx *c) {
  hipStreamSynchronize(c->stream);
  for (auto &pr : c->pending) {
    float ms = 0.f;
    hipEventElapsedTime(&ms, c->ev_pool[pr.second].first, c->ev_pool[pr.second].second);
    c->timers[pr.first].total_ms += ms;
    c->timers[pr.first].n += 1;
  }
  c->pending.clear();
}

// Parabolic interpolation of a uniform series, si3d_utils.f90:895-920.
static double parab(double frstpt, double x, const std::vector<double> &fx1, double dx) {
  // fx1 is 1-based (fx1[0] unused)
  int m = (int)((x - frstpt) / dx);
  double om = m;
  double theta = (x - frstpt - om * dx) / dx;
  if (m == 0) {
    m = 2;
    theta -= 1.0;
  } else {
    m = m + 1;
  }
  const int n = (int)fx1.size() - 1;
  if (m + 1 > n) m = n - 1;  // beyond the last record: extrapolate from the end
  return fx1[m] + 0.5 * theta * (fx1[m + 1] - fx1[m - 1] + theta * (fx1[m + 1] + fx1[m - 1] - 2.0 * fx1[m]));
}

// Values per record of the surface boundary series: ifsurfbc 20: cdw uair
// vair; 1: eta Qsw Qn cdw uair vair; 2: eta Qsw Ta Pa Rh Cc cdw uair vair;
// 3: eta Qsw Ta Pa Rh Qlw cdw uair vair (si3d_boundaryconditions.f90:4160-4300,
// 4450-4484; nvSurfbcW/P/R, si3d_types.f90:173-175).
int surfbc_nvar(int ifsurfbc, int nmet = 0) {
  switch (ifsurfbc) {
    case 20: return 3;
    case 1: return 6;
    case 2: case 3: return 9;
    case 10: case 11: return 6 * nmet + 2;  // eta Pa, 6 per station (4301-4450)
    default: return 0;
  }
}

// Every variable of the series at time thrs (hours); false with fewer than
// three records (parab needs three).
static bool interp_series(const std::vector<double> &sb, int nvar, double dtsurfbc, double thrs, double *v) {
  const int np = nvar > 0 ? (int)(sb.size() / nvar) : 0;
  if (np < 3) return false;
  std::vector<double> f(np + 1);
  for (int q = 0; q < nvar; ++q) {
    for (int j = 0; j < np; ++j) f[j + 1] = sb[(size_t)nvar * j + q];
    v[q] = parab(0., thrs, f, dtsurfbc / 3600.);
  }
  return true;
}

// Saturated vapour pressure (Pa) at T (K), si3d_boundaryconditions.f90:5488-5496.
static double sat_vapor_pressure(double T) { return 2.1718e10 * std::exp(-4157. / (T - 33.91)); }

// Bulk transfer coefficients of heat (H) and vapour (E) for neutral
// conditions as a function of wind speed, Kondo (1975), the five wind ranges
// and constants of si3d_boundaryconditions.f90:5166-5178, 5235-5243.
static void kondo_coefficients(double ws, double &cE, double &cH) {
  static const double lim[6] = {0.0, 2.2, 5.0, 8.0, 25.0, 50.0};
  static const double ah[5] = {0.000, 0.9270, 1.1500, 1.1700, 1.652}, bh[5] = {1.185, 0.0521, 0.0100, 0.0075, -0.017},
                      ch[5] = {0.000, 0.0000, 0.0000, -4.5E-4, 0.000}, ph[5] = {-0.157, 1., 1., 1., 1.};
  static const double ae[5] = {0.000, 0.9690, 1.1800, 1.1960, 1.680}, be[5] = {1.230, 0.0521, 0.0100, 0.0080, -0.016},
                      ce[5] = {0.000, 0.0000, 0.0000, -4.0E-4, 0.000}, pe[5] = {-0.160, 1., 1., 1., 1.};
  int r = 4;
  for (int q = 0; q < 4; ++q)
    if (ws > lim[q] && ws <= lim[q + 1]) {
      r = q;
      break;
    }
  cH = (ah[r] + bh[r] * std::pow(ws, ph[r]) + ch[r] * (ws - 8.) * (ws - 8.)) * 1.e-3;
  cE = (ae[r] + be[r] * std::pow(ws, pe[r]) + ce[r] * (ws - 8.) * (ws - 8.)) * 1.e-3;
}

// The column-independent part of the heat budget (ifsurfbc 1-3,
// si3d_boundaryconditions.f90:5182-5320) from the interpolated series values.
static s3g::HeatForce heat_force_of(int mode, const double *v) {
  constexpr double kStefanBoltzmann = 5.6697e-8, kAlbedoLw = 0.03;
  s3g::HeatForce H{};
  H.mode = mode;
  H.eta = v[0];
  H.Qsw = v[1];
  const double uair = v[surfbc_nvar(mode) - 2], vair = v[surfbc_nvar(mode) - 1];
  if (mode == 1) {
    H.Qn = v[2];
    return H;
  }
  H.Ta = v[2];
  H.Pa = v[3];
  const double Rh = v[4];
  const double TaK = H.Ta + 273.;
  H.ea = sat_vapor_pressure(TaK) * Rh;
  if (mode == 2) {  // longwave from air temperature, humidity and cloud cover
    const double Cc = v[5];
    const double em_air = 0.642 * std::pow(H.ea / TaK, 0.14285714) * (1. + 0.17 * Cc * Cc);
    H.Qbra = em_air * kStefanBoltzmann * TaK * TaK * TaK * TaK * (1. - kAlbedoLw);
  } else {  // measured incoming longwave
    H.Qbra = v[5] * (1. - kAlbedoLw);
  }
  H.ws = std::sqrt(uair * uair + vair * vair);
  kondo_coefficients(H.ws, H.coeffE, H.coeffH);
  return H;
}

// Station weights of the met-station modes (InitializeInterpolationMethods,
// si3d_boundaryconditions.f90:5510-5646) at every wet column (i, j), in
// grid units: 1 Barnes, exp(-ln(100) d^2/da^2) normalised with da the mean
// distance to the stations; 2 two-pass Barnes (Koch et al. 1983),
// exp(-d^2/kappa) unnormalised, kappa from the mean nearest-station distance,
// gammaB and delNfactor, and the same weights between stations (at the cell
// nearest each station); 3 inverse squared distance (1 at the station's own
// cell), normalised.  Device layout [m][LP], zero for the sentinel.
static int met_setup(s3g_ctx *c) {
  const int nm = (int)c->metxy.size() / 2, LP = c->LP;
  auto d2 = [&](int m, double x, double y) {
    const double dx = c->metxy[2 * m] - x, dy = c->metxy[2 * m + 1] - y;
    return dx * dx + dy * dy;
  };
  double kappa = 0.0;
  if (c->iinterp == 2) {
    double da = 0.0;
    for (int m = 0; m < nm; ++m) {
      double nearest = HUGE_VAL;  // no other station: the largest value
      for (int q = 0; q < nm; ++q)
        if (q != m) nearest = std::min(nearest, std::sqrt(d2(q, c->metxy[2 * m], c->metxy[2 * m + 1])));
      da += nearest;
    }
    da /= nm;
    const double g = c->gammaB;
    const double f = g <= 0.2 ? 5.0515 : g <= 0.3 ? 3.5132 : g <= 0.5 ? 2.3838 : g <= 0.6 ? 2.1153 : 1.7826;
    const double pi = 3.14159265358979;
    kappa = f * 4.0 / (pi * pi) * (c->delNfactor * da) * (c->delNfactor * da);
    c->met_wn.assign((size_t)nm * nm, 0.0);
    for (int m = 0; m < nm; ++m) {
      const double x = std::nearbyint(c->metxy[2 * m]), y = std::nearbyint(c->metxy[2 * m + 1]);
      for (int q = 0; q < nm; ++q) c->met_wn[(size_t)m * nm + q] = std::exp(-d2(q, x, y) / kappa);
    }
  }
  std::vector<double> w((size_t)nm * LP, 0.0);
  for (int col = 0; col < c->lm; ++col) {
    const double x = c->l2i[col], y = c->l2j[col];
    double sum = 0.0, da = 0.0;
    if (c->iinterp == 1) {
      for (int m = 0; m < nm; ++m) da += std::sqrt(d2(m, x, y));
      da /= nm;
    }
    for (int m = 0; m < nm; ++m) {
      const double dd = d2(m, x, y);
      double v;
      if (c->iinterp == 1) v = std::exp(-4.60517018598809 * dd / (da * da));
      else if (c->iinterp == 2) v = std::exp(-dd / kappa);
      else v = dd > 0.0 ? 1.0 / dd : 1.0;
      w[(size_t)m * LP + col] = v;
      sum += v;
    }
    if (c->iinterp != 2)
      for (int m = 0; m < nm; ++m) w[(size_t)m * LP + col] /= sum;
  }
  if (hipMalloc(&c->d_wmet, w.size() * sizeof(double)) != hipSuccess) return hip_check(hipErrorOutOfMemory, "met");
  c->allocs.push_back(c->d_wmet);
  return hip_check(hipMemcpy(c->d_wmet, w.data(), w.size() * sizeof(double), hipMemcpyHostToDevice), "met weights");
}

// surfbc cases 10 and 11 (si3d_boundaryconditions.f90:4596-4650, on leapfrog
// steps and step 1): every series value at thrs, then the first pass's
// residuals at the stations for the two-pass scheme (CalculateMetResiduals,
// 5650-5700); the columns are done on the device.  Mode 10 has no longwave
// series to interpolate (its fourth station value is cloud cover, which the
// column budget does not read).
static void met_surfbc(s3g_ctx *c, int n, int istep) {
  if (n > 1 && istep == 2) return;
  const int nm = (int)c->metxy.size() / 2, nv = 6 * nm + 2;
  std::vector<double> v(nv);
  if (!interp_series(c->surfbc, nv, c->P.dtsurfbc, c->thrs, v.data())) return;
  s3g::MetArgs &M = c->met;
  M.mode = c->P.ifsurfbc;
  M.nmet = nm;
  M.iinterp = c->iinterp;
  M.icw = c->icw;
  M.eta = v[0];
  M.Pa = v[1];
  M.gamma = c->gammaB;
  M.cw = c->P.cw;
  for (int i = 0; i < 6 * nm; ++i) M.st[i] = v[2 + i], M.res[i] = 0.0;
  if (c->iinterp == 2)
    for (int k = 0; k < nm; ++k) {
      double sumw = 0.0;
      for (int m = 0; m < nm; ++m) sumw += c->met_wn[(size_t)k * nm + m];
      for (int q = 0; q < 6; ++q) {
        double first = 0.0;
        for (int m = 0; m < nm; ++m) first += c->met_wn[(size_t)k * nm + m] / sumw * M.st[6 * m + q];
        M.res[6 * k + q] = M.st[6 * k + q] - first;
      }
    }
  c->met_due = true;
}

// surfbc cases 0, 1-3 and 20 (si3d_boundaryconditions.f90:4490-4669): the
// forcing is uniform in space, so wind is kept as three scalars; the heat
// budget cases also flag HeatSource for recomputation on the device.
static void forcing_of(const s3g_ctx::LakeForcing &F, double thrs, int n, int istep, double *out) {
  if (F.ifsurfbc == 0) {
    out[0] = -F.wa * std::sin(3.1415926535897932 * F.phi / 180.);
    out[1] = -F.wa * std::cos(3.1415926535897932 * F.phi / 180.);
    out[2] = F.cw;
  } else if (F.ifsurfbc == 20) {
    if (n > 1 && istep == 2) return;
    double v[3];
    if (!interp_series(F.surfbc, 3, F.dtsurfbc, thrs, v)) return;
    out[2] = v[0];
    out[0] = v[1];
    out[1] = v[2];
  }
}

void surfbc(s3g_ctx *c, int n, int istep) {
  if (c->nlakes > 1) {
    for (int i = 0; i < c->nlakes; ++i) forcing_of(c->lake_forcing[i], c->thrs, n, istep, &c->lake_force[3 * i]);
    set_forcing(c);
    return;
  }
  const int mode = c->P.ifsurfbc;
  if (mode == 10 || mode == 11) {
    met_surfbc(c, n, istep);
    return;
  }
  if (mode == 0) {
    c->uair = -c->P.wa * std::sin(3.1415926535897932 * c->P.phi / 180.);
    c->vair = -c->P.wa * std::cos(3.1415926535897932 * c->P.phi / 180.);
    c->cdw = c->P.cw;
    return;
  }
  if (n > 1 && istep == 2) return;  // only on leapfrog steps (and step 1)
  const int nv = surfbc_nvar(mode);
  double v[9];
  if (!interp_series(c->surfbc, nv, c->P.dtsurfbc, c->thrs, v)) return;
  c->cdw = v[nv - 3];
  c->uair = v[nv - 2];
  c->vair = v[nv - 1];
  if (mode >= 1 && mode <= 3) {
    c->heat_force = heat_force_of(mode, v);
    c->heat_due = true;
  }
}

// ---------------------------------------------------------------------------
// Field table: name -> device storage.
enum Kind { K3R, K3T, K2D, KI };  // K3T: turbulence state (storage type, held only with iturb > 0)
struct Field {
  const char *name;
  Kind kind;
  void *ptr;
};

template <typename R>
static std::vector<Field> field_table(s3g_ctx *c, Dev<R> &D) {
  std::vector<Field> t = {
      {"h", K3R, D.h},         {"hp", K3R, D.hp},       {"hpp", K3R, D.hpp},     {"hu", K3R, D.hu},
      {"hup", K3R, D.hup},     {"hupp", K3R, D.hupp},   {"hv", K3R, D.hv},       {"hvp", K3R, D.hvp},
      {"hvpp", K3R, D.hvpp},   {"uh", K3R, D.uh},       {"uhp", K3R, D.uhp},     {"uhpp", K3R, D.uhpp},
      {"u", K3R, D.u},         {"up", K3R, D.up},       {"upp", K3R, D.upp},     {"vh", K3R, D.vh},
      {"vhp", K3R, D.vhp},     {"vhpp", K3R, D.vhpp},   {"v", K3R, D.v},         {"vp", K3R, D.vp},
      {"vpp", K3R, D.vpp},     {"wp", K3R, D.wp},       {"sal", K3R, D.sal},     {"salp", K3R, D.salp},
      {"salpp", K3R, D.salpp}, {"rhop", K3R, D.rhop},   {"Av", K3R, D.Av},       {"Dv", K3R, D.Dv},
      {"kh", K3R, D.kh},       {"ex", K3R, D.ex},       {"ey", K3R, D.ey},       {"agx", K3R, D.agx},     {"arx", K3R, D.arx},
      {"agy", K3R, D.agy},     {"ary", K3R, D.ary},     {"bex", K3R, D.bex},     {"q2", K3T, D.q2},
      {"q2p", K3T, D.q2p},     {"q2pp", K3T, D.q2pp},   {"q2l", K3T, D.q2l},     {"q2lp", K3T, D.q2lp},
      {"q2lpp", K3T, D.q2lpp}, {"s", K2D, D.s},         {"sp", K2D, D.sp},       {"spp", K2D, D.spp},
      {"sx", K2D, D.sx},       {"sy", K2D, D.sy},       {"dd", K2D, D.dd},       {"qq", K2D, D.qq},
      {"rr", K2D, D.rr},       {"eagx", K2D, D.eagx},   {"earx", K2D, D.earx},   {"eagy", K2D, D.eagy},
      {"eary", K2D, D.eary},   {"hhs", K2D, (void *)D.hhs}, {"hhu", K2D, (void *)D.hhu},
      {"hhv", K2D, (void *)D.hhv}, {"kmz", KI, (void *)D.kmz}, {"k1z", KI, D.k1z}, {"k1u", KI, D.k1u},
      {"k1v", KI, D.k1v},      {"HeatSource", K3R, D.heat},
  };
  const size_t n3 = (size_t)c->KP * c->LP;
  for (int nt = 0; nt < c->ntr; ++nt) {
    t.push_back({c->tracer_names[2 * nt].c_str(), K3R, D.tr + nt * n3});
    t.push_back({c->tracer_names[2 * nt + 1].c_str(), K3R, D.trpp + nt * n3});
  }
  return t;
}

static bool find_field(s3g_ctx *c, const char *name, Field &out) {
  auto tab = (c->precision == S3G_FP64) ? field_table(c, c->d64) : field_table(c, c->d32);
  for (auto &f : tab)
    if (std::strcmp(f.name, name) == 0) {
      out = f;
      return true;
    }
  return false;
}

// Reference layout (km1, lm1) k-fastest  <->  device planes [k][c].
template <typename HT, typename DT>
static int upload3(s3g_ctx *c, DT *dst, const HT *src) {
  const int km1 = c->km1, lm1 = c->lm + 1;
  std::vector<DT> buf((size_t)c->KP * c->LP, DT(0));
  for (int l = 0; l < lm1; ++l)
    for (int k = 1; k <= km1; ++k) buf[(size_t)k * c->LP + l] = (DT)src[(size_t)l * km1 + (k - 1)];
  return hip_check(hipMemcpy(dst, buf.data(), buf.size() * sizeof(DT), hipMemcpyHostToDevice), "upload3");
}

template <typename HT, typename DT>
static int download3(s3g_ctx *c, HT *dst, const DT *src) {
  const int km1 = c->km1, lm1 = c->lm + 1;
  std::vector<DT> buf((size_t)c->KP * c->LP);
  if (int e = hip_check(hipMemcpy(buf.data(), src, buf.size() * sizeof(DT), hipMemcpyDeviceToHost), "download3")) return e;
  for (int l = 0; l < lm1; ++l)
    for (int k = 1; k <= km1; ++k) dst[(size_t)l * km1 + (k - 1)] = (HT)buf[(size_t)k * c->LP + l];
  return 0;
}

// Fields that the configuration does not allocate (turbulence state with
// iturb = 0, Smagorinsky coefficients with ihd <= 1) cannot be accessed.
static bool allocated(const s3g_ctx *c, const Field &f) {
  if (!f.ptr) return false;
  if (f.kind == K3T) return c->P.iturb > 0;
  return true;
}

template <typename HT>
static int set_field(s3g_ctx *c, const char *name, const HT *src) {
  Field f;
  if (!find_field(c, name, f)) {
    set_error(std::string("unknown field ") + name);
    return S3G_ERR_ARG;
  }
  if (!allocated(c, f)) {
    set_error(std::string(name) + " is not held in this configuration");
    return S3G_ERR_ARG;
  }
  if (!std::strcmp(name, "agx") || !std::strcmp(name, "arx") || !std::strcmp(name, "agy") ||
      !std::strcmp(name, "ary")) {
    set_error(std::string(name) + " is formed by matmom on every fd call and cannot be set");
    return S3G_ERR_ARG;
  }
  const int lm1 = c->lm + 1;
  switch (f.kind) {
    case K3R:
    case K3T:
      if (c->precision == S3G_FP64) return upload3(c, (double *)f.ptr, src);
      return upload3(c, (float *)f.ptr, src);
    case K2D: {
      std::vector<double> b(c->LP, 0.0);
      for (int l = 0; l < lm1; ++l) b[l] = (double)src[l];
      return hip_check(hipMemcpy(f.ptr, b.data(), c->LP * sizeof(double), hipMemcpyHostToDevice), "set2");
    }
    case KI: {
      std::vector<int> b(c->LP, c->km1);
      for (int l = 0; l < lm1; ++l) b[l] = (int)src[l];
      return hip_check(hipMemcpy(f.ptr, b.data(), c->LP * sizeof(int), hipMemcpyHostToDevice), "seti");
    }
  }
  return S3G_ERR_ARG;
}

template <typename HT>
static int get_field(s3g_ctx *c, const char *name, HT *dst) {
  Field f;
  if (!find_field(c, name, f)) {
    set_error(std::string("unknown field ") + name);
    return S3G_ERR_ARG;
  }
  if (!allocated(c, f)) {
    set_error(std::string(name) + " is not held in this configuration");
    return S3G_ERR_ARG;
  }
  if (int e = hip_check(hipStreamSynchronize(c->stream), "sync")) return e;
  const int lm1 = c->lm + 1;
  // agx/arx/agy/ary hold matmom's downward-pass factors (vel back-substitutes
  // them); the solution the reference holds there is formed on request
  const bool mx = !std::strcmp(name, "agx") || !std::strcmp(name, "arx");
  if (mx || !std::strcmp(name, "agy") || !std::strcmp(name, "ary")) {
    const size_t bytes = (size_t)c->KP * c->LP * (c->precision == S3G_FP64 ? sizeof(double) : sizeof(float));
    void *ag = nullptr, *ar = nullptr;
    int rc = hip_check(hipMalloc(&ag, bytes), "momentum solution");
    if (!rc) rc = hip_check(hipMalloc(&ar, bytes), "momentum solution");
    if (!rc) rc = hip_check(hipMemsetAsync(ag, 0, bytes, c->stream), "momentum solution");
    if (!rc) rc = hip_check(hipMemsetAsync(ar, 0, bytes, c->stream), "momentum solution");
    if (!rc)
      rc = (c->precision == S3G_FP64) ? momentum_solution<double>(c, mx, ag, ar) : momentum_solution<float>(c, mx, ag, ar);
    if (!rc) rc = hip_check(hipStreamSynchronize(c->stream), "sync");
    const void *src = (name[1] == 'g') ? ag : ar;
    if (!rc)
      rc = (c->precision == S3G_FP64) ? download3(c, dst, (const double *)src) : download3(c, dst, (const float *)src);
    if (ag) hipFree(ag);
    if (ar) hipFree(ar);
    return rc;
  }
  switch (f.kind) {
    case K3R:
    case K3T:
      if (c->precision == S3G_FP64) return download3(c, dst, (const double *)f.ptr);
      return download3(c, dst, (const float *)f.ptr);
    case K2D: {
      std::vector<double> b(c->LP);
      if (int e = hip_check(hipMemcpy(b.data(), f.ptr, c->LP * sizeof(double), hipMemcpyDeviceToHost), "get2")) return e;
      for (int l = 0; l < lm1; ++l) dst[l] = (HT)b[l];
      return 0;
    }
    case KI: {
      std::vector<int> b(c->LP);
      if (int e = hip_check(hipMemcpy(b.data(), f.ptr, c->LP * sizeof(int), hipMemcpyDeviceToHost), "geti")) return e;
      for (int l = 0; l < lm1; ++l) dst[l] = (HT)b[l];
      return 0;
    }
  }
  return S3G_ERR_ARG;
}

// Allocate the device state of one lake in storage type R.
template <typename R>
static int alloc_state(s3g_ctx *c, Dev<R> &D, const int *lE, const int *lW, const int *lN, const int *lS,
                       const int *kmz, const int *k1z, const int *k1u, const int *k1v, const double *zl,
                       const double *hhs, const double *hhu, const double *hhv, const int *lake_of,
                       const int *lake_km1) {
  const size_t n3 = (size_t)c->KP * c->LP, n2 = c->LP;
  D.lm = c->lm;
  D.LP = c->LP;
  D.KP = c->KP;
  D.km1 = c->km1;
  D.km = c->km;
  D.k1 = c->k1;
  D.P = c->P;
  auto up_i = [&](const int *src) -> int * {
    int *d = dalloc<int>(c, n2);
    if (d) hipMemcpy(d, src, n2 * sizeof(int), hipMemcpyHostToDevice);
    return d;
  };
  auto up_d = [&](const double *src, size_t n) -> double * {
    double *d = dalloc<double>(c, n);
    if (d) hipMemcpy(d, src, n * sizeof(double), hipMemcpyHostToDevice);
    return d;
  };
  D.lE = up_i(lE);
  D.lW = up_i(lW);
  D.lN = up_i(lN);
  D.lS = up_i(lS);
  D.kmz = up_i(kmz);
  {
    const int S = c->lm;
    std::vector<int> kmu(n2, -1), kmv(n2, -1);
    for (int l = 0; l < S; ++l) {
      if (lE[l] != S) kmu[l] = std::min(kmz[lE[l]], kmz[l]);
      if (lN[l] != S) kmv[l] = std::min(kmz[lN[l]], kmz[l]);
    }
    D.kmu = up_i(kmu.data());
    D.kmv = up_i(kmv.data());
  }
  D.k1z = up_i(k1z);
  D.k1u = up_i(k1u);
  D.k1v = up_i(k1v);
  D.zlevel = up_d(zl, (size_t)c->KP * c->nlakes);
  D.KPz = c->KP;
  D.lake = nullptr;
  D.lake_km1 = nullptr;
  // the uniform wind of one lake lives in device memory too (three doubles
  // written before every fd call), so the kernel sequences of a step take the
  // same arguments on every call and can be replayed from a captured graph
  if (!c->d_force) c->d_force = dalloc<double>(c, 3 * (size_t)c->nlakes);
  D.force = c->d_force;
  if (!D.force) return S3G_ERR_HIP;
  if (c->nlakes > 1) {
    D.lake = up_i(lake_of);
    int *k = dalloc<int>(c, c->nlakes);
    if (k) hipMemcpy(k, lake_km1, c->nlakes * sizeof(int), hipMemcpyHostToDevice);
    D.lake_km1 = k;
    if (!D.lake || !D.lake_km1) return S3G_ERR_HIP;
  }
  D.hhs = up_d(hhs, n2);
  D.hhu = up_d(hhu, n2);
  D.hhv = up_d(hhv, n2);
  double **d2[] = {&D.s, &D.sp, &D.spp, &D.sx, &D.sy, &D.dd, &D.qq, &D.rr, &D.eagx, &D.earx, &D.eagy, &D.eary};
  for (double **p : d2)
    if (!(*p = dalloc<double>(c, n2))) return S3G_ERR_HIP;
  R **d3[] = {&D.h,    &D.hp,   &D.hpp,  &D.hu,  &D.hup,  &D.hupp, &D.hv,  &D.hvp,  &D.hvpp, &D.uh,
              &D.uhp,  &D.uhpp, &D.u,    &D.up,  &D.upp,  &D.vh,   &D.vhp, &D.vhpp, &D.v,    &D.vp,
              &D.vpp,  &D.wp,   &D.sal,  &D.salp, &D.salpp, &D.rhop, &D.Av, &D.Dv,   &D.ex, &D.ey,
              &D.agx,  &D.arx,  &D.agy,  &D.ary, &D.fx_ew, &D.fx_ns, &D.bex, &D.t0, &D.cfx, &D.cfy};
  for (R **p : d3)
    if (!(*p = dalloc<R>(c, n3))) return S3G_ERR_HIP;
  // Smagorinsky coefficients only for ihd > 1, turbulence state only for iturb > 0
  D.kh = nullptr;
  D.t1 = nullptr;
  if (c->P.ihd > 1 && !(D.kh = dalloc<R>(c, n3))) return S3G_ERR_HIP;
  D.heat = nullptr;
  const bool budget = (c->P.ifsurfbc >= 1 && c->P.ifsurfbc <= 3) || c->P.ifsurfbc == 10 || c->P.ifsurfbc == 11;
  if (budget && !(D.heat = dalloc<R>(c, n3))) return S3G_ERR_HIP;
  D.wind = nullptr;
  if (c->P.ifsurfbc == 10 || c->P.ifsurfbc == 11) {  // per-column wind of the met-station modes
    if (!c->d_wind) c->d_wind = dalloc<double>(c, 3 * (size_t)c->LP);
    D.wind = c->d_wind;
    if (!D.wind) return S3G_ERR_HIP;
  }
  D.ntr = c->ntr;
  D.tr = D.trpp = nullptr;
  if (c->ntr > 0) {
    if (!(D.tr = dalloc<R>(c, (size_t)c->ntr * n3)) || !(D.trpp = dalloc<R>(c, (size_t)c->ntr * n3)))
      return S3G_ERR_HIP;
  }
  R **q3[] = {&D.q2, &D.q2p, &D.q2pp, &D.q2l, &D.q2lp, &D.q2lpp, &D.d0, &D.d1, &D.d2};
  for (R **p : q3) {  // d0, d1 also hold the fused turbulence sweep's factor pairs: 2 n3
    const size_t n = (c->P.iturb > 0 ? n3 : 1) * ((p == &D.d0 || p == &D.d1) ? 2 : 1);
    if (!(*p = dalloc<R>(c, n))) return S3G_ERR_HIP;
  }
  D.tkb = nullptr;
  D.tke_part = D.tke = nullptr;
  if (c->P.iturb > 0) {
    if (!(D.tkb = dalloc<float2>(c, n3)) || !(D.tke_part = dalloc<double>(c, 4 * ((size_t)c->lm / kBlock + 1))) ||
        !(D.tke = dalloc<double>(c, 8)))
      return S3G_ERR_HIP;
  }
  std::vector<double> one(n2, 1.0);
  hipMemcpy(D.rr, one.data(), n2 * sizeof(double), hipMemcpyHostToDevice);
  return hip_check(hipGetLastError(), "alloc");
}

}  // namespace s3g

using namespace s3g;

extern "C" {

const char *s3g_last_error(void) { return g_err.c_str(); }

int s3g_init(const s3g_grid *g, const s3g_params *p, s3g_ctx **out) {
  return s3g::init_ctx(g, p, 1, nullptr, nullptr, nullptr, out);
}

}  // extern "C"

int s3g::init_ctx(const s3g_grid *g, const s3g_params *p, int nlakes, const double *zl_lakes, const int *lake_of,
                  const int *lake_km1, s3g_ctx **out) {
  if (!g || !p || !out || g->lm <= 0 || g->km1 < 3) {
    set_error("s3g_init: invalid grid");
    return S3G_ERR_ARG;
  }
  if (p->precision != S3G_FP64 && p->precision != S3G_MIXED) {
    set_error("s3g_init: unknown precision");
    return S3G_ERR_ARG;
  }
  if (p->itrmom < 1 || p->itrmom > 4) {
    set_error("s3g_init: itrmom must be 1..4");
    return S3G_ERR_UNSUPPORTED;
  }
  if (p->iturb < 0) {
    set_error("s3g_init: iturb < 0 (GOTM) is not supported");
    return S3G_ERR_UNSUPPORTED;
  }
  if (p->ifsurfbc != 0 && p->ifsurfbc != 20 && (p->ifsurfbc < 1 || p->ifsurfbc > 3) && p->ifsurfbc != 10 &&
      p->ifsurfbc != 11) {
    set_error("s3g_init: ifsurfbc must be 0, 1, 2, 3, 10, 11 or 20");
    return S3G_ERR_UNSUPPORTED;
  }
  if ((p->ifsurfbc == 10 || p->ifsurfbc == 11) &&
      (p->nmetstat < 1 || p->nmetstat > s3g::kMaxMet || !p->metxy || p->iinterp < 1 || p->iinterp > 3 ||
       (p->iinterp == 2 && !(p->gammaB > 0.0)) || p->icw < 0 || p->icw > 2)) {
    set_error("s3g_init: ifsurfbc 10/11 need 1-16 met stations with their positions, iinterp 1-3 (gammaB > 0 "
              "for 2) and icw 0-2");
    return S3G_ERR_ARG;
  }
  if (nlakes > 1 && p->ifsurfbc != 0 && p->ifsurfbc != 20) {
    set_error("s3g_create_ensemble: the surface heat budget (ifsurfbc 1-3, 10, 11) is not supported in an ensemble");
    return S3G_ERR_UNSUPPORTED;
  }
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) {
    set_error("s3g_init: no HIP device is visible");
    return S3G_ERR_HIP;
  }
  if (p->ntr < 0 || p->ntr > 64) {
    set_error("s3g_init: ntr must be 0..64");
    return S3G_ERR_ARG;
  }
  s3g_ctx *c = new s3g_ctx();
  c->nlakes = nlakes;
  c->ntr = p->ntr;
  for (int nt = 1; nt <= c->ntr; ++nt) {
    c->tracer_names.push_back("tracer" + std::to_string(nt));
    c->tracer_names.push_back("tracerpp" + std::to_string(nt));
  }
  c->precision = p->precision;
  c->debug = std::getenv("S3G_DEBUG") != nullptr;
  {
    const char *t = std::getenv("S3G_TUNE");
    c->tune = !(t && t[0] == '0');
    if (const char *m = std::getenv("S3G_MG")) {
      int a = 0, b = 0;
      double cs = 1.0;
      const int got = std::sscanf(m, "%d,%d,%lf", &a, &b, &cs);
      if (got >= 2 && a >= 0 && b >= 1) c->mg_small = a, c->mg_sweeps = b;
      if (got == 3 && cs > 0.0 && cs <= 1.0) c->mg_cs = cs;
    }
    if (const char *gs = std::getenv("S3G_GUESS")) c->pcg_guess = (gs[0] == '1') ? 1 : 0;
    if (const char *gr = std::getenv("S3G_GRAPH")) c->cg_graph = (gr[0] != '0');
    if (const char *gs = std::getenv("S3G_STEP_GRAPH")) c->step_graph = (gs[0] != '0');
    if (const char *x = std::getenv("S3G_EXSAL")) c->exsal_split = std::strcmp(x, "fused") != 0;
    if (const char *x = std::getenv("S3G_TURB")) c->turb_split = std::strcmp(x, "split") == 0;
    if (const char *x = std::getenv("S3G_BUDGET")) c->tke_budget = x[0] != '0';
    if (const char *x = std::getenv("S3G_WINDOW")) c->window = x[0] != '0';
    // the windowed settrap / save kernels (not tuned: they rewrite their
    // inputs) take 8 layers per lane
    if (c->window) c->kc[s3g::TILE_SETTRAP] = c->kc[s3g::TILE_SAVE] = 8;
    // S3G_TILES="w0,w1,...": fixed tile words (chunk | order << 8), no tuning
    if (const char *w = std::getenv("S3G_TILES")) {
      c->tune = false;
      for (int i = 0; i < s3g::TILE_NUM && *w; ++i) {
        char *end = nullptr;
        const long v = std::strtol(w, &end, 10);
        if (end == w) break;
        if ((v & 255) >= 1) c->kc[i] = (int)v;
        w = (*end == ',') ? end + 1 : end;
      }
    }
  }
  c->prec_kind = p->preconditioner;
  c->im1 = g->im1;
  c->jm1 = g->jm1;
  c->km1 = g->km1;
  c->km = g->km1 - 1;
  c->k1 = 2;
  c->lm = g->lm;
  c->LP = ((g->lm + 1 + 255) / 256) * 256;
  c->KP = g->km1 + 3;
  c->pcg_tol = p->pcg_tol > 0 ? p->pcg_tol : 1e-10;
  c->pcg_maxit = p->pcg_maxit > 0 ? p->pcg_maxit : 4000;
  Params &P = c->P;
  P.dt = p->dt;
  P.dx = p->dx;
  P.dy = p->dy;
  P.twodt = 2.0 * P.dt;
  P.dtdx = P.dt / P.dx;
  P.dtdy = P.dt / P.dy;
  P.gdtdx = 9.806 * P.dtdx;
  P.gdtdy = 9.806 * P.dtdy;
  P.gdt2dx2 = P.gdtdx * P.dtdx;
  P.gdt2dy2 = P.gdtdy * P.dtdy;
  P.twodx = 2.0 * P.dx;
  P.twody = 2.0 * P.dy;
  P.fourdx = 2.0 * P.twodx;
  P.fourdy = 2.0 * P.twody;
  P.dxdx = P.dx * P.dx;
  P.dydy = P.dy * P.dy;
  P.dxdy = P.dx * P.dy;
  P.beta = p->beta;
  P.beta2 = p->beta / 2.0;
  P.itrap = p->itrap;
  P.niter = p->niter;
  P.ismooth = p->ismooth;
  P.iturb = p->iturb;
  P.iadv = p->iadv;
  P.itrmom = p->itrmom;
  P.ihd = p->ihd;
  P.ibc = p->ibc;
  P.isal = p->isal;
  P.ifsurfbc = p->ifsurfbc;
  P.Av0 = p->Av0;
  P.Dv0 = p->Dv0;
  P.Ax0 = p->Ax0;
  P.Ay0 = p->Ay0;
  P.f = p->f;
  P.theta = p->theta;
  P.cd = p->cd;
  P.cw = p->cw;
  P.wa = p->wa;
  P.phi = p->phi;
  P.dtsurfbc = p->dtsurfbc;
  // closure constants, si3d_mixing.f90:641-645
  const double A1 = 0.92, A2 = 0.74, B1 = 16.6, B2 = 10.1, C2 = 0.70, C3 = 0.20;
  P.t_1 = A2 * (1. - 6. * A1 / B1);
  P.t_2 = 3. * A2 * (6. * A1 + B2 * (1. - C3));
  P.t_3 = std::pow(B1, -1. / 3.);
  P.t_4 = 9. * A1 * (2. * A1 + A2 * (1. - C2));
  P.t_5 = 9. * A1 * A2;
  const int nmet = (p->ifsurfbc == 10 || p->ifsurfbc == 11) ? p->nmetstat : 0;
  const int nv = surfbc_nvar(p->ifsurfbc, nmet);
  if (p->n_surfbc > 0 && p->surfbc && nv > 0) c->surfbc.assign(p->surfbc, p->surfbc + (size_t)nv * p->n_surfbc);
  if (nmet > 0) {
    c->metxy.assign(p->metxy, p->metxy + 2 * (size_t)nmet);
    c->iinterp = p->iinterp;
    c->icw = p->icw;
    c->gammaB = p->gammaB;
    c->delNfactor = p->delNfactor;
  } else if (nv > 0 && (int)c->surfbc.size() >= nv) {  // momentum flux terms at t = 0 (surfbc0)
    c->cdw = c->surfbc[nv - 3];
    c->uair = c->surfbc[nv - 2];
    c->vair = c->surfbc[nv - 1];
  }
  if (hip_check(hipStreamCreate(&c->stream), "stream") < 0) {
    delete c;
    return S3G_ERR_HIP;
  }
  hipHostMalloc((void **)&c->h_iters, sizeof(int) * 4, hipHostMallocDefault);
  hipHostMalloc((void **)&c->h_scal, sizeof(double) * 4096, hipHostMallocDefault);
  // geometry in device numbering (0-based columns, sentinel = lm)
  const int lm = c->lm, LP = c->LP, S = lm;
  std::vector<int> lE(LP, S), lW(LP, S), lN(LP, S), lS(LP, S), kmz(LP, c->km1), k1z(LP, c->km1),
      k1u(LP, c->km1), k1v(LP, c->km1);
  std::vector<double> hhs(LP, 0.0), hhu(LP, 0.0), hhv(LP, 0.0), zl(c->KP, 1e30);
  std::vector<int> lake(LP, 0);
  long nwet = 0;
  for (int l = 0; l < lm; ++l) {
    auto nb = [&](const int *a) { int v = a[l] - 1; return (v < 0 || v >= lm) ? S : v; };
    lE[l] = nb(g->lEC);
    lW[l] = nb(g->lWC);
    lN[l] = nb(g->lNC);
    lS[l] = nb(g->lSC);
    kmz[l] = g->kmz[l];
    k1z[l] = g->k1z[l];
    k1u[l] = g->k1u[l];
    k1v[l] = g->k1v[l];
    hhs[l] = g->hhs[l];
    hhu[l] = g->hhu[l];
    hhv[l] = g->hhv[l];
    nwet += std::max(0, g->kmz[l] - g->k1z[l] + 1);
  }
  if (zl_lakes) {
    zl.assign(zl_lakes, zl_lakes + (size_t)c->KP * nlakes);
    for (int l = 0; l <= lm; ++l) lake[l] = lake_of[l];
  } else {
    zl[0] = -100.0;
    for (int k = 1; k <= c->km1; ++k) zl[k] = g->zlevel[k - 1];
  }
  c->n_wet = nwet;
  c->nb_symmetric = true;
  for (int l = 0; l < lm && c->nb_symmetric; ++l)
    c->nb_symmetric = (lE[l] == S || lW[lE[l]] == l) && (lW[l] == S || lE[lW[l]] == l) &&
                      (lN[l] == S || lS[lN[l]] == l) && (lS[l] == S || lN[lS[l]] == l);
  int rc = (c->precision == S3G_FP64)
               ? alloc_state(c, c->d64, lE.data(), lW.data(), lN.data(), lS.data(), kmz.data(), k1z.data(),
                             k1u.data(), k1v.data(), zl.data(), hhs.data(), hhu.data(), hhv.data(), lake.data(),
                             lake_km1)
               : alloc_state(c, c->d32, lE.data(), lW.data(), lN.data(), lS.data(), kmz.data(), k1z.data(),
                             k1u.data(), k1v.data(), zl.data(), hhs.data(), hhu.data(), hhv.data(), lake.data(),
                             lake_km1);
  Pcg &G = c->pcg;
  const size_t n2 = LP;
  double **pv[] = {&G.r, &G.r1, &G.z, &G.p0, &G.p1, &G.q, &G.x, &G.lc, &G.ld};
  for (double **q : pv)
    if (rc == 0 && !(*q = dalloc<double>(c, n2))) rc = S3G_ERR_HIP;
  float **pf[] = {&G.rf, &G.rf1, &G.zf, &G.pf0, &G.pf1, &G.qf, &G.sxf, &G.syf, &G.rrf, &G.bf, &G.df, &G.lcf, &G.ldf};
  for (float **q : pf)
    if (rc == 0 && !(*q = dalloc<float>(c, c->precision == S3G_MIXED ? n2 : 1))) rc = S3G_ERR_HIP;
  c->l2i.assign(g->l2i, g->l2i + c->lm);
  c->l2j.assign(g->l2j, g->l2j + c->lm);
  if (rc == 0 && !c->metxy.empty()) rc = met_setup(c);
  if (rc == 0) rc = pcg_setup(c, lN.data(), g->l2i, g->l2j);
  if (rc == 0 && !(G.part = dalloc<double>(c, 10 * (size_t)G.nblk + 64))) rc = S3G_ERR_HIP;  // 10 partial slots
  if (rc == 0 && !(G.iters = dalloc<int>(c, 4))) rc = S3G_ERR_HIP;
  if (rc == 0 && !(G.tol2 = dalloc<double>(c, 1))) rc = S3G_ERR_HIP;
  if (rc != 0) {
    if (g_err.empty()) set_error("s3g_init: device allocation failed");
    s3g_finalize(c);
    return rc;
  }
  c->stats.lm = lm;
  c->stats.km1 = c->km1;
  c->stats.n_wet = nwet;
  *out = c;
  return hip_check(hipDeviceSynchronize(), "init");
}

namespace s3g {
// Streaming copy for s3g_probe_stream: 16-byte loads and stores, grid-stride.
__global__ void __launch_bounds__(256) k_stream_copy(const float4 *__restrict__ a, float4 *__restrict__ b, size_t n) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    b[i] = a[i];
}
}  // namespace s3g

extern "C" {

int s3g_set_field_f32(s3g_ctx *c, const char *n, const float *h) { return set_field(c, n, h); }
int s3g_set_field_f64(s3g_ctx *c, const char *n, const double *h) { return set_field(c, n, h); }
int s3g_set_field_i32(s3g_ctx *c, const char *n, const int *h) { return set_field(c, n, h); }
int s3g_get_field_f32(s3g_ctx *c, const char *n, float *h) { return get_field(c, n, h); }
int s3g_get_field_f64(s3g_ctx *c, const char *n, double *h) { return get_field(c, n, h); }
int s3g_get_field_i32(s3g_ctx *c, const char *n, int *h) { return get_field(c, n, h); }

int s3g_fd(s3g_ctx *c, int n, int istep, int iter, int lastiter, double thrs) {
  if (!c || (istep != 1 && istep != 2)) return S3G_ERR_ARG;
  c->thrs = thrs;
  c->solve_slot = (istep == 1) ? 0 : std::min(std::max(iter, 1), 3);
  c->stats.fd_calls++;
  return c->precision == S3G_FP64 ? s3g::fd<double>(c, n, istep, lastiter) : s3g::fd<float>(c, n, istep, lastiter);
}

int s3g_settrap(s3g_ctx *c) {
  if (!c) return S3G_ERR_ARG;
  return c->precision == S3G_FP64 ? s3g::settrap<double>(c) : s3g::settrap<float>(c);
}

int s3g_settrap2(s3g_ctx *c) {
  if (!c) return S3G_ERR_ARG;
  return c->precision == S3G_FP64 ? s3g::settrap2<double>(c) : s3g::settrap2<float>(c);
}

int s3g_save(s3g_ctx *c, int istep) {
  if (!c || (istep != 1 && istep != 2)) return S3G_ERR_ARG;
  return c->precision == S3G_FP64 ? s3g::save<double>(c, istep) : s3g::save<float>(c, istep);
}

int s3g_phase(s3g_ctx *c, int ph, int n, int istep, double thrs) {
  if (!c || (istep != 1 && istep != 2)) return S3G_ERR_ARG;
  c->thrs = thrs;
  return c->precision == S3G_FP64 ? s3g::phase<double>(c, ph, n, istep) : s3g::phase<float>(c, ph, n, istep);
}

// One step of the si3d.f90 time loop (324-393).
int s3g_step(s3g_ctx *c, int n) {
  if (!c) return S3G_ERR_ARG;
  const double thrs = (double)n * c->P.dt / 3600.;
  int niter1 = c->P.niter, lastiter = 0, rc = 0, worst = 0;
  auto acc = [&](int r) {
    if (r < 0) return false;
    worst = std::max(worst, r);
    return true;
  };
  if (n == 1) {
    niter1 = std::max(c->P.niter, 2);
  } else {
    const bool lf_only = (c->P.itrap == 0) || (n % std::max(c->P.itrap, 1) != 0);
    if (lf_only) lastiter = 1;
    if (!acc(rc = s3g_fd(c, n, 1, 0, lastiter, thrs))) return rc;
    if (lf_only) {
      if (!acc(rc = s3g_save(c, 1))) return rc;
      c->stats.steps++;
      return worst;
    }
    if (!acc(rc = s3g_settrap(c))) return rc;
  }
  for (int iter = 1; iter <= niter1; ++iter) {
    if (iter == niter1) lastiter = 1;
    if (!acc(rc = s3g_fd(c, n, 2, iter, lastiter, thrs))) return rc;
    if (iter < niter1)
      if (!acc(rc = s3g_settrap2(c))) return rc;
  }
  if (!acc(rc = s3g_save(c, 2))) return rc;
  c->stats.steps++;
  return worst;
}

int s3g_run(s3g_ctx *c, int n0, int nsteps) {
  int worst = 0;
  for (int n = n0; n < n0 + nsteps; ++n) {
    int rc = s3g_step(c, n);
    if (rc < 0) return rc;
    worst = std::max(worst, rc);
  }
  return worst;
}

int s3g_synchronize(s3g_ctx *c) { return c ? hip_check(hipStreamSynchronize(c->stream), "sync") : S3G_ERR_ARG; }

int s3g_get_stats(s3g_ctx *c, s3g_stats *out) {
  if (!c || !out) return S3G_ERR_ARG;
  std::vector<double> s(c->lm + 1);
  if (int e = get_field(c, "s", s.data())) return e;
  double v = 0.0;
  for (int l = 0; l < c->lm; ++l) v += s[l];
  c->stats.zeta_volume = v;
  c->stats.graph_replays = c->seg_replays;
  *out = c->stats;
  return 0;
}

int s3g_energy_balance(s3g_ctx *c, double *out, int reset) {
  if (!c || !out) return S3G_ERR_ARG;
  if (c->P.iturb < 1 || c->nlakes != 1) {
    set_error("s3g_energy_balance: the basin energy budget needs the turbulence closure (iturb >= 1) and one lake");
    return S3G_ERR_UNSUPPORTED;
  }
  const double *tke = c->precision == S3G_FP64 ? c->d64.tke : c->d32.tke;
  double v[7];
  if (int e = hip_check(hipMemcpyAsync(v, tke, sizeof v, hipMemcpyDeviceToHost, c->stream), "energy balance"))
    return e;
  if (int e = hip_check(hipStreamSynchronize(c->stream), "energy balance")) return e;
  out[0] = v[3];
  out[1] = v[4];
  out[2] = v[5];
  out[3] = v[6];
  if (reset) {  // OutScalarEnergyBalance clears the integrals after writing them (si3d_utils.f90:3809-3811)
    static const double zero[3] = {0.0, 0.0, 0.0};
    if (int e = hip_check(hipMemcpyAsync(const_cast<double *>(tke) + 4, zero, sizeof zero, hipMemcpyHostToDevice,
                                         c->stream), "energy balance reset"))
      return e;
    return hip_check(hipStreamSynchronize(c->stream), "energy balance reset");
  }
  return 0;
}

int s3g_reset_timers(s3g_ctx *c, int enable) {
  if (!c) return S3G_ERR_ARG;
  if (!c->pending.empty()) timers_flush(c);
  c->timers.clear();
  c->timing = enable != 0;
  return 0;
}

int s3g_get_tiles(s3g_ctx *c, int max, int *kc) {
  if (!c || !kc) return S3G_ERR_ARG;
  const int n = std::min<int>(max, s3g::TILE_NUM);
  for (int i = 0; i < n; ++i) kc[i] = c->kc[i];
  return n;
}

int s3g_probe_stream(size_t bytes, int reps, double *gbs) {
  if (!gbs || reps < 1 || bytes < 16) return S3G_ERR_ARG;
  const size_t n = bytes / 16;
  float4 *a = nullptr, *b = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int rc = hip_check(hipMalloc(&a, n * 16), "s3g_probe_stream");
  if (!rc) rc = hip_check(hipMalloc(&b, n * 16), "s3g_probe_stream");
  if (!rc) rc = hip_check(hipMemset(a, 0, n * 16), "s3g_probe_stream");
  if (!rc) rc = hip_check(hipEventCreate(&e0), "s3g_probe_stream");
  if (!rc) rc = hip_check(hipEventCreate(&e1), "s3g_probe_stream");
  if (!rc) {
    const dim3 grid(8192), block(256);  // 32 workgroups per CU
    hipLaunchKernelGGL(s3g::k_stream_copy, grid, block, 0, nullptr, (const float4 *)a, b, n);
    hipEventRecord(e0, nullptr);
    for (int r = 0; r < reps; ++r) hipLaunchKernelGGL(s3g::k_stream_copy, grid, block, 0, nullptr, (const float4 *)a, b, n);
    hipEventRecord(e1, nullptr);
    rc = hip_check(hipEventSynchronize(e1), "s3g_probe_stream");
    float ms = 0.f;
    if (!rc) rc = hip_check(hipEventElapsedTime(&ms, e0, e1), "s3g_probe_stream");
    if (!rc) *gbs = ms > 0.f ? 2.0 * (double)n * 16.0 * reps / (ms * 1e-3) / 1e9 : 0.0;
  }
  if (e0) hipEventDestroy(e0);
  if (e1) hipEventDestroy(e1);
  if (a) hipFree(a);
  if (b) hipFree(b);
  return rc;
}

int s3g_kernel_times(s3g_ctx *c, int max, char (*names)[32], double *avg_ms, long *launches) {
  if (!c) return S3G_ERR_ARG;
  timers_flush(c);
  int n = std::min<int>(max, (int)c->timers.size());
  for (int i = 0; i < n; ++i) {
    std::snprintf(names[i], 32, "%s", c->timers[i].name.c_str());
    avg_ms[i] = c->timers[i].n ? c->timers[i].total_ms / c->timers[i].n : 0.0;
    launches[i] = c->timers[i].n;
  }
  return n;
}

void s3g_finalize(s3g_ctx *c) {
  if (!c) return;
  if (c->stream) hipStreamSynchronize(c->stream);
  for (auto &e : c->cg_exec)
    if (e) hipGraphExecDestroy(e);
  for (auto &a : c->seg_exec)
    for (auto &b : a)
      for (auto &e : b)
        if (e) hipGraphExecDestroy(e);
  s3g::learn_destroy(c);
  for (void *p : c->allocs) hipFree(p);
  for (auto &e : c->ev_pool) {
    hipEventDestroy(e.first);
    hipEventDestroy(e.second);
  }
  if (c->h_iters) hipHostFree(c->h_iters);
  if (c->h_scal) hipHostFree(c->h_scal);
  if (c->pcg.lake_off) hipFree(c->pcg.lake_off);
  if (c->pcg.lake_part) hipFree(c->pcg.lake_part);
  if (c->pcg.h_lake) hipHostFree(c->pcg.h_lake);
  if (c->stream) hipStreamDestroy(c->stream);
  delete c;
}

}  // extern "C"
