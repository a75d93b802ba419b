// s3g_deck.cpp -- standalone driver set-up: what the Fortran host does before
// the time loop, restated in C++ for callers without the Fortran driver.
//
//   ZGridDimensions  si3d_utils.f90:318-382   (zlevel, km1)
//   bathy            si3d_utils.f90:595-892   (mask2d, l numbering, neighbours,
//                                              kmz, hhs, hhu, hhv)
//   init             si3d_procedures.f90:23-181 (zeta, hp/hup/hvp, k1z/k1u/k1v)
//   InitializeScalarFields si3d_procedures.f90:292-377 (profile from
//                                              si3d_init.txt; density on device)
//   InitializeTurbulenceModel si3d_mixing.f90:615-724 (cases 0 and 1)
//
// The grid arithmetic is FP32, as the reference computes it, so that masks and
// layer indices come out bit-identical to the Fortran driver's.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/si3d_gpu.h"

namespace s3g {
void set_error(const std::string &m);
int init_density(s3g_ctx *c);
}  // namespace s3g

namespace {

struct HostLake {
  int im1 = 0, jm1 = 0, km1 = 0, km = 0, lm = 0;
  std::vector<float> zlevel;  // [km1], 0-based: zlevel[k-1]
  std::vector<int> l2i, l2j, lEC, lWC, lNC, lSC, kmz, k1z, k1u, k1v;  // [lm1], 1-based values
  std::vector<float> hhs, hhu, hhv;
  std::vector<float> s;                       // [lm1]
  std::vector<float> hp, hup, hvp;            // (km1, lm1) k-fastest
};

bool build_host_lake(const s3g_deck *d, HostLake &L) {
  const int k1 = 2;
  // ZGridDimensions
  if (d->ibathyf < 0) {
    if (!d->zlevel_file || d->km1_file < 3) return false;
    L.km1 = d->km1_file;
    for (int k = 0; k < L.km1; ++k) L.zlevel.push_back((float)d->zlevel_file[k]);
  } else {
    const float zl = (float)d->zl, ddz = (float)d->ddz;
    L.km1 = (int)std::ceil(zl / ddz) + 2;
    L.zlevel.assign(L.km1, 0.0f);
    for (int k = k1 + 1; k <= L.km1; ++k) L.zlevel[k - 1] = L.zlevel[k - 2] + ddz;
    L.zlevel[k1 - 1] = -100.0f;
    L.zlevel[0] = -100.0f;
  }
  L.km = L.km1 - 1;
  auto Z = [&](int k) { return L.zlevel[k - 1]; };
  // bathy
  const int im = d->imm + 1, jm = d->jmm + 1;
  L.im1 = im + 1;
  L.jm1 = jm + 1;
  const int W = L.jm1 + 2;
  std::vector<float> h4((size_t)(L.im1 + 2) * W, 0.0f);
  for (int i = 2; i <= im; ++i)
    for (int j = 2; j <= jm; ++j)
      h4[(size_t)i * W + j] = (float)d->bathy_dm[(size_t)(i - 2) * d->jmm + (j - 2)] * 0.1f + (float)d->datadj;
  for (int j = 1; j <= L.jm1; ++j) h4[(size_t)1 * W + j] = h4[(size_t)L.im1 * W + j] = 0.0f;
  for (int i = 1; i <= L.im1; ++i) h4[(size_t)i * W + 1] = h4[(size_t)i * W + L.jm1] = 0.0f;
  auto wet = [&](int i, int j) { return h4[(size_t)i * W + j] > 0.0f; };
  std::vector<int> ij2l((size_t)(L.im1 + 2) * W, 0);
  int l = 0;
  for (int i = 1; i <= L.im1; ++i)
    for (int j = 1; j <= L.jm1; ++j)
      if (wet(i, j)) ij2l[(size_t)i * W + j] = ++l;
  L.lm = l;
  const int lm1 = l + 1;
  for (auto &v : ij2l)
    if (v == 0) v = lm1;
  for (auto *v : {&L.l2i, &L.l2j, &L.kmz}) v->assign(lm1, 0);
  for (auto *v : {&L.lEC, &L.lWC, &L.lNC, &L.lSC}) v->assign(lm1, lm1);
  for (int i = 1; i <= L.im1; ++i)
    for (int j = 1; j <= L.jm1; ++j) {
      const int ll = ij2l[(size_t)i * W + j];
      if (ll == lm1) continue;
      L.l2i[ll - 1] = i;
      L.l2j[ll - 1] = j;
      L.lEC[ll - 1] = ij2l[(size_t)(i + 1) * W + j];
      L.lWC[ll - 1] = ij2l[(size_t)(i - 1) * W + j];
      L.lNC[ll - 1] = ij2l[(size_t)i * W + j + 1];
      L.lSC[ll - 1] = ij2l[(size_t)i * W + j - 1];
    }
  L.hhs.assign(lm1, 0.0f);
  const float ddz = (float)d->ddz, dzmin = (float)d->dzmin;
  for (int ll = 1; ll <= L.lm; ++ll) {
    const float hs1 = h4[(size_t)L.l2i[ll - 1] * W + L.l2j[ll - 1]];
    int kz = 0;
    float hh = 0.0f;
    if (d->ibathyf >= 0) {
      kz = (int)std::floor(hs1 / ddz);
      if ((hs1 - (float)kz * ddz) > dzmin) {
        kz += 1;
        hh = hs1;
      } else {
        hh = (float)kz * ddz;
      }
      kz += 1;
    } else {
      for (int k = k1; k <= L.km; ++k)
        if (Z(k + 1) >= hs1) {
          hh = Z(k) + std::max(dzmin, hs1 - Z(k));
          kz = k;
          break;
        }
      if (kz == 0) {
        s3g::set_error("bathymetry deeper than the deepest layer");
        return false;
      }
    }
    L.kmz[ll - 1] = kz;
    L.hhs[ll - 1] = hh;
  }
  L.kmz[lm1 - 1] = L.km1;
  L.hhu.assign(lm1, 0.0f);
  L.hhv.assign(lm1, 0.0f);
  for (int ll = 1; ll <= L.lm; ++ll) {
    const int e = L.lEC[ll - 1], n = L.lNC[ll - 1];
    if (e != lm1) L.hhu[ll - 1] = std::min(L.hhs[e - 1], L.hhs[ll - 1]);
    if (n != lm1) L.hhv[ll - 1] = std::min(L.hhs[n - 1], L.hhs[ll - 1]);
  }
  // init: level surface, layer thicknesses, top indices
  const int km1 = L.km1;
  L.s.assign(lm1, (float)d->zetainit);
  L.hp.assign((size_t)km1 * lm1, 0.0f);
  L.hup.assign((size_t)km1 * lm1, 0.0f);
  L.hvp.assign((size_t)km1 * lm1, 0.0f);
  L.k1z.assign(lm1, km1);
  L.k1u.assign(lm1, km1);
  L.k1v.assign(lm1, km1);
  auto at = [&](std::vector<float> &a, int k, int ll) -> float & { return a[(size_t)(ll - 1) * km1 + (k - 1)]; };
  for (int ll = 1; ll <= L.lm; ++ll) {
    const int kms = L.kmz[ll - 1];
    const float sl = L.s[ll - 1];
    for (int k = k1; k <= kms; ++k) {
      float v = std::min(Z(k + 1), L.hhs[ll - 1]) - std::max(Z(k), -sl);
      if (v > 0.0f) {
        if (L.k1z[ll - 1] == km1) L.k1z[ll - 1] = k;
        at(L.hp, k, ll) = v;
      }
    }
    if (kms + 1 <= km1) at(L.hp, kms + 1, ll) = 0.5f;  // sediment layer, sed_h
    if (L.k1z[ll - 1] == km1) L.s[ll - 1] = -L.hhs[ll - 1];
    const int e = L.lEC[ll - 1], n = L.lNC[ll - 1];
    if (e != lm1) {
      const int kmx = std::min(kms, L.kmz[e - 1]);
      for (int k = k1; k <= kmx; ++k) {
        float v = std::min(Z(k + 1), L.hhu[ll - 1]) - std::max(Z(k), -(sl + L.s[e - 1]) / 2.0f);
        if (v > 0.0f) {
          if (L.k1u[ll - 1] == km1) L.k1u[ll - 1] = k;
          at(L.hup, k, ll) = v;
        }
      }
    }
    if (n != lm1) {
      const int kmy = std::min(kms, L.kmz[n - 1]);
      for (int k = k1; k <= kmy; ++k) {
        float v = std::min(Z(k + 1), L.hhv[ll - 1]) - std::max(Z(k), -(sl + L.s[n - 1]) / 2.0f);
        if (v > 0.0f) {
          if (L.k1v[ll - 1] == km1) L.k1v[ll - 1] = k;
          at(L.hvp, k, ll) = v;
        }
      }
    }
  }
  return true;
}

}  // namespace

extern "C" int s3g_create_from_deck(const s3g_deck *d, const s3g_params *p, s3g_ctx **out) {
  if (!d || !p || !out || !d->bathy_dm || !d->init_profile) {
    s3g::set_error("s3g_create_from_deck: missing deck arrays");
    return S3G_ERR_ARG;
  }
  HostLake L;
  if (!build_host_lake(d, L)) return S3G_ERR_ARG;
  if (L.lm == 0) {
    s3g::set_error("s3g_create_from_deck: no wet cells");
    return S3G_ERR_ARG;
  }
  s3g_grid g{};
  g.im1 = L.im1;
  g.jm1 = L.jm1;
  g.km1 = L.km1;
  g.lm = L.lm;
  g.l2i = L.l2i.data();
  g.l2j = L.l2j.data();
  g.lEC = L.lEC.data();
  g.lWC = L.lWC.data();
  g.lNC = L.lNC.data();
  g.lSC = L.lSC.data();
  g.kmz = L.kmz.data();
  g.k1z = L.k1z.data();
  g.k1u = L.k1u.data();
  g.k1v = L.k1v.data();
  g.zlevel = L.zlevel.data();
  g.hhs = L.hhs.data();
  g.hhu = L.hhu.data();
  g.hhv = L.hhv.data();
  s3g_ctx *c = nullptr;
  int rc = s3g_init(&g, p, &c);
  if (rc < 0) return rc;
  const int km1 = L.km1, lm1 = L.lm + 1;
  // zeta and thicknesses at all three time levels (init 42-65, 136-138)
  for (const char *n : {"s", "sp", "spp"}) rc |= s3g_set_field_f32(c, n, L.s.data());
  for (const char *n : {"hp", "hpp", "h"}) rc |= s3g_set_field_f32(c, n, L.hp.data());
  for (const char *n : {"hup", "hupp", "hu"}) rc |= s3g_set_field_f32(c, n, L.hup.data());
  for (const char *n : {"hvp", "hvpp", "hv"}) rc |= s3g_set_field_f32(c, n, L.hvp.data());
  // active scalar profile on every column (InitializeScalarFields 350-353)
  std::vector<double> sal((size_t)km1 * lm1);
  for (int ll = 0; ll < lm1; ++ll)
    for (int k = 0; k < km1; ++k) sal[(size_t)ll * km1 + k] = d->init_profile[k];
  for (const char *n : {"sal", "salp", "salpp"}) rc |= s3g_set_field_f64(c, n, sal.data());
  // eddy coefficients / turbulence
  std::vector<double> a((size_t)km1 * lm1, 0.0);
  if (p->iturb == 0) {
    std::vector<double> dv((size_t)km1 * lm1, 0.0);
    for (int ll = 1; ll <= L.lm; ++ll)
      for (int k = L.k1z[ll - 1] + 1; k <= L.kmz[ll - 1]; ++k) {
        a[(size_t)(ll - 1) * km1 + (k - 1)] = p->Av0;
        dv[(size_t)(ll - 1) * km1 + (k - 1)] = p->Dv0;
      }
    rc |= s3g_set_field_f64(c, "Av", a.data());
    rc |= s3g_set_field_f64(c, "Dv", dv.data());
  } else {
    const double q2min = 2.e-6, kappa = 0.41;
    std::vector<double> q2((size_t)km1 * lm1, q2min), q2l((size_t)km1 * lm1, 0.0);
    for (int ll = 1; ll <= L.lm; ++ll) {
      const int kms = L.kmz[ll - 1], k1s = L.k1z[ll - 1];
      if (kms <= k1s) continue;
      double htot = 0.0;
      for (int k = 1; k <= km1; ++k) htot += L.hp[(size_t)(ll - 1) * km1 + (k - 1)];
      double zt = 0.0;
      for (int k = k1s + 1; k <= kms; ++k) {
        zt += L.hp[(size_t)(ll - 1) * km1 + (k - 2)];
        const double mixl = kappa * zt * std::sqrt(std::max(1. - zt / htot, 0.0));
        q2l[(size_t)(ll - 1) * km1 + (k - 1)] = q2min * mixl;
      }
    }
    for (const char *n : {"q2", "q2p", "q2pp"}) rc |= s3g_set_field_f64(c, n, q2.data());
    for (const char *n : {"q2l", "q2lp", "q2lpp"}) rc |= s3g_set_field_f64(c, n, q2l.data());
  }
  if (rc == 0) rc = s3g::init_density(c);
  if (rc != 0) {
    s3g_finalize(c);
    return rc < 0 ? rc : S3G_ERR_HIP;
  }
  *out = c;
  return 0;
}
