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
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "s3g_internal.h"

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

// Initial state of one lake (init 42-65, 136-138; InitializeScalarFields
// 350-353; InitializeTurbulenceModel cases 0/1), arrays (km1, lm1) k fastest.
struct HostFields {
  std::vector<float> s, hp, hup, hvp;
  std::vector<double> sal, Av, Dv, q2, q2l;
  std::vector<std::vector<double>> tr;  // per tracer
};

void lake_fields(const HostLake &L, const s3g_deck *d, const s3g_params *p, HostFields &F) {
  const int km1 = L.km1, lm1 = L.lm + 1;
  F.s = L.s;
  F.hp = L.hp;
  F.hup = L.hup;
  F.hvp = L.hvp;
  F.sal.assign((size_t)km1 * lm1, 0.0);
  for (int ll = 0; ll < lm1; ++ll)
    for (int k = 0; k < km1; ++k) F.sal[(size_t)ll * km1 + k] = d->init_profile[k];
  // tracers: profile column nn+1 of si3d_init.txt on every column (330-356)
  F.tr.assign(std::max(0, p->ntr), std::vector<double>((size_t)km1 * lm1, 0.0));
  for (int nt = 0; nt < p->ntr && d->tracer_profiles; ++nt)
    for (int ll = 0; ll < lm1; ++ll)
      for (int k = 0; k < km1; ++k) F.tr[nt][(size_t)ll * km1 + k] = d->tracer_profiles[(size_t)nt * km1 + k];
  F.Av.assign((size_t)km1 * lm1, 0.0);
  F.Dv.assign((size_t)km1 * lm1, 0.0);
  if (p->iturb == 0) {
    for (int ll = 1; ll <= L.lm; ++ll)
      for (int k = L.k1z[ll - 1] + 1; k <= L.kmz[ll - 1]; ++k) {
        F.Av[(size_t)(ll - 1) * km1 + (k - 1)] = p->Av0;
        F.Dv[(size_t)(ll - 1) * km1 + (k - 1)] = p->Dv0;
      }
  } else {
    const double q2min = 2.e-6, kappa = 0.41;
    F.q2.assign((size_t)km1 * lm1, q2min);
    F.q2l.assign((size_t)km1 * lm1, 0.0);
    for (int ll = 1; ll <= L.lm; ++ll) {
      const int kms = L.kmz[ll - 1], k1s = L.k1z[ll - 1];
      if (kms <= k1s) continue;
      double htot = 0.0;
      for (int k = 1; k <= km1; ++k) htot += L.hp[(size_t)(ll - 1) * km1 + (k - 1)];
      double zt = 0.0;
      for (int k = k1s + 1; k <= kms; ++k) {
        zt += L.hp[(size_t)(ll - 1) * km1 + (k - 2)];
        const double mixl = kappa * zt * std::sqrt(std::max(1. - zt / htot, 0.0));
        F.q2l[(size_t)(ll - 1) * km1 + (k - 1)] = q2min * mixl;
      }
    }
  }
}

int upload_fields(s3g_ctx *c, const HostFields &F, int iturb) {
  int rc = 0;
  for (const char *n : {"s", "sp", "spp"}) rc |= s3g_set_field_f32(c, n, F.s.data());
  for (const char *n : {"hp", "hpp", "h"}) rc |= s3g_set_field_f32(c, n, F.hp.data());
  for (const char *n : {"hup", "hupp", "hu"}) rc |= s3g_set_field_f32(c, n, F.hup.data());
  for (const char *n : {"hvp", "hvpp", "hv"}) rc |= s3g_set_field_f32(c, n, F.hvp.data());
  for (const char *n : {"sal", "salp", "salpp"}) rc |= s3g_set_field_f64(c, n, F.sal.data());
  if (iturb == 0) {
    rc |= s3g_set_field_f64(c, "Av", F.Av.data());
    rc |= s3g_set_field_f64(c, "Dv", F.Dv.data());
  } else {
    for (const char *n : {"q2", "q2p", "q2pp"}) rc |= s3g_set_field_f64(c, n, F.q2.data());
    for (const char *n : {"q2l", "q2lp", "q2lpp"}) rc |= s3g_set_field_f64(c, n, F.q2l.data());
  }
  for (size_t nt = 0; nt < F.tr.size(); ++nt) {
    rc |= s3g_set_field_f64(c, ("tracer" + std::to_string(nt + 1)).c_str(), F.tr[nt].data());
    rc |= s3g_set_field_f64(c, ("tracerpp" + std::to_string(nt + 1)).c_str(), F.tr[nt].data());
  }
  if (rc == 0) rc = s3g::init_density(c);
  return rc;
}

s3g_grid grid_of(const HostLake &L) {
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
  return g;
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
  const s3g_grid g = grid_of(L);
  s3g_ctx *c = nullptr;
  int rc = s3g_init(&g, p, &c);
  if (rc < 0) return rc;
  HostFields F;
  lake_fields(L, d, p, F);
  rc = upload_fields(c, F, p->iturb);
  if (rc != 0) {
    s3g_finalize(c);
    return rc < 0 ? rc : S3G_ERR_HIP;
  }
  *out = c;
  return 0;
}

// The lakes of an ensemble share one column space: lake i's columns follow
// lake i-1's, its neighbours point inside the lake or to the one shared dry
// sentinel, and its (i, j) box is placed beside the previous lake's (at an
// i offset aligned to 64) so that the multigrid's aggregates of the first
// levels never span two lakes.  Layers are numbered as in each lake; planes
// beyond a lake's km1 are zero padding.  Each lake keeps its own layer depths
// (zlevel) and forcing.
extern "C" int s3g_create_ensemble(int nlakes, const s3g_deck *decks, const s3g_params *params, s3g_ctx **out) {
  if (nlakes < 1 || !decks || !params || !out) {
    s3g::set_error("s3g_create_ensemble: invalid arguments");
    return S3G_ERR_ARG;
  }
  const s3g_params &p0 = params[0];
  for (int i = 1; i < nlakes; ++i) {
    const s3g_params &q = params[i];
    const bool same = q.dt == p0.dt && q.dx == p0.dx && q.dy == p0.dy && q.itrap == p0.itrap &&
                      q.niter == p0.niter && q.ismooth == p0.ismooth && q.iturb == p0.iturb && q.iadv == p0.iadv &&
                      q.itrmom == p0.itrmom && q.ihd == p0.ihd && q.ibc == p0.ibc && q.isal == p0.isal &&
                      q.beta == p0.beta && q.Av0 == p0.Av0 && q.Dv0 == p0.Dv0 && q.Ax0 == p0.Ax0 &&
                      q.Ay0 == p0.Ay0 && q.f == p0.f && q.theta == p0.theta && q.cd == p0.cd &&
                      q.precision == p0.precision && q.preconditioner == p0.preconditioner &&
                      q.pcg_tol == p0.pcg_tol && q.pcg_maxit == p0.pcg_maxit && q.ntr == p0.ntr;
    if (!same) {
      s3g::set_error("s3g_create_ensemble: lakes may differ only in their decks and forcing");
      return S3G_ERR_UNSUPPORTED;
    }
  }
  std::vector<HostLake> Ls(nlakes);
  int km1 = 0, lm = 0, ioff = 0, jm1 = 0;
  std::vector<int> off(nlakes + 1, 0), io(nlakes, 0);
  for (int i = 0; i < nlakes; ++i) {
    if (!decks[i].bathy_dm || !decks[i].init_profile || !build_host_lake(&decks[i], Ls[i])) {
      s3g::set_error("s3g_create_ensemble: lake " + std::to_string(i) + ": invalid deck");
      return S3G_ERR_ARG;
    }
    if (Ls[i].lm == 0) {
      s3g::set_error("s3g_create_ensemble: lake " + std::to_string(i) + " has no wet cells");
      return S3G_ERR_ARG;
    }
    km1 = std::max(km1, Ls[i].km1);
    off[i] = lm;
    lm += Ls[i].lm;
    io[i] = ioff;
    ioff = (ioff + Ls[i].im1 + 2 + 63) / 64 * 64;
    jm1 = std::max(jm1, Ls[i].jm1);
  }
  off[nlakes] = lm;
  const int lm1 = lm + 1, KP = km1 + 3;
  HostLake G;
  G.im1 = ioff;
  G.jm1 = jm1;
  G.km1 = km1;
  G.km = km1 - 1;
  G.lm = lm;
  G.zlevel = Ls[0].zlevel;
  G.zlevel.resize(km1, 1e30f);
  for (auto *v : {&G.l2i, &G.l2j, &G.kmz, &G.k1z, &G.k1u, &G.k1v}) v->assign(lm1, 0);
  for (auto *v : {&G.lEC, &G.lWC, &G.lNC, &G.lSC}) v->assign(lm1, lm1);
  for (auto *v : {&G.hhs, &G.hhu, &G.hhv}) v->assign(lm1, 0.0f);
  std::vector<double> zl((size_t)nlakes * KP, 1e30);
  std::vector<int> lake_of(lm1, 0), lake_km1(nlakes);
  HostFields GF;
  GF.s.assign(lm1, 0.0f);
  for (auto *v : {&GF.hp, &GF.hup, &GF.hvp}) v->assign((size_t)km1 * lm1, 0.0f);
  for (auto *v : {&GF.sal, &GF.Av, &GF.Dv}) v->assign((size_t)km1 * lm1, 0.0);
  if (p0.iturb > 0) {
    GF.q2.assign((size_t)km1 * lm1, 0.0);
    GF.q2l.assign((size_t)km1 * lm1, 0.0);
  }
  GF.tr.assign(std::max(0, p0.ntr), std::vector<double>((size_t)km1 * lm1, 0.0));
  const bool verbose = std::getenv("S3G_DEBUG") != nullptr;
  for (int i = 0; i < nlakes; ++i) {
    const HostLake &L = Ls[i];
    const int li1 = L.lm + 1, o = off[i];
    if (verbose) std::fprintf(stderr, "[s3g] ensemble: lake %d of %d (%d columns)\n", i + 1, nlakes, L.lm);
    lake_km1[i] = L.km1;
    zl[(size_t)i * KP] = -100.0;
    for (int k = 1; k <= L.km1; ++k) zl[(size_t)i * KP + k] = L.zlevel[k - 1];
    auto nb = [&](int v) { return v == li1 ? lm1 : v + o; };
    for (int l = 0; l < L.lm; ++l) {
      const int g = o + l;
      G.l2i[g] = L.l2i[l] + io[i];
      G.l2j[g] = L.l2j[l];
      G.lEC[g] = nb(L.lEC[l]);
      G.lWC[g] = nb(L.lWC[l]);
      G.lNC[g] = nb(L.lNC[l]);
      G.lSC[g] = nb(L.lSC[l]);
      G.kmz[g] = L.kmz[l];
      G.k1z[g] = L.k1z[l];
      G.k1u[g] = L.k1u[l];
      G.k1v[g] = L.k1v[l];
      G.hhs[g] = L.hhs[l];
      G.hhu[g] = L.hhu[l];
      G.hhv[g] = L.hhv[l];
      lake_of[g] = i;
    }
    HostFields F;
    lake_fields(L, &decks[i], &params[i], F);
    // columns 0..lm-1 of the lake, plus lake 0's sentinel column for the shared one
    for (int l = 0; l <= L.lm; ++l) {
      if (l == L.lm && i != 0) continue;
      const int g = (l == L.lm) ? lm : o + l;
      if (l < L.lm || i == 0) GF.s[g] = F.s[l];
      for (int k = 0; k < L.km1; ++k) {
        const size_t src = (size_t)l * L.km1 + k, dst = (size_t)g * km1 + k;
        GF.hp[dst] = F.hp[src];
        GF.hup[dst] = F.hup[src];
        GF.hvp[dst] = F.hvp[src];
        GF.sal[dst] = F.sal[src];
        GF.Av[dst] = F.Av[src];
        GF.Dv[dst] = F.Dv[src];
        if (p0.iturb > 0) {
          GF.q2[dst] = F.q2[src];
          GF.q2l[dst] = F.q2l[src];
        }
        for (size_t nt = 0; nt < GF.tr.size(); ++nt) GF.tr[nt][dst] = F.tr[nt][src];
      }
    }
  }
  G.kmz[lm] = km1;
  G.k1z[lm] = G.k1u[lm] = G.k1v[lm] = km1;
  const s3g_grid g = grid_of(G);
  s3g_ctx *c = nullptr;
  int rc = s3g::init_ctx(&g, &p0, nlakes, zl.data(), lake_of.data(), lake_km1.data(), &c);
  if (rc < 0) return rc;
  c->lake_off = off;
  c->lake_km1 = lake_km1;
  c->lake_forcing.resize(nlakes);
  c->lake_force.assign(3 * (size_t)nlakes, 0.0);
  for (int i = 0; i < nlakes; ++i) {
    auto &F = c->lake_forcing[i];
    const s3g_params &q = params[i];
    F.ifsurfbc = q.ifsurfbc;
    F.cw = q.cw;
    F.wa = q.wa;
    F.phi = q.phi;
    F.dtsurfbc = q.dtsurfbc;
    if (q.n_surfbc > 0 && q.surfbc) F.surfbc.assign(q.surfbc, q.surfbc + 3 * q.n_surfbc);
    if (q.ifsurfbc != 0 && q.ifsurfbc != 20) {
      s3g_finalize(c);
      s3g::set_error("s3g_create_ensemble: only ifsurfbc = 0 or 20 is supported");
      return S3G_ERR_UNSUPPORTED;
    }
    if (q.ifsurfbc == 20 && F.surfbc.size() >= 3) {
      c->lake_force[3 * i + 2] = F.surfbc[0];
      c->lake_force[3 * i] = F.surfbc[1];
      c->lake_force[3 * i + 1] = F.surfbc[2];
    }
  }
  rc = s3g::set_forcing(c);
  if (rc == 0) rc = upload_fields(c, GF, p0.iturb);
  if (rc != 0) {
    s3g_finalize(c);
    return rc < 0 ? rc : S3G_ERR_HIP;
  }
  *out = c;
  return 0;
}

extern "C" int s3g_ensemble_info(s3g_ctx *c, int max, int *col_offset, int *km1) {
  if (!c) return S3G_ERR_ARG;
  const int n = c->nlakes;
  if (max < n) return n;
  for (int i = 0; i < n; ++i) {
    if (col_offset) col_offset[i] = c->lake_off.empty() ? 0 : c->lake_off[i];
    if (km1) km1[i] = c->lake_km1.empty() ? c->km1 : c->lake_km1[i];
  }
  if (col_offset) col_offset[n] = c->lake_off.empty() ? c->lm : c->lake_off[n];
  return n;
}
