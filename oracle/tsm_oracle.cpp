/*
 * tsm_oracle.cpp -- CPU oracle for the vertex-centred FV residual + RK loop.
 *
 * TEST INFRASTRUCTURE ONLY (see tsm_oracle.h).  This file restates, in plain
 * C++ and in the reference's own loop and summation order, the algorithm of
 * deepray/TENSUM for one mesh partition (NPART = 1).  Each function names the
 * reference lines it follows.  The z components of the reference's 3-D
 * vectors are identically zero in 2-D; every term they contribute is an exact
 * +-0, so dropping them leaves every result bit-for-bit unchanged.
 *
 * Build: oracle/Makefile (g++ -O2, no FMA contraction on x86-64 by default,
 * which is also how the reference is compiled).
 */
#include "tsm_oracle.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <string>
#include <vector>

namespace {

thread_local std::string g_err;

struct V2 { double x, y; };
struct Prim { double T, u, v, p; };
struct Ent { double f, vx, vy, l; };
struct Flx { double m, mx, my, e; };

enum { BC_NONE = 0, BC_SLIP, BC_NOSLIP, BC_FARFIELD, BC_INLET, BC_OUTLET,
       BC_PRESSURE, BC_PERIODIC };
enum { F_ROE = 0, F_RUS, F_KEPEC, F_ROEEC, F_KEP, F_AVG };
enum { R_FIRST = 0, R_SECOND, R_MUSCL, R_ALBADA, R_MINMOD, R_TVB, R_ENO2 };

const double a_rk[] = {0.0, 3.0 / 4.0, 1.0 / 3.0};   // src/constants.h:4
const double b_rk[] = {1.0, 1.0 / 4.0, 2.0 / 3.0};   // src/constants.h:6
const double rk4_rks[] = {1.0, 1.0 / 2.0, 1.0 / 2.0, 1.0};  // src/constants.h:9
const double KKK = 1.0 / 3.0;                        // src/constants.h:12
const double EPSILON = 1.0e-14;                      // src/reconstruct.cc:5

struct Mat {
  orc_params p;
  double Cp;
  double density(const Prim& s) const { return s.p / (p.gas_const * s.T); }
  double sound_speed(const Prim& s) const { return std::sqrt(p.gamma * p.gas_const * s.T); }
  double entropy(const Prim& s) const { return std::log(s.p) - p.gamma * std::log(density(s)); }
  double mach(const Prim& s) const {
    double c = sound_speed(s);
    return std::sqrt(s.u * s.u + s.v * s.v) / c;
  }
  // src/material.h:551-568
  double viscosity(double T) const {
    switch (p.mu_model) {
      case 0: return p.mu_ref;
      case 1: return p.mu_ref * std::pow(T, 1.5) / (T + p.T_0);
      default: return p.mu_ref * std::pow(T / p.T_ref, p.omega);
    }
  }
};

// src/material.h:24-42
inline double logavg(double a, double b) {
  double xi = b / a;
  double f = (xi - 1.0) / (xi + 1.0);
  double u = f * f;
  double F;
  if (u < 1.0e-2) {
    double u2 = u * u;
    double u3 = u2 * u;
    F = 1.0 + u / 3.0 + u2 / 5.0 + u3 / 7.0;
  } else {
    F = std::log(xi) / 2.0 / f;
  }
  return 0.5 * (a + b) / F;
}

// src/material.h:501-513
Flx prim2con(const Mat& M, const Prim& w) {
  Flx c;
  double rho = M.density(w);
  c.m = rho;
  c.mx = w.u * rho;
  c.my = w.v * rho;
  c.e = w.p / (M.p.gamma - 1.0) + 0.5 * (w.u * w.u + w.v * w.v) * rho;
  return c;
}
// src/material.h:518-529
Prim con2prim(const Mat& M, const Flx& c) {
  Prim w;
  double ir = 1.0 / c.m;
  w.u = c.mx * ir;
  w.v = c.my * ir;
  w.p = (M.p.gamma - 1.0) * (c.e - 0.5 * (c.mx * c.mx + c.my * c.my) / c.m);
  w.T = w.p / (M.p.gas_const * c.m);
  return w;
}
// src/material.h:534-546
Ent prim2ent(const Mat& M, const Prim& w) {
  Ent e;
  double rho = M.density(w);
  double beta = rho / (2 * w.p);
  double s = M.entropy(w);
  e.f = (M.p.gamma - s) / (M.p.gamma - 1) - (w.u * w.u + w.v * w.v) * beta;
  e.vx = w.u * (2 * beta);
  e.vy = w.v * (2 * beta);
  e.l = -2 * beta;
  return e;
}

// src/limiter.h:14-27, :32-43, :66-77
inline double minmod_tvd(double a, double b) {
  if (a < 0.0 && b < 0.0) return -std::min(-a, -b);
  if (a > 0.0 && b > 0.0) return std::min(a, b);
  return 0.0;
}
inline double minabs(double a, double b) { return (std::fabs(a) < std::fabs(b)) ? a : b; }
inline double minmod_tvb(double a, double b, double h, double Mv) {
  double c = 0.5 * (a + b);
  if (std::fabs(c) < Mv * std::pow(h, 2)) return c;
  return minmod_tvd(a, b);
}
// src/reconstruct.cc:101-134
double limited_slope(const orc_params& p, double ul, double ur, double h) {
  switch (p.reconstruct) {
    case R_ALBADA: {
      double s = std::max(0.0, (2 * ul * ur + EPSILON) / (ul * ul + ur * ur + EPSILON));
      return 0.25 * s * ((1.0 - KKK * s) * ul + (1.0 + KKK * s) * ur);
    }
    case R_MINMOD: return 0.5 * minmod_tvd(ul, ur);
    case R_ENO2: return 0.5 * minabs(ul, ur);
    case R_TVB: return 0.5 * minmod_tvb(ul, ur, h, p.tvb_M);
  }
  return 0.0;
}
// src/reconstruct.cc:13-96, :139-188 (4 active components)
void reconstruct(const orc_params& p, double h, const double* zl, const double* zr,
                 const double* gl, const double* gr, double* rl, double* rr) {
  for (int i = 0; i < 4; ++i) {
    switch (p.reconstruct) {
      case R_FIRST:
        rl[i] = zl[i]; rr[i] = zr[i];
        break;
      case R_SECOND:
        rl[i] = zl[i] + 0.5 * gl[i];
        rr[i] = zr[i] - 0.5 * gr[i];
        break;
      case R_MUSCL: {
        double ds = zr[i] - zl[i];
        double dl = 2.0 * gl[i] - ds;
        rl[i] = zl[i] + (dl * (1.0 - KKK) + ds * (1.0 + KKK)) * 0.25;
        double dr = 2.0 * gr[i] - ds;
        rr[i] = zr[i] - (dr * (1.0 - KKK) + ds * (1.0 + KKK)) * 0.25;
        break;
      }
      default: {
        double ds = zr[i] - zl[i];
        double dl = 2.0 * gl[i] - ds;
        double si = limited_slope(p, dl, ds, h);
        rl[i] = zl[i] + si;
        double dr = 2.0 * gr[i] - ds;
        double sj = limited_slope(p, dr, ds, h);
        rr[i] = zr[i] - sj;
      }
    }
  }
}

// src/kepec.cc:10-38 (per unit normal)
Flx kepec_flux(const Mat& M, const Prim& l, const Prim& r, V2 n) {
  double area = std::sqrt(n.x * n.x + n.y * n.y);
  double ia = 1.0 / area;
  V2 un{n.x * ia, n.y * ia};
  double rhol = M.density(l), rhor = M.density(r);
  double rho = logavg(rhol, rhor);
  double vel2 = 0.5 * ((l.u * l.u + l.v * l.v) + (r.u * r.u + r.v * r.v));
  double betal = 0.5 / (M.p.gas_const * l.T);
  double betar = 0.5 / (M.p.gas_const * r.T);
  double beta = logavg(betal, betar);
  double p = 0.5 * (rhol + rhor) / (betal + betar);
  V2 vel{l.u * 0.5 + r.u * 0.5, l.v * 0.5 + r.v * 0.5};
  double vn = vel.x * un.x + vel.y * un.y;
  Flx F;
  F.m = rho * vn;
  F.mx = un.x * p + vel.x * F.m;
  F.my = un.y * p + vel.y * F.m;
  F.e = 0.5 * (1.0 / ((M.p.gamma - 1.0) * beta) - vel2) * F.m + (F.mx * vel.x + F.my * vel.y);
  return F;
}
// src/kep.cc:10-38
Flx kep_flux(const Mat& M, const Prim& l, const Prim& r, V2 n) {
  double area = std::sqrt(n.x * n.x + n.y * n.y);
  double ia = 1.0 / area;
  V2 un{n.x * ia, n.y * ia};
  double rhol = M.density(l), rhor = M.density(r);
  double g1 = M.p.gamma - 1.0;
  double Hl = std::pow(M.sound_speed(l), 2.0) / g1 + 0.5 * (l.u * l.u + l.v * l.v);
  double Hr = std::pow(M.sound_speed(r), 2.0) / g1 + 0.5 * (r.u * r.u + r.v * r.v);
  double p = 0.5 * (l.p + r.p);
  double rho = 0.5 * (rhol + rhor);
  double H = 0.5 * (Hl + Hr);
  V2 vel{l.u * 0.5 + r.u * 0.5, l.v * 0.5 + r.v * 0.5};
  double vn = vel.x * un.x + vel.y * un.y;
  Flx F;
  F.m = rho * vn;
  F.mx = un.x * p + vel.x * F.m;
  F.my = un.y * p + vel.y * F.m;
  F.e = rho * H * vn;
  return F;
}
// src/roe_ec.cc:10-44
Flx roe_ec_flux(const Mat& M, const Prim& l, const Prim& r, V2 n) {
  double area = std::sqrt(n.x * n.x + n.y * n.y);
  double ia = 1.0 / area;
  V2 un{n.x * ia, n.y * ia};
  double rhol = M.density(l), rhor = M.density(r);
  double pl = l.p, pr = r.p;
  double bl = std::sqrt(rhol / pl);
  double br = std::sqrt(rhor / pr);
  double z1b = 0.5 * (bl + br);
  V2 zvb{l.u * (0.5 * bl) + r.u * (0.5 * br), l.v * (0.5 * bl) + r.v * (0.5 * br)};
  double z5b = 0.5 * (bl * pl + br * pr);
  double zvbn = zvb.x * un.x + zvb.y * un.y;
  double z1l = logavg(bl, br);
  double z5l = logavg(bl * pl, br * pr);
  double g = M.p.gamma;
  Flx F;
  F.m = z5l * zvbn;
  double c1 = z5b / z1b, c2 = F.m / z1b;
  F.mx = un.x * c1 + zvb.x * c2;
  F.my = un.y * c1 + zvb.y * c2;
  F.e = 0.5 * ((g + 1.0) / (g - 1.0) * F.m / z1l + (F.mx * zvb.x + F.my * zvb.y)) / z1b;
  return F;
}
// src/simple_avg.cc:10-46
Flx avg_flux(const Mat& M, const Prim& l, const Prim& r, V2 n) {
  double area = std::sqrt(n.x * n.x + n.y * n.y);
  double ia = 1.0 / area;
  V2 un{n.x * ia, n.y * ia};
  double g = M.p.gamma;
  double rhol = M.density(l), rhor = M.density(r);
  double El = 0.5 * rhol * (l.u * l.u + l.v * l.v) + l.p / (g - 1);
  double Er = 0.5 * rhor * (r.u * r.u + r.v * r.v) + r.p / (g - 1);
  double vnl = l.u * un.x + l.v * un.y;
  Flx fl{rhol * vnl, 0, 0, 0};
  fl.mx = un.x * l.p + l.u * fl.m;
  fl.my = un.y * l.p + l.v * fl.m;
  fl.e = (El + l.p) * vnl;
  double vnr = r.u * un.x + r.v * un.y;
  Flx fr{rhor * vnr, 0, 0, 0};
  fr.mx = un.x * r.p + r.u * fr.m;
  fr.my = un.y * r.p + r.v * fr.m;
  fr.e = (Er + r.p) * vnr;
  Flx F{0.0, 0.0, 0.0, 0.0};
  F.m += fl.m * 0.5; F.mx += fl.mx * 0.5; F.my += fl.my * 0.5; F.e += fl.e * 0.5;
  F.m += fr.m * 0.5; F.mx += fr.mx * 0.5; F.my += fr.my * 0.5; F.e += fr.e * 0.5;
  return F;
}

// src/kepes_tecno_roe.cc:11-173 and src/kepes_tecno_rusanov.cc:11-164.
// dEl/dEr: gradients of (f, vx, vy, l) as (x,y) pairs.
Flx tecno_flux(const Mat& M, V2 cl, V2 cr, const Prim& l, const Prim& r,
               const V2* dEl, const V2* dEr, V2 n, double ssw, bool roe) {
  const double g = M.p.gamma, R = M.p.gas_const;
  double area = std::sqrt(n.x * n.x + n.y * n.y);
  double ia = 1.0 / area;
  V2 un{n.x * ia, n.y * ia};
  double rhol = M.density(l), rhor = M.density(r);
  double rho = logavg(rhol, rhor);
  double betal = 0.5 / (R * l.T);
  double betar = 0.5 / (R * r.T);
  double beta = logavg(betal, betar);
  double a = std::sqrt(0.5 * g / beta);
  double p = 0.5 * (rhol + rhor) / (betal + betar);
  V2 vel{l.u * 0.5 + r.u * 0.5, l.v * 0.5 + r.v * 0.5};
  double vn = vel.x * un.x + vel.y * un.y;

  Flx F = kepec_flux(M, l, r, n);
  if (!M.p.nd_active) return F;

  double vsq = vel.x * vel.x + vel.y * vel.y;
  double H = a * a / (g - 1.0) + 0.5 * vsq;
  double v1 = vel.x * un.y - vel.y * un.x;
  // rows/cols: 0 mass, 1 x-mom, 2 y-mom, 3 energy  (reference index 0,1,2,4)
  double Rm[4][4] = {
      {1, 1, 0, 1},
      {vel.x - a * un.x, vel.x, un.y, vel.x + a * un.x},
      {vel.y - a * un.y, vel.y, -un.x, vel.y + a * un.y},
      {H - a * vn, 0.5 * vsq, v1, H + a * vn}};
  double const1 = std::sqrt(rho / g);
  double psqrt = std::sqrt(p);
  double S[4] = {std::sqrt(0.5) * const1, std::sqrt(g - 1.0) * const1, psqrt,
                 std::sqrt(0.5) * const1};
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) Rm[j][i] *= S[i];

  double Lam[4];
  if (roe) {
    double vnl = l.u * un.x + l.v * un.y;
    double vnr = r.u * un.x + r.v * un.y;
    double al = M.sound_speed(l), ar = M.sound_speed(r);
    double L0l = vnl - al, L0r = vnr - ar, L4l = vnl + al, L4r = vnr + ar;
    double l2 = std::fabs(vn), l3 = std::fabs(vn);
    if (ssw > 0) l3 = std::max(l3, a);
    Lam[0] = (1 + M.p.ev_beta) * std::fabs(vn - a) + M.p.ev_alpha * std::fabs(L0l - L0r);
    Lam[1] = l2;
    Lam[2] = l3;
    Lam[3] = (1 + M.p.ev_beta) * std::fabs(vn + a) + M.p.ev_alpha * std::fabs(L4l - L4r);
  } else {
    double lr = std::fabs(vn) + std::fabs(a);
    Lam[0] = Lam[1] = Lam[2] = Lam[3] = lr;
  }
  double sl = M.entropy(l), sr = M.entropy(r);
  double bl = 0.5 / (R * l.T), br = 0.5 / (R * r.T);
  double Vl[4] = {(g - sl) / (g - 1.0) - bl * (l.u * l.u + l.v * l.v), 2 * l.u * bl,
                  2 * l.v * bl, -2 * bl};
  double Vr[4] = {(g - sr) / (g - 1.0) - br * (r.u * r.u + r.v * r.v), 2 * r.u * br,
                  2 * r.v * br, -2 * br};
  double rx = cr.x - cl.x, ry = cr.y - cl.y;
  double h = std::sqrt(std::pow(rx, 2) + std::pow(ry, 2) + std::pow(0.0, 2));
  double gl[4], gr[4];
  for (int i = 0; i < 4; ++i) {
    gl[i] = dEl[i].x * rx + dEl[i].y * ry;
    gr[i] = dEr[i].x * rx + dEr[i].y * ry;
  }
  double Zl[4] = {0, 0, 0, 0}, Zr[4] = {0, 0, 0, 0};
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) {
      Zl[i] += Rm[j][i] * Vl[j];
      Zr[i] += Rm[j][i] * Vr[j];
    }
  double sgl[4] = {0, 0, 0, 0}, sgr[4] = {0, 0, 0, 0};
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) {
      sgl[i] += Rm[j][i] * gl[j];
      sgr[i] += Rm[j][i] * gr[j];
    }
  double Zij[4], Zji[4];
  reconstruct(M.p, h, Zl, Zr, sgl, sgr, Zij, Zji);
  double dZ[4], D[4];
  for (int i = 0; i < 4; ++i) { dZ[i] = Zji[i] - Zij[i]; D[i] = 0; }
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) D[i] += Rm[i][j] * Lam[j] * dZ[j];
  F.m -= 0.5 * D[0];
  F.mx -= 0.5 * D[1];
  F.my -= 0.5 * D[2];
  F.e -= 0.5 * D[3];
  return F;
}

// src/material.cc:27-74
Flx num_flux(const Mat& M, V2 cl, V2 cr, const Prim& l, const Prim& r, const V2* dEl,
             const V2* dEr, V2 n, double ssw) {
  Flx F;
  switch (M.p.flux) {
    case F_ROE: F = tecno_flux(M, cl, cr, l, r, dEl, dEr, n, ssw, true); break;
    case F_RUS: F = tecno_flux(M, cl, cr, l, r, dEl, dEr, n, ssw, false); break;
    case F_KEPEC: F = kepec_flux(M, l, r, n); break;
    case F_KEP: F = kep_flux(M, l, r, n); break;
    case F_ROEEC: F = roe_ec_flux(M, l, r, n); break;
    default: F = avg_flux(M, l, r, n); break;
  }
  double area = std::sqrt(n.x * n.x + n.y * n.y);
  F.m *= area; F.mx *= area; F.my *= area; F.e *= area;
  return F;
}

// src/material.cc:110-127
Flx inviscid_bflux(const Mat& M, const Prim& w, const Prim& wb, V2 n) {
  double vnb = wb.u * n.x + wb.v * n.y;
  double vn = w.u * n.x + w.v * n.y;
  double rho = M.density(w);
  Flx F;
  F.m = rho * vnb;
  F.mx = n.x * wb.p + w.u * F.m;
  F.my = n.y * wb.p + w.v * F.m;
  F.e = (0.5 * rho * (w.u * w.u + w.v * w.v) + M.p.gamma * w.p / (M.p.gamma - 1.0)) * vnb +
        (wb.p - w.p) * vn;
  return F;
}

// src/material.cc:158-230 (z terms are exact zeros); with `wall` set, the
// boundary flux of src/material.cc:236-302 through the single normal nrm[0]:
// the same terms, minus Cor = -(0 + c2 e4^2 (dE_l . n)) in energy when the wall
// is adiabatic
void viscous_flux(const Mat& M, const Ent& s, const V2* dE, const V2* nrm, Flx* out,
                  bool wall = false, bool adiabatic = false) {
  const double R = M.p.gas_const;
  double mu = M.viscosity(-1 / (R * s.l));
  double k = mu * M.Cp / M.p.prandtl;
  if (!M.p.visc_active) mu = 0.0;
  if (!M.p.heat_active) k = 0.0;
  double e1 = s.vx / s.l, e2 = s.vy / s.l, e4 = 1.0 / s.l;
  double c1 = 4.0 * mu / 3.0, c2 = k / R, c3 = 2 * mu / 3.0;
  // dE index: 0 f, 1 vx, 2 vy, 3 l  (reference 0,1,2,4)
  double K11[4], K12[4], K21[4], K22[4];
  K11[1] = c1 * (-dE[1].x + e1 * dE[3].x);
  K11[2] = mu * (-dE[2].x + e2 * dE[3].x);
  K11[3] = c1 * e1 * dE[1].x + mu * (e2 * dE[2].x) +
           (-c1 * e1 * e1 - mu * (e2 * e2) + c2 * e4) * dE[3].x;
  K12[1] = c3 * (dE[2].y - e2 * dE[3].y);
  K12[2] = mu * (-dE[1].y + e1 * dE[3].y);
  K12[3] = mu * e2 * dE[1].y - c3 * e1 * dE[2].y - (mu - c3) * e1 * e2 * dE[3].y;
  K21[1] = mu * (-dE[2].x + e2 * dE[3].x);
  K21[2] = c3 * (dE[1].x - e1 * dE[3].x);
  K21[3] = -c3 * e2 * dE[1].x + mu * e1 * dE[2].x - (mu - c3) * e1 * e2 * dE[3].x;
  K22[1] = mu * (-dE[1].y + e1 * dE[3].y);
  K22[2] = c1 * (-dE[2].y + e2 * dE[3].y);
  K22[3] = mu * (e1 * dE[1].y) + c1 * e2 * dE[2].y +
           (-mu * (e1 * e1) - c1 * e2 * e2 + c2 * e4) * dE[3].y;
  for (int j = 0; j < (wall ? 1 : 3); ++j) {
    V2 n = nrm[j];
    out[j].m = 0.0;
    out[j].mx = -((K11[1] + K12[1]) * n.x + (K21[1] + K22[1]) * n.y) * e4;
    out[j].my = -((K11[2] + K12[2]) * n.x + (K21[2] + K22[2]) * n.y) * e4;
    out[j].e = -((K11[3] + K12[3]) * n.x + (K21[3] + K22[3]) * n.y) * e4;
    if (wall) {
      double Cor = adiabatic ? -1.0 * (0.0 + c2 * e4 * e4 * (dE[3].x * n.x + dE[3].y * n.y)) : 0.0;
      out[j].e = out[j].e - Cor;
    }
  }
}

struct Face {
  int v[2];
  int type;        // bc tag (>=0) or -1 interior
  bool periodic;
  int lcell, rcell;
  V2 centroid, normal;
  double measure;
};
struct Cell {
  int v[3];
  int face[3];
  V2 centroid;
  double area;
  V2 normal[3];
  std::vector<int> nbr_bface;
};
struct BC { int type; Prim state; bool adiabatic; };

}  // namespace

struct orc_solver {
  Mat M;
  int nv, nc;
  std::vector<V2> x;
  std::vector<Cell> cell;
  std::vector<Face> face, bface;
  std::vector<double> mcarea, dcarea;
  std::vector<std::vector<int>> per;   // periodic sets P_ij
  std::vector<double> pweight;         // 1/|P| for periodic vertices else 1
  std::map<int, BC> bc;
  // state
  std::vector<Prim> W;
  std::vector<Ent> E;
  std::vector<Flx> Uold, Rs, R2;
  std::vector<double> dt, ssw;
  std::vector<V2> dE;       // nv*4
  // dE is not yet that of the current set-up: the reference's loop computes
  // the gradients of the initial state before its first output (src/fv.cc:2238)
  bool dE_stale = true;
  std::vector<V2> dEc;      // nc*4
  double dt_global = 0, elapsed = 0, res_total = 1e20, res0 = 0;
  long iter = 0;
  unsigned counter = 0;
  bool reached = false;
  std::vector<double> time_instance;
  // boundary states per (bface, end point) of this run's sample, set by
  // orc_set_bc_state (the reference evaluates each boundary expression at the
  // end point's coordinates with the sample's random numbers, src/bc.h:411-562);
  // empty: every bface takes its tag's constant state
  std::vector<Prim> bst;
  // boundary values of time (orc_set_bc_time_fn): callback, input bface end
  // points and the values in the input layout
  orc_bc_time_fn bt_fn = nullptr;
  void* bt_user = nullptr;
  std::vector<int> bt_bv;
  std::vector<double> bt_in;
  std::vector<int> face_bface;  // the bface of each boundary edge's face, else -1
  Prim bstate_at(int bi, int vertex) const {
    if (bst.empty()) return bc.at(bface[bi].type).state;
    return bst[2 * (size_t)bi + (bface[bi].v[1] == vertex ? 1 : 0)];
  }
  const BC& bcof(int tag) const { return bc.at(tag); }
};

namespace {

// src/grid_preproc.cc:396-454 (add_face) and :461-590 (make_faces)
bool build_geometry(orc_solver* s, const orc_mesh* m) {
  const orc_params& P = s->M.p;
  s->nv = m->n_vertex;
  s->nc = m->n_cell;
  s->x.resize(s->nv);
  for (int i = 0; i < s->nv; ++i) s->x[i] = V2{m->x[i], m->y[i]};
  for (int i = 0; i < m->n_bc; ++i) {
    BC b;
    b.type = m->bc_type[i];
    b.state = Prim{m->bc_state[4 * i], m->bc_state[4 * i + 1], m->bc_state[4 * i + 2],
                   m->bc_state[4 * i + 3]};
    // src/bc.h:233-276: slip walls are adiabatic; a noslip wall is unless it
    // names a temperature
    b.adiabatic = b.type == BC_SLIP || (b.type == BC_NOSLIP && m->bc_adiabatic && m->bc_adiabatic[i]);
    s->bc[m->bc_tag[i]] = b;
  }
  s->cell.resize(s->nc);
  for (int c = 0; c < s->nc; ++c)
    for (int j = 0; j < 3; ++j) s->cell[c].v[j] = m->cell_v[3 * c + j];
  std::vector<Face>& face = s->face;
  face.resize(m->n_bface);
  std::vector<std::vector<int>> node_face(s->nv);
  for (int i = 0; i < m->n_bface; ++i) {
    Face& f = face[i];
    f.type = m->bface_tag[i];
    f.v[0] = m->bface_v[2 * i];
    f.v[1] = m->bface_v[2 * i + 1];
    if (!s->bc.count(f.type)) { g_err = "boundary tag without bc"; return false; }
    f.periodic = s->bcof(f.type).type == BC_PERIODIC;
    f.lcell = f.rcell = -1;
    node_face[f.v[0]].push_back(i);
    node_face[f.v[1]].push_back(i);
  }
  auto add_face = [&](int a, int b, int c) {
    bool found = false;
    for (size_t n = 0; n < node_face[a].size() && !found; ++n) {
      Face& f = face[node_face[a][n]];
      bool same = (f.v[0] == a && f.v[1] == b) || (f.v[0] == b && f.v[1] == a);
      if (f.lcell == -1) {
        if (same) { f.lcell = c; found = true; }
      } else if (f.rcell == -1) {
        if (same) { f.rcell = c; found = true; }
      }
    }
    if (!found) {
      Face f;
      f.type = -1; f.v[0] = a; f.v[1] = b; f.lcell = c; f.rcell = -1; f.periodic = false;
      int id = (int)face.size();
      face.push_back(f);
      node_face[a].push_back(id);
      node_face[b].push_back(id);
    }
  };
  for (int c = 0; c < s->nc; ++c) {
    const int* v = s->cell[c].v;
    add_face(v[0], v[1], c);
    add_face(v[1], v[2], c);
    add_face(v[2], v[0], c);
  }
  // drop interior-typed faces with a single cell (ghost-domain faces)
  std::vector<Face> kept;
  kept.reserve(face.size());
  for (auto& f : face)
    if (!(f.type == -1 && f.rcell == -1)) kept.push_back(f);
  face.swap(kept);
  for (auto& f : face)
    if (f.lcell == -1) { g_err = "face without cell"; return false; }
  for (auto& f : face)
    if (f.type != -1) s->bface.push_back(f);
  // src/grid_preproc.cc:600-623 find_cell_faces
  for (auto& c : s->cell) c.face[0] = c.face[1] = c.face[2] = -1;
  for (int i = 0; i < (int)face.size(); ++i) {
    Cell& L = s->cell[face[i].lcell];
    int j = 0; while (L.face[j] != -1) ++j; L.face[j] = i;
    if (face[i].rcell != -1) {
      Cell& Rr = s->cell[face[i].rcell];
      j = 0; while (Rr.face[j] != -1) ++j; Rr.face[j] = i;
    }
  }
  // src/grid_preproc.cc:165-184 face centroids
  for (auto& f : face) {
    f.centroid.x = s->x[f.v[0]].x * 0.5 + s->x[f.v[1]].x * 0.5;
    f.centroid.y = s->x[f.v[0]].y * 0.5 + s->x[f.v[1]].y * 0.5;
  }
  for (auto& f : s->bface) {
    f.centroid.x = s->x[f.v[0]].x * 0.5 + s->x[f.v[1]].x * 0.5;
    f.centroid.y = s->x[f.v[0]].y * 0.5 + s->x[f.v[1]].y * 0.5;
  }
  // src/grid_preproc.cc:18-150 cell centroids
  const double third = 1.0 / 3.0;
  for (auto& c : s->cell) {
    const V2 &a = s->x[c.v[0]], &b = s->x[c.v[1]], &d = s->x[c.v[2]];
    if (P.cell_type == 0) {
      c.centroid.x = a.x * third + b.x * third + d.x * third;
      c.centroid.y = a.y * third + b.y * third + d.y * third;
    } else {
      double x1 = a.x, y1 = a.y, x2 = b.x, y2 = b.y, x3 = d.x, y3 = d.y;
      double l1 = std::pow(x2 - x3, 2) + std::pow(y2 - y3, 2);
      double l2 = std::pow(x3 - x1, 2) + std::pow(y3 - y1, 2);
      double l3 = std::pow(x1 - x2, 2) + std::pow(y1 - y2, 2);
      double be1 = std::max(0.0, l2 + l3 - l1);
      double be2 = std::max(0.0, l3 + l1 - l2);
      double be3 = std::max(0.0, l1 + l2 - l3);
      if (be1 < be2 / 2.0 && be1 < be3 / 2.0) be1 = 0.0;
      if (be2 < be3 / 2.0 && be2 < be1 / 2.0) be2 = 0.0;
      if (be3 < be1 / 2.0 && be3 < be2 / 2.0) be3 = 0.0;
      double det = (x2 - x1) * (y3 - y1) - (x3 - x1) * (y2 - y1);
      double b1 = 0.5 * ((x2 - x1) * (x2 + x1) + (y2 - y1) * (y2 + y1));
      double b2 = 0.5 * ((x3 - x1) * (x3 + x1) + (y3 - y1) * (y3 + y1));
      double xc = ((y3 - y1) * b1 - (y2 - y1) * b2) / det;
      double yc = (-(x3 - x1) * b1 + (x2 - x1) * b2) / det;
      if (be1 == 0.0) { c.centroid.x = 0.5 * (x2 + x3); c.centroid.y = 0.5 * (y2 + y3); }
      else if (be2 == 0.0) { c.centroid.x = 0.5 * (x3 + x1); c.centroid.y = 0.5 * (y3 + y1); }
      else if (be3 == 0.0) { c.centroid.x = 0.5 * (x1 + x2); c.centroid.y = 0.5 * (y1 + y2); }
      else { c.centroid.x = xc; c.centroid.y = yc; }
    }
  }
  if (P.cell_type == 1) {
    for (auto& bf : s->bface) {
      Cell& c = s->cell[bf.lcell];
      double dx = bf.centroid.x * 1.0 + c.centroid.x * -1.0;
      double dy = bf.centroid.y * 1.0 + c.centroid.y * -1.0;
      if (std::sqrt(dx * dx + dy * dy) < 1.0e-14) {
        const V2 &a = s->x[c.v[0]], &b = s->x[c.v[1]], &d = s->x[c.v[2]];
        c.centroid.x = a.x * third + b.x * third + d.x * third;
        c.centroid.y = a.y * third + b.y * third + d.y * third;
      }
    }
  }
  // src/grid_preproc.cc:175-254 areas
  s->mcarea.assign(s->nv, 0.0);
  s->dcarea.assign(s->nv, 0.0);
  for (auto& c : s->cell) {
    const V2 &a = s->x[c.v[0]], &b = s->x[c.v[1]], &d = s->x[c.v[2]];
    double l1x = b.x * 1.0 + a.x * -1.0, l1y = b.y * 1.0 + a.y * -1.0;
    double l2x = d.x * 1.0 + a.x * -1.0, l2y = d.y * 1.0 + a.y * -1.0;
    double z = l1x * l2y - l1y * l2x;
    c.area = 0.5 * z;
    if (!(c.area > 0.0)) { g_err = "non-positive cell area"; return false; }
    for (int j = 0; j < 3; ++j) s->mcarea[c.v[j]] += c.area / 3.0;
  }
  if (P.cell_type == 0) {
    s->dcarea = s->mcarea;
  } else {
    for (auto& c : s->cell) {
      for (int j = 0; j < 3; ++j) {
        int n0 = c.v[j];
        int n1 = (j == 0) ? c.v[2] : c.v[j - 1];
        int n2 = (j == 2) ? c.v[0] : c.v[j + 1];
        V2 pt[5];
        pt[0] = s->x[n0];
        pt[1] = V2{s->x[n0].x * 0.5 + s->x[n2].x * 0.5, s->x[n0].y * 0.5 + s->x[n2].y * 0.5};
        pt[2] = c.centroid;
        pt[3] = V2{s->x[n0].x * 0.5 + s->x[n1].x * 0.5, s->x[n0].y * 0.5 + s->x[n1].y * 0.5};
        pt[4] = pt[0];
        double ar = 0;
        for (int k = 0; k < 4; ++k) ar += pt[k].x * pt[k + 1].y - pt[k + 1].x * pt[k].y;
        if (ar < 0.0) { g_err = "dual cell area is non-positive"; return false; }
        ar *= 0.5;
        s->dcarea[n0] += ar;
      }
    }
  }
  // periodic sets, src/reduce_dual_area.cc:18-133
  s->per.resize(m->n_periodic);
  for (int i = 0; i < m->n_periodic; ++i)
    for (int k = m->periodic_ptr[i]; k < m->periodic_ptr[i + 1]; ++k)
      s->per[i].push_back(m->periodic_v[k]);
  for (auto& P_ : s->per) {
    double mc = 0, dc = 0;
    for (int v : P_) { mc += s->mcarea[v] * 1.0; dc += s->dcarea[v] * 1.0; }
    s->mcarea[P_[0]] = mc;
    s->dcarea[P_[0]] = dc;
  }
  for (auto& P_ : s->per)
    for (size_t j = 1; j < P_.size(); ++j) {
      s->mcarea[P_[j]] = s->mcarea[P_[0]];
      s->dcarea[P_[j]] = s->dcarea[P_[0]];
    }
  // src/grid_preproc.cc:259-365 normals
  for (auto& f : face) {
    int v0 = f.v[0], v1 = f.v[1];
    double drx = s->x[v1].x * 1.0 + s->x[v0].x * -1.0;
    double dry = s->x[v1].y * 1.0 + s->x[v0].y * -1.0;
    double nx = +dry, ny = -drx;
    const Cell& L = s->cell[f.lcell];
    double dcx = f.centroid.x * 1.0 + L.centroid.x * -1.0;
    double dcy = f.centroid.y * 1.0 + L.centroid.y * -1.0;
    double d = dcx * nx + dcy * ny;
    if (d < 0.0) { f.v[0] = v1; f.v[1] = v0; }
  }
  for (auto& f : face) {
    const Cell& L = s->cell[f.lcell];
    double drx = f.centroid.x * 1.0 + L.centroid.x * -1.0;
    double dry = f.centroid.y * 1.0 + L.centroid.y * -1.0;
    if (f.type == -1) {
      const Cell& Rr = s->cell[f.rcell];
      drx += Rr.centroid.x * 1.0 + f.centroid.x * -1.0;
      dry += Rr.centroid.y * 1.0 + f.centroid.y * -1.0;
    }
    f.normal.x = -dry;
    f.normal.y = drx;
    f.measure = std::sqrt(f.normal.x * f.normal.x + f.normal.y * f.normal.y);
  }
  for (auto& f : s->bface) {
    int v0 = f.v[0], v1 = f.v[1];
    double drx = s->x[v1].x * 1.0 + s->x[v0].x * -1.0;
    double dry = s->x[v1].y * 1.0 + s->x[v0].y * -1.0;
    f.normal.x = dry;
    f.normal.y = -drx;
    f.measure = std::sqrt(f.normal.x * f.normal.x + f.normal.y * f.normal.y);
    const Cell& L = s->cell[f.lcell];
    double dcx = f.centroid.x * 1.0 + L.centroid.x * -1.0;
    double dcy = f.centroid.y * 1.0 + L.centroid.y * -1.0;
    double d = dcx * f.normal.x + dcy * f.normal.y;
    if (d < 0.0) {
      f.v[0] = v1; f.v[1] = v0;
      f.normal.x *= -1.0; f.normal.y *= -1.0;
    }
  }
  for (auto& c : s->cell) {
    const V2 &p0 = s->x[c.v[0]], &p1 = s->x[c.v[1]], &p2 = s->x[c.v[2]];
    V2 d01{(p0.y - p1.y), -(p0.x - p1.x)};
    V2 d12{(p1.y - p2.y), -(p1.x - p2.x)};
    V2 d20{(p2.y - p0.y), -(p2.x - p0.x)};
    c.normal[0] = d12; c.normal[1] = d20; c.normal[2] = d01;
  }
  // src/grid_preproc.cc:640-700 remove_empty_faces (voronoi only)
  if (P.cell_type == 1) {
    std::vector<Face> tmp;
    for (auto& f : face) if (f.measure > 0.0) tmp.push_back(f);
    face.swap(tmp);
  }
  // src/grid_preproc.cc:811-831 mark_cell_bface.  The slot ids in cell.face
  // were assigned before remove_empty_faces, and the reference reads them
  // against the shortened list; that is reproduced here (median duals never
  // remove a face, so there the ids are exact).
  for (auto& c : s->cell) {
    c.nbr_bface.clear();
    for (int j = 0; j < 3; ++j) {
      int f = c.face[j];
      if (f < 0 || f >= (int)face.size()) continue;
      if (face[f].type != -1 && s->bcof(face[f].type).type != BC_PERIODIC)
        c.nbr_bface.push_back(j);
    }
  }
  {
    std::map<std::pair<int, int>, int> pair_bface;
    for (size_t bi = 0; bi < s->bface.size(); ++bi) {
      const Face& b = s->bface[bi];
      pair_bface[{std::min(b.v[0], b.v[1]), std::max(b.v[0], b.v[1])}] = (int)bi;
    }
    s->face_bface.assign(face.size(), -1);
    for (size_t i = 0; i < face.size(); ++i) {
      if (face[i].type == -1) continue;
      auto it = pair_bface.find({std::min(face[i].v[0], face[i].v[1]), std::max(face[i].v[0], face[i].v[1])});
      if (it != pair_bface.end()) s->face_bface[i] = it->second;
    }
  }
  s->pweight.assign(s->nv, 1.0);
  for (auto& P_ : s->per) {
    double w = 1.0 / (double(P_.size()) * 1.0);
    for (int v : P_) s->pweight[v] = w;
  }
  return true;
}

}  // namespace

namespace {

// Boundary state for the two-state apply, src/bc.h:569-610; bs = the
// boundary expressions' values at the state's vertex
void bc_apply2(const orc_solver* s, const BC& b, const Prim& bs, const Face& bf, Prim* st) {
  switch (b.type) {
    case BC_SLIP: {
      double im = 1.0 / bf.measure;
      V2 un{bf.normal.x * im, bf.normal.y * im};
      double fact = st[0].u * un.x + st[0].v * un.y;
      st[0].u += un.x * -fact;
      st[0].v += un.y * -fact;
      st[1] = st[0];
      break;
    }
    case BC_NOSLIP:  // src/bc.h:411-431
      st[0].u = bs.u; st[0].v = bs.v;
      st[1].u = st[0].u; st[1].v = st[0].v;
      st[1].p = st[0].p;
      if (b.adiabatic) {
        st[1].T = st[0].T;
      } else {
        st[0].T = bs.T; st[1].T = bs.T;
      }
      break;
    case BC_PRESSURE: st[1].p = bs.p; break;
    case BC_INLET: st[0] = bs; st[1] = st[0]; break;
    case BC_OUTLET: st[1] = st[0]; break;
    case BC_FARFIELD: st[1] = bs; break;
    default: break;
  }
  (void)s;
}
// One-state apply, src/bc.h:615-654
void bc_apply1(const BC& b, const Prim& bs, const Face& bf, Prim& st) {
  switch (b.type) {
    case BC_SLIP: {
      double im = 1.0 / bf.measure;
      V2 un{bf.normal.x * im, bf.normal.y * im};
      double fact = st.u * un.x + st.v * un.y;
      st.u += un.x * -fact;
      st.v += un.y * -fact;
      break;
    }
    case BC_NOSLIP:  // src/bc.h:438-451
      st.u = bs.u; st.v = bs.v;
      if (!b.adiabatic) st.T = bs.T;
      break;
    case BC_PRESSURE: st.p = bs.p; break;
    case BC_INLET: st = bs; break;
    default: break;
  }
}

inline void addF(Flx& a, const Flx& f, double c) {
  a.m += f.m * c; a.mx += f.mx * c; a.my += f.my * c; a.e += f.e * c;
}

// src/fv.cc:699-1059 (NPART = 1: every cell is in cell_mpi_int, every
// boundary face in d_bnd_unshared_face)
void compute_gradients(orc_solver* s) {
  const Mat& M = s->M;
  int nv = s->nv, nc = s->nc;
  s->dE.assign((size_t)nv * 4, V2{0.0, 0.0});
  s->dE_stale = false;
  s->dEc.assign((size_t)nc * 4, V2{0.0, 0.0});
  for (int i = 0; i < nc; ++i) {
    const Cell& c = s->cell[i];
    const Ent &e0 = s->E[c.v[0]], &e1 = s->E[c.v[1]], &e2 = s->E[c.v[2]];
    const double a0[4] = {e0.f, e0.vx, e0.vy, e0.l};
    const double a1[4] = {e1.f, e1.vx, e1.vy, e1.l};
    const double a2[4] = {e2.f, e2.vx, e2.vy, e2.l};
    for (int k = 0; k < 4; ++k) {
      V2& g = s->dEc[(size_t)i * 4 + k];
      g.x = c.normal[0].x * a0[k] + c.normal[1].x * a1[k] + c.normal[2].x * a2[k];
      g.y = c.normal[0].y * a0[k] + c.normal[1].y * a1[k] + c.normal[2].y * a2[k];
    }
  }
  for (size_t bi = 0; bi < s->bface.size(); ++bi) {
    const Face& bf = s->bface[bi];
    if (bf.periodic) continue;
    const BC& b = s->bcof(bf.type);
    int v0 = bf.v[0], v1 = bf.v[1];
    Prim st0 = s->W[v0], st1 = s->W[v1];
    bc_apply1(b, s->bstate_at((int)bi, v0), bf, st0);
    bc_apply1(b, s->bstate_at((int)bi, v1), bf, st1);
    Ent q0 = prim2ent(M, st0), q1 = prim2ent(M, st1);
    const Ent &E0 = s->E[v0], &E1 = s->E[v1];
    double d[4] = {(q0.f + q1.f - E0.f - E1.f) / 2.0, (q0.vx + q1.vx - E0.vx - E1.vx) / 2.0,
                   (q0.vy + q1.vy - E0.vy - E1.vy) / 2.0, (q0.l + q1.l - E0.l - E1.l) / 2.0};
    for (int k = 0; k < 4; ++k) {
      V2& g = s->dEc[(size_t)bf.lcell * 4 + k];
      g.x += bf.normal.x * d[k];
      g.y += bf.normal.y * d[k];
    }
  }
  for (int i = 0; i < nc; ++i) {
    const Cell& c = s->cell[i];
    for (int j = 0; j < 3; ++j)
      for (int k = 0; k < 4; ++k) {
        V2& g = s->dE[(size_t)c.v[j] * 4 + k];
        const V2& h = s->dEc[(size_t)i * 4 + k];
        g.x += h.x; g.y += h.y;
      }
  }
  // src/reduce_gradient_periodic.cc:14-135
  for (auto& P : s->per) {
    V2 sum[4] = {{0, 0}, {0, 0}, {0, 0}, {0, 0}};
    for (int v : P)
      for (int k = 0; k < 4; ++k) {
        sum[k].x += s->dE[(size_t)v * 4 + k].x * 1.0;
        sum[k].y += s->dE[(size_t)v * 4 + k].y * 1.0;
      }
    for (int k = 0; k < 4; ++k) s->dE[(size_t)P[0] * 4 + k] = sum[k];
  }
  for (auto& P : s->per)
    for (size_t j = 1; j < P.size(); ++j)
      for (int k = 0; k < 4; ++k) s->dE[(size_t)P[j] * 4 + k] = s->dE[(size_t)P[0] * 4 + k];
  for (int i = 0; i < nv; ++i) {
    double fact = 1.0 / (6.0 * s->mcarea[i]);
    for (int k = 0; k < 4; ++k) { s->dE[(size_t)i * 4 + k].x *= fact; s->dE[(size_t)i * 4 + k].y *= fact; }
  }
  for (int i = 0; i < nc; ++i) {
    double fact = 1.0 / (2.0 * s->cell[i].area);
    for (int k = 0; k < 4; ++k) { s->dEc[(size_t)i * 4 + k].x *= fact; s->dEc[(size_t)i * 4 + k].y *= fact; }
  }
}

// src/fv.cc:1066-1192
void compute_inviscid_residual(orc_solver* s) {
  const Mat& M = s->M;
  for (auto& f : s->face) {
    int cl = f.v[0], cr = f.v[1];
    double ssw = s->ssw[cl] + s->ssw[cr];
    Flx F = num_flux(M, s->x[cl], s->x[cr], s->W[cl], s->W[cr], &s->dE[(size_t)cl * 4],
                     &s->dE[(size_t)cr * 4], f.normal, ssw);
    addF(s->Rs[cl], F, 1.0 * 1.0);
    addF(s->Rs[cr], F, -1.0 * 1.0);
  }
  const V2 zero4[4] = {{0, 0}, {0, 0}, {0, 0}, {0, 0}};
  for (size_t bi = 0; bi < s->bface.size(); ++bi) {
    const Face& bf = s->bface[bi];
    if (bf.periodic) continue;
    const BC& b = s->bcof(bf.type);
    for (int side = 0; side < 2; ++side) {
      int c = bf.v[side];
      Prim st[2] = {s->W[c], s->W[c]};
      bc_apply2(s, b, s->bstate_at((int)bi, c), bf, st);
      Flx F;
      if (b.type == BC_FARFIELD)
        F = num_flux(M, s->x[c], s->x[c], st[0], st[1], zero4, zero4, bf.normal, 0.0);
      else
        F = inviscid_bflux(M, st[0], st[1], bf.normal);
      addF(s->Rs[c], F, 0.5 * 1.0);
    }
  }
}

// src/fv.cc:1197-1232 (cells without a non-periodic boundary face), then
// src/fv.cc:1236-1312 (cells with one, once per such boundary face)
bool compute_viscous_residual(orc_solver* s) {
  const Mat& M = s->M;
  const double fact = 1.0 / 3.0;
  for (int i = 0; i < s->nc; ++i) {
    const Cell& c = s->cell[i];
    if (!c.nbr_bface.empty()) continue;
    const Ent &e0 = s->E[c.v[0]], &e1 = s->E[c.v[1]], &e2 = s->E[c.v[2]];
    Ent a;
    a.f = e0.f * fact + e1.f * fact + e2.f * fact;
    a.vx = e0.vx * fact + e1.vx * fact + e2.vx * fact;
    a.vy = e0.vy * fact + e1.vy * fact + e2.vy * fact;
    a.l = e0.l * fact + e1.l * fact + e2.l * fact;
    Flx out[3];
    viscous_flux(M, a, &s->dEc[(size_t)i * 4], c.normal, out);
    for (int j = 0; j < 3; ++j) addF(s->Rs[c.v[j]], out[j], -(0.5 * 1.0));
  }
  for (const auto& bf : s->bface) {
    const Cell& c = s->cell[bf.lcell];
    if (c.nbr_bface.empty() || bf.periodic) continue;
    // mean entropy variables of the boundary states of the cell's boundary
    // edges, each edge's end points in face order; the states come from the
    // one-state apply with the edge's face (its half dual face normal)
    Ent a{0.0, 0.0, 0.0, 0.0};
    for (int ind : c.nbr_bface) {
      const Face& f = s->face[c.face[ind]];
      const BC& b = s->bcof(f.type);
      const int fb = s->face_bface[c.face[ind]];
      for (int k = 0; k < 2; ++k) {
        Prim st = s->W[f.v[k]];
        bc_apply1(b, s->bstate_at(fb, f.v[k]), f, st);
        Ent e = prim2ent(M, st);
        a.f += e.f; a.vx += e.vx; a.vy += e.vy; a.l += e.l;
      }
    }
    const double sc = 0.5 * 1.0 / c.nbr_bface.size();
    a.f *= sc; a.vx *= sc; a.vy *= sc; a.l *= sc;
    const V2* dEc = &s->dEc[(size_t)bf.lcell * 4];
    Flx out[3];
    viscous_flux(M, a, dEc, c.normal, out);
    for (int j = 0; j < 3; ++j) addF(s->Rs[c.v[j]], out[j], -(0.5 * 1.0));
    Flx fb;
    viscous_flux(M, a, dEc, &bf.normal, &fb, true, s->bcof(bf.type).adiabatic);
    addF(s->Rs[bf.v[0]], fb, 0.5 * 1.0);
    addF(s->Rs[bf.v[1]], fb, 0.5 * 1.0);
  }
  return true;
}

// src/fv.cc:1467-1494, src/reduce_residual_periodic.cc:14-132
bool compute_residual(orc_solver* s) {
  s->Rs.assign(s->nv, Flx{0, 0, 0, 0});
  compute_gradients(s);
  compute_inviscid_residual(s);
  if (s->M.p.model == 1 && !compute_viscous_residual(s)) return false;
  for (auto& P : s->per) {
    Flx sum{0, 0, 0, 0};
    for (int v : P) addF(sum, s->Rs[v], 1.0);
    s->Rs[P[0]] = sum;
  }
  for (auto& P : s->per)
    for (size_t j = 1; j < P.size(); ++j) s->Rs[P[j]] = s->Rs[P[0]];
  return true;
}

// src/compute_ssw.cc:7-39
void compute_ssw(orc_solver* s) {
  if (s->M.p.flux != F_ROE) return;
  for (auto& v : s->ssw) v = 0;
  for (auto& f : s->face) {
    int cl = f.v[0], cr = f.v[1];
    double ui = s->W[cl].u * f.normal.x + s->W[cl].v * f.normal.y;
    double uj = s->W[cr].u * f.normal.x + s->W[cr].v * f.normal.y;
    double ai = s->M.sound_speed(s->W[cl]), aj = s->M.sound_speed(s->W[cr]);
    double area = std::sqrt(f.normal.x * f.normal.x + f.normal.y * f.normal.y);
    double l1i = ui - ai * area, l1j = uj - aj * area;
    double l5i = ui + ai * area, l5j = uj + aj * area;
    if (l1i > 0 && l1j < 0) s->ssw[cl] = s->ssw[cr] = 1;
    if (l5i > 0 && l5j < 0) s->ssw[cl] = s->ssw[cr] = 1;
  }
}

// src/fv.cc:1499-1621
void compute_dt(orc_solver* s) {
  const Mat& M = s->M;
  const orc_params& P = M.p;
  if (P.time_step > 0.0) {
    s->dt_global = P.time_step;
    for (auto& d : s->dt) d = s->dt_global;
    return;
  }
  for (auto& d : s->dt) d = 0.0;
  for (auto& f : s->face) {
    double area = f.measure;
    int cl = f.v[0], cr = f.v[1];
    double vl = s->W[cl].u * f.normal.x + s->W[cl].v * f.normal.y;
    double c_l = M.sound_speed(s->W[cl]);
    s->dt[cl] += (std::fabs(vl) + c_l * area) * 1.0;
    double vr = s->W[cr].u * f.normal.x + s->W[cr].v * f.normal.y;
    double c_r = M.sound_speed(s->W[cr]);
    s->dt[cr] += (std::fabs(vr) + c_r * area) * 1.0;
  }
  for (auto& bf : s->bface) {
    if (bf.periodic) continue;
    double area = bf.measure;
    int cl = bf.v[0], cr = bf.v[1];
    double vl = s->W[cl].u * bf.normal.x + s->W[cl].v * bf.normal.y;
    double c_l = M.sound_speed(s->W[cl]);
    s->dt[cl] += 0.5 * (std::fabs(vl) + c_l * area);
    double vr = s->W[cr].u * bf.normal.x + s->W[cr].v * bf.normal.y;
    double c_r = M.sound_speed(s->W[cr]);
    s->dt[cr] += 0.5 * (std::fabs(vr) + c_r * area);
  }
  if (P.model == 1) {
    for (auto& f : s->face) {
      double area = f.measure;
      int cl = f.v[0], cr = f.v[1];
      double dx = s->x[cr].x * 1.0 + s->x[cl].x * -1.0;
      double dy = s->x[cr].y * 1.0 + s->x[cl].y * -1.0;
      double dl = std::sqrt(dx * dx + dy * dy);
      double mu_l = M.viscosity(s->W[cl].T), mu_r = M.viscosity(s->W[cr].T);
      double rho_l = M.density(s->W[cl]), rho_r = M.density(s->W[cr]);
      s->dt[cl] += (mu_l * area / dl / rho_l) * 1.0;
      s->dt[cr] += (mu_r * area / dl / rho_r) * 1.0;
    }
  }
  s->dt_global = 1.0e20;
  for (auto& Pv : s->per) {
    double sum = 0;
    for (int v : Pv) sum += s->dt[v] * 1.0;
    s->dt[Pv[0]] = sum;
  }
  for (auto& Pv : s->per)
    for (size_t j = 1; j < Pv.size(); ++j) s->dt[Pv[j]] = s->dt[Pv[0]];
  for (int i = 0; i < s->nv; ++i) {
    s->dt[i] = P.cfl * s->dcarea[i] / s->dt[i];
    s->dt_global = std::min(s->dt_global, s->dt[i]);
  }
  if (P.unsteady) {
    // the reference indexes past time_instance when elapsed_time falls an ulp
    // short of final_time (src/fv.cc:1610); the last stamp is used instead
    unsigned ci = std::min<unsigned>(s->counter, (unsigned)s->time_instance.size() - 1);
    if (s->elapsed + s->dt_global > s->time_instance[ci]) {
      s->dt_global = s->time_instance[ci] - s->elapsed;
      s->reached = true;
    }
    for (auto& d : s->dt) d = s->dt_global;
  }
}

// src/fv.cc:1801-1879 (returns the unnormalised total)
double residual_norm(orc_solver* s) {
  double n[4] = {0, 0, 0, 0};
  for (int i = 0; i < s->nv; ++i) {
    double area = s->dcarea[i], w = s->pweight[i];
    const Flx& r = s->Rs[i];
    n[0] += std::pow(r.m / area, 2) * w;
    n[1] += std::pow(r.mx / area, 2) * w;
    n[2] += std::pow(r.my / area, 2) * w;
    n[3] += std::pow(r.e / area, 2) * w;
  }
  for (int k = 0; k < 4; ++k) n[k] = std::sqrt(n[k]) / s->nv;
  double tot = std::pow(n[0], 2) + (n[1] * n[1] + n[2] * n[2]) + std::pow(n[3], 2);
  return std::sqrt(tot);
}

// src/fv.cc:1635-1707 (ssprk3 and rk4)
void update_solution(orc_solver* s, unsigned r) {
  const Mat& M = s->M;
  for (int i = 0; i < s->nv; ++i) {
    double factor = s->dt[i] / (1.0 * s->dcarea[i]);
    Flx U;
    if (M.p.time_scheme == 0 || M.p.time_scheme == 2) {
      // SSP-RK3, or Heun's method (src/fv.cc:1675-1686) as it is meant:
      // U = (1 - r/2) U_r + (r/2) U_old - (1 - r/2) dt/A R.  The reference
      // overwrites the first term (Vector::equ where sadd adds), which gives
      // NaN after one step; this is the two-stage SSP scheme its
      // coefficients describe.
      const bool heun = M.p.time_scheme == 2;
      U = prim2con(M, s->W[i]);
      double b = heun ? (1.0 - r * 0.5) : b_rk[r], a = heun ? r * 0.5 : a_rk[r];
      double c = -factor * b;
      U.m *= b; U.mx *= b; U.my *= b; U.e *= b;
      const Flx &Uo = s->Uold[i], &R = s->Rs[i];
      U.m += Uo.m * a + R.m * c;
      U.mx += Uo.mx * a + R.mx * c;
      U.my += Uo.my * a + R.my * c;
      U.e += Uo.e * a + R.e * c;
    } else {
      addF(s->R2[i], s->Rs[i], 1.0 / rk4_rks[r]);
      const Flx& Uo = s->Uold[i];
      const Flx& RR = (r < 3) ? s->Rs[i] : s->R2[i];
      double c = (r < 3) ? -factor * rk4_rks[r + 1] : -factor / 6.0;
      U.m = Uo.m * 1.0 + RR.m * c;
      U.mx = Uo.mx * 1.0 + RR.mx * c;
      U.my = Uo.my * 1.0 + RR.my * c;
      U.e = Uo.e * 1.0 + RR.e * c;
    }
    s->W[i] = con2prim(M, U);
    s->E[i] = prim2ent(M, s->W[i]);
  }
}

// bc_scheme strong, src/fv.cc:1805-1819: before the last stage's residual norm
// and update, the momentum residual of both end points of every noslip bface
// is set to zero.
void strong_zero_noslip(orc_solver* s) {
  for (const Face& f : s->bface) {
    if (s->bcof(f.type).type != BC_NOSLIP) continue;
    for (int e = 0; e < 2; ++e) { s->Rs[f.v[e]].mx = 0; s->Rs[f.v[e]].my = 0; }
  }
}

// bc_scheme strong, src/fv.cc:1694-1704: after each stage's update every bface,
// in bface order, imposes its boundary condition on both end points' primitive
// state; the entropy variables keep the values the update gave them.
void strong_apply(orc_solver* s) {
  for (size_t b = 0; b < s->bface.size(); ++b) {
    const Face& f = s->bface[b];
    const BC& bc = s->bcof(f.type);
    bc_apply1(bc, s->bstate_at((int)b, f.v[0]), f, s->W[f.v[0]]);
    bc_apply1(bc, s->bstate_at((int)b, f.v[1]), f, s->W[f.v[1]]);
  }
}

}  // namespace

extern "C" {

const char* orc_error(void) { return g_err.c_str(); }

orc_solver* orc_create(const orc_mesh* m, const orc_params* p) {
  orc_solver* s = new orc_solver();
  s->M.p = *p;
  s->M.Cp = p->gamma * p->gas_const / (p->gamma - 1.0);
  if (!build_geometry(s, m)) { delete s; return nullptr; }
  s->W.assign(s->nv, Prim{1, 0, 0, 1});
  s->E.resize(s->nv);
  s->Uold.resize(s->nv);
  s->Rs.assign(s->nv, Flx{0, 0, 0, 0});
  s->R2.assign(s->nv, Flx{0, 0, 0, 0});
  s->dt.assign(s->nv, 0.0);
  s->ssw.assign(s->nv, 0.0);
  // src/fv.cc:50-60 time stamps
  if (p->unsteady) {
    s->time_instance.resize(p->n_time_stamps + 1);
    double tdt = p->final_time / (p->n_time_stamps);
    for (size_t i = 0; i < s->time_instance.size(); ++i) s->time_instance[i] = i * tdt;
  } else {
    s->time_instance.assign(2, 0.0);
  }
  return s;
}
void orc_destroy(orc_solver* s) { delete s; }

int orc_n_face(const orc_solver* s) { return (int)s->face.size(); }
int orc_n_bface(const orc_solver* s) { return (int)s->bface.size(); }
void orc_faces(const orc_solver* s, int* fv, double* fn, int* ft) {
  for (size_t i = 0; i < s->face.size(); ++i) {
    fv[2 * i] = s->face[i].v[0]; fv[2 * i + 1] = s->face[i].v[1];
    fn[2 * i] = s->face[i].normal.x; fn[2 * i + 1] = s->face[i].normal.y;
    ft[i] = s->face[i].type;
  }
}
void orc_bfaces(const orc_solver* s, int* bv, double* bn, int* lc) {
  for (size_t i = 0; i < s->bface.size(); ++i) {
    bv[2 * i] = s->bface[i].v[0]; bv[2 * i + 1] = s->bface[i].v[1];
    bn[2 * i] = s->bface[i].normal.x; bn[2 * i + 1] = s->bface[i].normal.y;
    lc[i] = s->bface[i].lcell;
  }
}
void orc_areas(const orc_solver* s, double* dc, double* mc, double* ca) {
  for (int i = 0; i < s->nv; ++i) { dc[i] = s->dcarea[i]; mc[i] = s->mcarea[i]; }
  for (int i = 0; i < s->nc; ++i) ca[i] = s->cell[i].area;
}

// src/fv.cc:380-402: state set at t = 0, entropy variables from prim2ent
void orc_set_state(orc_solver* s, const double* w) {
  for (int i = 0; i < s->nv; ++i) {
    s->W[i] = Prim{w[4 * i], w[4 * i + 1], w[4 * i + 2], w[4 * i + 3]};
    s->E[i] = prim2ent(s->M, s->W[i]);
  }
  s->elapsed = 0; s->iter = 0; s->counter = 0; s->res_total = 1e20; s->res0 = 0;
  s->reached = false;
  s->dE_stale = true;
}
void orc_get_state(const orc_solver* s, double* w) {
  for (int i = 0; i < s->nv; ++i) {
    w[4 * i] = s->W[i].T; w[4 * i + 1] = s->W[i].u; w[4 * i + 2] = s->W[i].v; w[4 * i + 3] = s->W[i].p;
  }
}
void orc_gradients(orc_solver* s, double* out) {
  compute_gradients(s);
  for (int i = 0; i < s->nv; ++i)
    for (int k = 0; k < 4; ++k) {
      out[8 * i + 2 * k] = s->dE[(size_t)i * 4 + k].x;
      out[8 * i + 2 * k + 1] = s->dE[(size_t)i * 4 + k].y;
    }
}
void orc_residual(orc_solver* s, double* R) {
  if (s->M.p.liou_fix) compute_ssw(s);
  compute_residual(s);
  for (int i = 0; i < s->nv; ++i) {
    R[4 * i] = s->Rs[i].m; R[4 * i + 1] = s->Rs[i].mx; R[4 * i + 2] = s->Rs[i].my; R[4 * i + 3] = s->Rs[i].e;
  }
}
double orc_compute_dt(orc_solver* s) {
  if (s->counter == 0) s->counter = 1;
  compute_dt(s);
  return s->dt_global;
}
void orc_local_dt(const orc_solver* s, double* dt) {
  for (int i = 0; i < s->nv; ++i) dt[i] = s->dt[i];
}

// src/fv.cc:2230-2395 loop body
long orc_run(orc_solver* s, long nsteps, double* log) {
  const orc_params& P = s->M.p;
  if (s->counter == 0) s->counter++;
  long done = 0;
  long term_iter = P.unsteady ? 0 : s->iter;
  while (done < nsteps && s->res_total > P.min_residue && term_iter < P.max_iter &&
         s->elapsed < (P.unsteady ? P.final_time : 1.0e20)) {
    for (int i = 0; i < s->nv; ++i) s->Uold[i] = prim2con(s->M, s->W[i]);
    compute_dt(s);
    if (P.liou_fix) compute_ssw(s);
    unsigned n_rks = P.time_scheme == 0 ? 3 : (P.time_scheme == 1 ? 4 : 2);
    if (P.time_scheme == 1)
      for (auto& r : s->R2) r = Flx{0, 0, 0, 0};
    for (unsigned r = 0; r < n_rks; ++r) {
      if (s->bt_fn && P.time_scheme == 2) {  // int_step_time of stage r (src/fv.cc:2296-2298)
        s->bt_fn(s->bt_user, s->elapsed + r * s->dt_global, s->bt_in.data());
        if (orc_set_bc_state(s, s->bt_bv.data(), s->bt_in.data())) return -1;
      }
      if (!compute_residual(s)) return -1;
      if (r == n_rks - 1) {
        if (P.strong_bc) strong_zero_noslip(s);
        double tot = residual_norm(s);
        if (s->iter == 0) {
          s->res0 = tot;
          if (s->res0 == 0.0) s->res0 = 1.0;
        }
        s->res_total = tot / s->res0;
      }
      update_solution(s, r);
      if (P.strong_bc) strong_apply(s);
    }
    ++s->iter;
    if (!P.unsteady) term_iter = s->iter;
    s->elapsed += s->dt_global;
    if (log) {
      log[4 * done] = (double)s->iter; log[4 * done + 1] = s->dt_global;
      log[4 * done + 2] = s->elapsed; log[4 * done + 3] = s->res_total;
    }
    if (s->reached && P.unsteady) { s->counter++; s->reached = false; }
    ++done;
  }
  return done;
}
double orc_time(const orc_solver* s) { return s->elapsed; }

// The loop clock a restart file carries (src/fv.cc:334-338, written by
// src/writer.cc:404-407): last iteration, stamp counter, elapsed time and the
// first step's residual norm that later ones are normalised by.  The loop
// then starts with residual_norm_total = 1e20 (src/fv.cc:2234).
void orc_get_clock(const orc_solver* s, long* iter, int* counter, double* t, double* res0) {
  *iter = s->iter; *counter = (int)s->counter; *t = s->elapsed; *res0 = s->res0;
}
void orc_set_clock(orc_solver* s, long iter, int counter, double t, double res0) {
  s->iter = iter; s->counter = (unsigned)counter; s->elapsed = t; s->res0 = res0;
  s->res_total = 1e20; s->reached = false;
}

namespace {
// Mean over a boundary face's two end points of the velocity and temperature
// gradients formed from the entropy-variable gradients dE and the end points'
// entropy variables: gradU = -dE[vx]/l + (vx/l^2) dE[l] (and likewise v),
// gradT = dE[l] / (R l^2) (src/force.cc:160-175, src/output_skin_friction.cc:
// 64-74, src/output_heatflux.cc:69-76).
struct WallGrad { V2 gu, gv, gT; };
WallGrad wall_gradients(const orc_solver* s, int v0, int v1, const Ent& e0, const Ent& e1) {
  const double f0 = 1.0 / e0.l, f1 = 1.0 / e1.l;
  const V2* g0 = &s->dE[(size_t)v0 * 4];
  const V2* g1 = &s->dE[(size_t)v1 * 4];
  const double a0 = f0 * e0.vx / e0.l, a1 = f1 * e1.vx / e1.l;
  const double b0 = f0 * e0.vy / e0.l, b1 = f1 * e1.vy / e1.l;
  V2 gu0{g0[1].x * -f0 + g0[3].x * a0, g0[1].y * -f0 + g0[3].y * a0};
  V2 gu1{g1[1].x * -f1 + g1[3].x * a1, g1[1].y * -f1 + g1[3].y * a1};
  V2 gv0{g0[2].x * -f0 + g0[3].x * b0, g0[2].y * -f0 + g0[3].y * b0};
  V2 gv1{g1[2].x * -f1 + g1[3].x * b1, g1[2].y * -f1 + g1[3].y * b1};
  const double R = s->M.p.gas_const;
  const double t0 = 1.0 / (R * std::pow(e0.l, 2.0)), t1 = 1.0 / (R * std::pow(e1.l, 2.0));
  WallGrad w;
  w.gu = V2{gu0.x * 0.5 + gu1.x * 0.5, gu0.y * 0.5 + gu1.y * 0.5};
  w.gv = V2{gv0.x * 0.5 + gv1.x * 0.5, gv0.y * 0.5 + gv1.y * 0.5};
  w.gT = V2{(g0[3].x * t0) * 0.5 + (g1[3].x * t1) * 0.5, (g0[3].y * t0) * 0.5 + (g1[3].y * t1) * 0.5};
  return w;
}
}  // namespace

// src/force.cc:118-217: pressure and viscous forces on the boundary faces whose
// tag is in tags, in bface order.  The reference takes the boundary condition
// of bface[force_index] (its loop index) rather than of the face summed; kept.
void orc_forces(orc_solver* s, int force_index, const int* tags, int ntags, double* out) {
  const Mat& M = s->M;
  if (M.p.model == 1) compute_gradients(s);
  const Face& bi = s->bface[force_index];
  const BC& b = s->bcof(bi.type);
  V2 inv{0.0, 0.0}, vis{0.0, 0.0};
  for (size_t f = 0; f < s->bface.size(); ++f) {
    const Face& bf = s->bface[f];
    bool in = false;
    for (int k = 0; k < ntags; ++k) in = in || tags[k] == bf.type;
    if (!in) continue;
    const int v0 = bf.v[0], v1 = bf.v[1];
    Prim st[2] = {s->W[v0], s->W[v1]};
    bc_apply1(b, s->bstate_at(force_index, v0), bi, st[0]);
    bc_apply1(b, s->bstate_at(force_index, v1), bi, st[1]);
    const Ent e0 = prim2ent(M, st[0]), e1 = prim2ent(M, st[1]);
    const double p = 0.5 * (st[0].p + st[1].p);
    inv.x += bf.normal.x * p;
    inv.y += bf.normal.y * p;
    if (M.p.model == 1) {
      const double T = (st[0].T + st[1].T) / 2.0;
      const double mu = M.viscosity(T);
      const WallGrad w = wall_gradients(s, v0, v1, e0, e1);
      const V2 gu = w.gu, gv = w.gv;
      const double div = gu.x + gv.y;
      const double sxx = 2.0 * mu * (gu.x - (1.0 / 3.0) * div);
      const double syy = 2.0 * mu * (gv.y - (1.0 / 3.0) * div);
      const double sxy = mu * (gu.y + gv.x);
      vis.x += -(sxx * bf.normal.x + sxy * bf.normal.y);
      vis.y += -(sxy * bf.normal.x + syy * bf.normal.y);
    }
  }
  out[0] = inv.x; out[1] = inv.y; out[2] = vis.x; out[3] = vis.y;
}

// The entropy-variable gradients the reference's dE holds: after a step those
// its last stage computed, of that stage's input state (src/fv.cc:1476, :2295-2304);
// before any step, of the current state.
void orc_stage_gradients(orc_solver* s, double* out) {
  if (s->dE_stale) compute_gradients(s);
  for (int i = 0; i < s->nv; ++i)
    for (int k = 0; k < 4; ++k) {
      out[8 * i + 2 * k] = s->dE[(size_t)i * 4 + k].x;
      out[8 * i + 2 * k + 1] = s->dE[(size_t)i * 4 + k].y;
    }
}

// src/output_skin_friction.cc:43-95 and src/output_heatflux.cc:43-90: for each
// boundary face whose tag is in tags, in bface order, the skin friction
// cf = f x n (f the viscous stress on the unit normal n) and the heat flux
// -k gradT.n, from the end points' states after the face's own boundary
// condition and the gradients dE holds.
void orc_surface(orc_solver* s, const int* tags, int ntags, double* out) {
  const Mat& M = s->M;
  if (s->dE_stale) compute_gradients(s);
  int j = 0;
  for (size_t f = 0; f < s->bface.size(); ++f) {
    const Face& bf = s->bface[f];
    bool in = false;
    for (int k = 0; k < ntags; ++k) in = in || tags[k] == bf.type;
    if (!in) continue;
    const int v0 = bf.v[0], v1 = bf.v[1];
    const BC& b = s->bcof(bf.type);
    Prim st[2] = {s->W[v0], s->W[v1]};
    bc_apply1(b, s->bstate_at((int)f, v0), bf, st[0]);
    bc_apply1(b, s->bstate_at((int)f, v1), bf, st[1]);
    const Ent e0 = prim2ent(M, st[0]), e1 = prim2ent(M, st[1]);
    const double T = (st[0].T + st[1].T) / 2.0;
    const double mu = M.viscosity(T);
    const double kc = mu * M.Cp / M.p.prandtl;
    const WallGrad w = wall_gradients(s, v0, v1, e0, e1);
    const double div = w.gu.x + w.gv.y;
    const double sxx = 2.0 * mu * (w.gu.x - (1.0 / 3.0) * div);
    const double syy = 2.0 * mu * (w.gv.y - (1.0 / 3.0) * div);
    const double sxy = mu * (w.gu.y + w.gv.x);
    const double im = 1.0 / bf.measure;
    const V2 n{bf.normal.x * im, bf.normal.y * im};
    const double fx = sxx * n.x + sxy * n.y;
    const double fy = sxy * n.x + syy * n.y;
    out[2 * j] = fx * n.y - fy * n.x;
    out[2 * j + 1] = (w.gT.x * n.x + w.gT.y * n.y) * -kc;
    ++j;
  }
}

int orc_set_bc_state(orc_solver* s, const int* bface_v, const double* state) {
  // state: per input bface b and end point k (bface_v[2b+k]) the values (T,u,v,p)
  std::map<std::pair<int, int>, int> in;
  const int nb = (int)s->bface.size();
  for (int b = 0; b < nb; ++b)
    in[{std::min(bface_v[2 * b], bface_v[2 * b + 1]), std::max(bface_v[2 * b], bface_v[2 * b + 1])}] = b;
  s->bst.assign(2 * (size_t)nb, Prim{0, 0, 0, 0});
  s->dE_stale = true;
  for (int bi = 0; bi < nb; ++bi) {
    const Face& f = s->bface[bi];
    auto it = in.find({std::min(f.v[0], f.v[1]), std::max(f.v[0], f.v[1])});
    if (it == in.end()) { g_err = "bface not in the input list"; return -1; }
    const int b = it->second;
    for (int e = 0; e < 2; ++e) {
      const int k = bface_v[2 * b + 1] == f.v[e] ? 1 : 0;
      const double* q = state + 8 * (size_t)b + 4 * k;
      s->bst[2 * (size_t)bi + e] = Prim{q[0], q[1], q[2], q[3]};
    }
  }
  return 0;
}

int orc_set_bc_time_fn(orc_solver* s, const int* bface_v, orc_bc_time_fn fn, void* user) {
  s->bt_fn = fn;
  s->bt_user = user;
  if (!fn) return 0;
  const int nb = (int)s->bface.size();
  s->bt_bv.assign(bface_v, bface_v + 2 * (size_t)nb);
  // start from the values in force: per input bface end point
  s->bt_in.assign(8 * (size_t)nb, 0.0);
  std::map<std::pair<int, int>, int> at;
  for (int bi = 0; bi < nb; ++bi)
    at[{std::min(s->bface[bi].v[0], s->bface[bi].v[1]), std::max(s->bface[bi].v[0], s->bface[bi].v[1])}] = bi;
  for (int b = 0; b < nb; ++b) {
    auto it = at.find({std::min(bface_v[2 * b], bface_v[2 * b + 1]), std::max(bface_v[2 * b], bface_v[2 * b + 1])});
    if (it == at.end()) { g_err = "bface not in the input list"; return -1; }
    for (int k = 0; k < 2; ++k) {
      const Prim q = s->bstate_at(it->second, bface_v[2 * b + k]);
      double* o = &s->bt_in[8 * (size_t)b + 4 * k];
      o[0] = q.T; o[1] = q.u; o[2] = q.v; o[3] = q.p;
    }
  }
  return 0;
}

// src/fv.cc:1712-1795
void orc_error_norms(const orc_solver* s, const double* ex, double* out) {
  const Mat& M = s->M;
  double L1[8] = {0}, L2[8] = {0}, Li[8] = {0};
  for (int i = 0; i < s->nv; ++i) {
    double area = s->dcarea[i], w = s->pweight[i];
    Prim pe{ex[4 * i], ex[4 * i + 1], ex[4 * i + 2], ex[4 * i + 3]};
    const Prim& pw = s->W[i];
    double e[8];
    e[0] = std::fabs(pe.T - pw.T);
    e[1] = std::fabs(pe.p - pw.p);
    e[2] = std::fabs(pe.u - pw.u);
    e[3] = std::fabs(pe.v - pw.v);
    e[4] = 0.0;
    e[5] = std::fabs(M.density(pe) - M.density(pw));
    e[6] = std::fabs(M.mach(pe) - M.mach(pw));
    e[7] = std::fabs(M.entropy(pe) - M.entropy(pw));
    for (int k = 0; k < 5; ++k) {
      Li[k] = std::max(Li[k], e[k]);
      L1[k] += e[k] * (area * w);
      L2[k] += e[k] * e[k] * (area * w);
    }
    for (int k = 5; k < 8; ++k) {
      Li[k] = std::max(Li[k], e[k]);
      L1[k] += e[k] * area * w;
      L2[k] += std::pow(e[k], 2.0) * area * w;
    }
  }
  for (int k = 0; k < 8; ++k) {
    out[3 * k] = L1[k]; out[3 * k + 1] = std::sqrt(L2[k]); out[3 * k + 2] = Li[k];
  }
}

// src/fv.cc:2175-2225
void orc_global(const orc_solver* s, double* ke, double* ent) {
  const Mat& M = s->M;
  double K = 0, S = 0;
  for (int i = 0; i < s->nv; ++i) {
    const Prim& w = s->W[i];
    double wt = s->pweight[i];
    K += (0.5 * M.density(w) * (w.u * w.u + w.v * w.v) * s->dcarea[i]) * wt;
    S += (-M.density(w) * M.entropy(w) / (M.p.gamma - 1.0)) * s->dcarea[i] * wt;
  }
  *ke = K; *ent = S;
}

// Enstrophy, 1/2 sum_i omega_i^2 dcarea_i w_i, beside compute_global's
// integrals (src/fv.cc:2175-2225; the reference has no enstrophy): the
// vorticity of Material::Vorticity (src/material.h:622-627) from the entropy-
// variable gradients of the current state (compute_gradients),
// omega = R T (dE[vy].x + v dE[l].x - dE[vx].y - u dE[l].y).
double orc_enstrophy(orc_solver* s) {
  compute_gradients(s);
  const Mat& M = s->M;
  double Z = 0;
  for (int i = 0; i < s->nv; ++i) {
    const Prim& w = s->W[i];
    const V2* g = &s->dE[(size_t)i * 4];
    const double om = M.p.gas_const * w.T * (g[2].x + w.v * g[3].x - g[1].y - w.u * g[3].y);
    Z += (0.5 * om * om * s->dcarea[i]) * s->pweight[i];
  }
  return Z;
}

// the Liou sensor of the current state, 0/1 per vertex (compute_ssw)
void orc_ssw(orc_solver* s, int* out) {
  compute_ssw(s);
  for (int i = 0; i < s->nv; ++i) out[i] = s->ssw[i] > 0 ? 1 : 0;
}

// src/fv.cc:595-694 (increment_moments): first sample sets mean = x, S2 = x*x,
// M2 = 0; sample k >= 2 (1-based) updates primitives with PrimVar::equ
// (src/primvar.h:213-221): M = x*(1/k) + M*((k-1)/k), and derived fields with
// M = M + (x - M)/k; S2 += x*x and M2 += (x - M_old)*(x - M_new) for both.
void orc_moments(int nv, int S, const double* prim, const double* gamma, const double* gas_const,
                 double* mean, double* m2, double* s2) {
  (void)gamma;
  for (int s = 0; s < S; ++s) {
    const double k = s + 1.0;
    for (int i = 0; i < nv; ++i) {
      const double* w = prim + ((size_t)s * nv + i) * 4;
      const double x[5] = {w[0], w[1], w[2], w[3], w[3] / (gas_const[s] * w[0])};
      double* M = mean + 5 * (size_t)i;
      double* Q = m2 + 5 * (size_t)i;
      double* Z = s2 + 5 * (size_t)i;
      for (int q = 0; q < 5; ++q) {
        if (s == 0) {
          M[q] = x[q]; Z[q] = x[q] * x[q]; Q[q] = 0.0;
          continue;
        }
        const double mo = M[q];
        const double mn = q < 4 ? x[q] * (1.0 / k) + mo * ((k - 1.0) / k) : mo + (x[q] - mo) / k;
        M[q] = mn;
        Z[q] += x[q] * x[q];
        Q[q] += (x[q] - mo) * (x[q] - mn);
      }
    }
  }
}

double orc_logavg(double a, double b) { return logavg(a, b); }

void orc_viscous_flux(const orc_params* p, const double* ent, const double* dE, const double* n,
                      int wall, int adiabatic, double* F) {
  Mat M; M.p = *p; M.Cp = p->gamma * p->gas_const / (p->gamma - 1.0);
  V2 g[4];
  for (int k = 0; k < 4; ++k) g[k] = V2{dE[2 * k], dE[2 * k + 1]};
  V2 nrm[3] = {V2{n[0], n[1]}, V2{0, 0}, V2{0, 0}};
  Flx out[3];
  viscous_flux(M, Ent{ent[0], ent[1], ent[2], ent[3]}, g, nrm, out, wall != 0, adiabatic != 0);
  F[0] = out[0].m; F[1] = out[0].mx; F[2] = out[0].my; F[3] = out[0].e;
}

void orc_num_flux(const orc_params* p, const double* xl, const double* xr, const double* Wl,
                  const double* Wr, const double* dEl, const double* dEr, const double* n,
                  double ssw, double* F) {
  Mat M; M.p = *p; M.Cp = p->gamma * p->gas_const / (p->gamma - 1.0);
  V2 gl[4], gr[4];
  for (int k = 0; k < 4; ++k) { gl[k] = V2{dEl[2 * k], dEl[2 * k + 1]}; gr[k] = V2{dEr[2 * k], dEr[2 * k + 1]}; }
  Flx f = num_flux(M, V2{xl[0], xl[1]}, V2{xr[0], xr[1]}, Prim{Wl[0], Wl[1], Wl[2], Wl[3]},
                   Prim{Wr[0], Wr[1], Wr[2], Wr[3]}, gl, gr, V2{n[0], n[1]}, ssw);
  F[0] = f.m; F[1] = f.mx; F[2] = f.my; F[3] = f.e;
}
double orc_limited_slope(const orc_params* p, double ul, double ur, double h) {
  return limited_slope(*p, ul, ur, h);
}
void orc_reconstruct(const orc_params* p, double h, const double* zl, const double* zr,
                     const double* gl, const double* gr, double* rl, double* rr) {
  reconstruct(*p, h, zl, zr, gl, gr, rl, rr);
}

}  // extern "C"
