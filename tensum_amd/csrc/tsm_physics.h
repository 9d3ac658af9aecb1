// tsm_physics.h -- pointwise gas physics and numerical fluxes, device side.
//
// Restates, for 2-D FP64 states (T, u, v, p), the reference's pointwise
// functions with the same operation order (src/material.h, src/material.cc,
// src/kepec.cc, src/kep.cc, src/roe_ec.cc, src/simple_avg.cc,
// src/kepes_tecno_roe.cc, src/kepes_tecno_rusanov.cc, src/reconstruct.cc,
// src/limiter.h).  The reference's z components are identically zero in 2-D
// and every term they contribute is an exact +-0, so they are not carried.
// Compiled with -ffp-contract=off so that no multiply-add is fused, except in
// the KEPES-TeCNO flux, which opts in (see there).
#pragma once
#include <hip/hip_runtime.h>

namespace tsm {

enum { F_ROE = 0, F_RUS = 1, F_KEPEC = 2, F_ROEEC = 3, F_KEP = 4, F_AVG = 5 };
enum { R_FIRST = 0, R_SECOND = 1, R_MUSCL = 2, R_ALBADA = 3, R_MINMOD = 4,
       R_TVB = 5, R_ENO2 = 6 };

struct Gas {
  double gamma, R;
};

struct Prim { double T, u, v, p; };
struct Cons { double m, mx, my, e; };
struct Ent { double f, vx, vy, l; };

// Quotient and square root of normal-range FP64 operands.  These are the
// refinement steps of the IEEE sequences the compiler emits for gfx950 (the
// hardware reciprocal or reciprocal-square-root estimate, two Newton-Raphson
// steps, one residual correction), without those sequences' range scaling and
// special-value fix-up, which leave normal-range operands unchanged: the results
// are bit for bit those of `/` and sqrt() for every state the solver holds,
// in 3-5 fewer instructions and without the scale-flag dependency.  A zero or
// infinite operand gives a NaN or an infinity, as the blown-up states it
// comes from do anyway.
__device__ __forceinline__ double fdiv(double n, double d) {
  double r = __builtin_amdgcn_rcp(d);
  double e = __builtin_fma(-d, r, 1.0);
  r = __builtin_fma(r, e, r);
  e = __builtin_fma(-d, r, 1.0);
  r = __builtin_fma(r, e, r);
  const double q = n * r;
  e = __builtin_fma(-d, q, n);
  return __builtin_fma(e, r, q);
}
__device__ __forceinline__ double fsqrt(double x) {
  const double y = __builtin_amdgcn_rsq(x);
  double g = x * y;
  double h = y * 0.5;
  const double r = __builtin_fma(-h, g, 0.5);
  g = __builtin_fma(g, r, g);
  double d = __builtin_fma(-g, g, x);
  h = __builtin_fma(h, r, h);
  g = __builtin_fma(d, h, g);
  d = __builtin_fma(-g, g, x);
  return __builtin_fma(d, h, g);
}

__device__ __forceinline__ double density(const Gas& g, const Prim& w) {
  return fdiv(w.p, g.R * w.T);
}
__device__ __forceinline__ double sound_speed(const Gas& g, const Prim& w) {
  return fsqrt(g.gamma * g.R * w.T);
}
__device__ __forceinline__ double entropy(const Gas& g, const Prim& w) {
  return log(w.p) - g.gamma * log(density(g, w));
}

// Logarithmic mean (src/material.h:24-42): F(f) = atanh(f)/f with
// f = (b - a)/(b + a), by its series 1 + u/3 + u^2/5 + u^3/7 (u = f^2 < 1e-2)
// or by log(b/a)/(2f).  The reference forms f as (xi - 1)/(xi + 1) with
// xi = b/a and divides the series terms; forming f with one division and
// multiplying by the reciprocals gives the same value to within an ulp with
// two divisions instead of five on the series branch.
// F(f) of the logarithmic mean of a and b for a given f = (b - a)/(b + a)
__device__ __forceinline__ double logavg_F(double f, double a, double b) {
  const double u = f * f;
  if (u < 1.0e-2) {
    const double u2 = u * u;
    const double u3 = u2 * u;
    return 1.0 + u * (1.0 / 3.0) + u2 * (1.0 / 5.0) + u3 * (1.0 / 7.0);
  }
  return fdiv(log(fdiv(b, a)) * 0.5, f);
}
__device__ __forceinline__ double logavg(double a, double b) {
  const double sum = a + b;
  const double f = fdiv(b - a, sum);
  const double u = f * f;
  double F;
  if (u < 1.0e-2) {
    const double u2 = u * u;
    const double u3 = u2 * u;
    F = 1.0 + u * (1.0 / 3.0) + u2 * (1.0 / 5.0) + u3 * (1.0 / 7.0);
  } else {
    F = fdiv(log(fdiv(b, a)) * 0.5, f);
  }
  return fdiv(0.5 * sum, F);
}

// src/material.h:501-513
__device__ __forceinline__ Cons prim2con(const Gas& g, const Prim& w) {
  Cons c;
  double rho = density(g, w);
  c.m = rho;
  c.mx = w.u * rho;
  c.my = w.v * rho;
  c.e = fdiv(w.p, g.gamma - 1.0) + 0.5 * (w.u * w.u + w.v * w.v) * rho;
  return c;
}
// src/material.h:518-529
__device__ __forceinline__ Prim con2prim(const Gas& g, const Cons& c) {
  Prim w;
  double ir = fdiv(1.0, c.m);
  w.u = c.mx * ir;
  w.v = c.my * ir;
  w.p = (g.gamma - 1.0) * (c.e - fdiv(0.5 * (c.mx * c.mx + c.my * c.my), c.m));
  w.T = fdiv(w.p, g.R * c.m);
  return w;
}
// src/material.h:534-546
__device__ __forceinline__ Ent prim2ent(const Gas& g, const Prim& w) {
  Ent e;
  double rho = density(g, w);
  double beta = fdiv(rho, 2 * w.p);
  double s = entropy(g, w);
  e.f = fdiv(g.gamma - s, g.gamma - 1) - (w.u * w.u + w.v * w.v) * beta;
  e.vx = w.u * (2 * beta);
  e.vy = w.v * (2 * beta);
  e.l = -2 * beta;
  return e;
}

// src/limiter.h:14-77
__device__ __forceinline__ double minmod_tvd(double a, double b) {
  if (a < 0.0 && b < 0.0) return -fmin(-a, -b);
  if (a > 0.0 && b > 0.0) return fmin(a, b);
  return 0.0;
}
__device__ __forceinline__ double minabs(double a, double b) {
  return (fabs(a) < fabs(b)) ? a : b;
}

// src/reconstruct.cc:101-134
template <int REC>
__device__ __forceinline__ double limited_slope(double ul, double ur, double h, double tvbM) {
  if (REC == R_ALBADA) {
    const double eps = 1.0e-14;
    double s = fmax(0.0, fdiv(2 * ul * ur + eps, ul * ul + ur * ur + eps));
    const double K = 1.0 / 3.0;
    return 0.25 * s * ((1.0 - K * s) * ul + (1.0 + K * s) * ur);
  } else if (REC == R_MINMOD) {
    return 0.5 * minmod_tvd(ul, ur);
  } else if (REC == R_ENO2) {
    return 0.5 * minabs(ul, ur);
  } else {  // R_TVB
    double c = 0.5 * (ul + ur);
    double r = (fabs(c) < tvbM * (h * h)) ? c : minmod_tvd(ul, ur);
    return 0.5 * r;
  }
}

// src/reconstruct.cc:13-96: one component
template <int REC>
__device__ __forceinline__ void reconstruct1(double zl, double zr, double gl, double gr,
                                             double h, double tvbM, double& rl, double& rr) {
  if (REC == R_FIRST) {
    rl = zl; rr = zr;
  } else if (REC == R_SECOND) {
    rl = zl + 0.5 * gl;
    rr = zr - 0.5 * gr;
  } else if (REC == R_MUSCL) {
    const double K = 1.0 / 3.0;
    double ds = zr - zl;
    double dl = 2.0 * gl - ds;
    rl = zl + (dl * (1.0 - K) + ds * (1.0 + K)) * 0.25;
    double dr = 2.0 * gr - ds;
    rr = zr - (dr * (1.0 - K) + ds * (1.0 + K)) * 0.25;
  } else {
    double ds = zr - zl;
    double dl = 2.0 * gl - ds;
    rl = zl + limited_slope<REC>(dl, ds, h, tvbM);
    double dr = 2.0 * gr - ds;
    rr = zr - limited_slope<REC>(dr, ds, h, tvbM);
  }
}

struct FluxOpts {
  double ev_alpha, ev_beta, tvbM;
  int nd_active;
};

// Per-vertex quantities the two-point fluxes use, computed once per vertex with
// the reference's own expressions (Density, 0.5/(R T), Entropy, sound_speed).
// V: entropy variables of the TeCNO dissipation (src/kepes_tecno_roe.cc:
// 98-113), (gamma - s)/(gamma - 1) - beta |u|^2, 2 beta u, 2 beta v, -2 beta.
struct Side {
  Prim w;
  double rho, beta, c;
  double V[4];
};
template <bool NEED_V, bool NEED_C>
__device__ __forceinline__ Side make_side(const Gas& g, const Prim& w) {
  Side q;
  q.w = w;
  q.rho = density(g, w);
  q.beta = fdiv(0.5, g.R * w.T);
  if (NEED_V) {
    const double s = log(w.p) - g.gamma * log(q.rho);
    q.V[0] = fdiv(g.gamma - s, g.gamma - 1.0) - q.beta * (w.u * w.u + w.v * w.v);
    q.V[1] = 2 * w.u * q.beta;
    q.V[2] = 2 * w.v * q.beta;
    q.V[3] = -2 * q.beta;
  } else {
    q.V[0] = q.V[1] = q.V[2] = q.V[3] = 0.0;
  }
  q.c = NEED_C ? sound_speed(g, w) : 0.0;
  return q;
}

// src/kepec.cc:10-38 (per unit normal)
__device__ __forceinline__ Cons kepec_flux(const Gas& g, const Prim& l, const Prim& r,
                                           double unx, double uny) {
  double rhol = density(g, l), rhor = density(g, r);
  double rho = logavg(rhol, rhor);
  double vel2 = 0.5 * ((l.u * l.u + l.v * l.v) + (r.u * r.u + r.v * r.v));
  double betal = fdiv(0.5, g.R * l.T);
  double betar = fdiv(0.5, g.R * r.T);
  double beta = logavg(betal, betar);
  double p = fdiv(0.5 * (rhol + rhor), betal + betar);
  double vx = l.u * 0.5 + r.u * 0.5, vy = l.v * 0.5 + r.v * 0.5;
  double vn = vx * unx + vy * uny;
  Cons F;
  F.m = rho * vn;
  F.mx = unx * p + vx * F.m;
  F.my = uny * p + vy * F.m;
  F.e = 0.5 * (fdiv(1.0, (g.gamma - 1.0) * beta) - vel2) * F.m + (F.mx * vx + F.my * vy);
  return F;
}

// kepec_flux with the per-vertex density and beta already formed (same values)
__device__ __forceinline__ Cons kepec_flux_s(const Gas& g, const Side& L, const Side& Rr,
                                             double unx, double uny) {
  const Prim& l = L.w;
  const Prim& r = Rr.w;
  double rho = logavg(L.rho, Rr.rho);
  double vel2 = 0.5 * ((l.u * l.u + l.v * l.v) + (r.u * r.u + r.v * r.v));
  double beta = logavg(L.beta, Rr.beta);
  double p = fdiv(0.5 * (L.rho + Rr.rho), L.beta + Rr.beta);
  double vx = l.u * 0.5 + r.u * 0.5, vy = l.v * 0.5 + r.v * 0.5;
  double vn = vx * unx + vy * uny;
  Cons F;
  F.m = rho * vn;
  F.mx = unx * p + vx * F.m;
  F.my = uny * p + vy * F.m;
  F.e = 0.5 * (fdiv(1.0, (g.gamma - 1.0) * beta) - vel2) * F.m + (F.mx * vx + F.my * vy);
  return F;
}

// src/kep.cc:10-38
__device__ __forceinline__ Cons kep_flux(const Gas& g, const Prim& l, const Prim& r,
                                         double unx, double uny) {
  double rhol = density(g, l), rhor = density(g, r);
  double g1 = g.gamma - 1.0;
  double cl = sound_speed(g, l), cr = sound_speed(g, r);
  double Hl = fdiv(cl * cl, g1) + 0.5 * (l.u * l.u + l.v * l.v);
  double Hr = fdiv(cr * cr, g1) + 0.5 * (r.u * r.u + r.v * r.v);
  double p = 0.5 * (l.p + r.p);
  double rho = 0.5 * (rhol + rhor);
  double H = 0.5 * (Hl + Hr);
  double vx = l.u * 0.5 + r.u * 0.5, vy = l.v * 0.5 + r.v * 0.5;
  double vn = vx * unx + vy * uny;
  Cons F;
  F.m = rho * vn;
  F.mx = unx * p + vx * F.m;
  F.my = uny * p + vy * F.m;
  F.e = rho * H * vn;
  return F;
}

// src/roe_ec.cc:10-44
__device__ __forceinline__ Cons roe_ec_flux(const Gas& g, const Prim& l, const Prim& r,
                                            double unx, double uny) {
  double rhol = density(g, l), rhor = density(g, r);
  double pl = l.p, pr = r.p;
  double bl = fsqrt(fdiv(rhol, pl)), br = fsqrt(fdiv(rhor, pr));
  double z1b = 0.5 * (bl + br);
  double zx = l.u * (0.5 * bl) + r.u * (0.5 * br);
  double zy = l.v * (0.5 * bl) + r.v * (0.5 * br);
  double z5b = 0.5 * (bl * pl + br * pr);
  double zn = zx * unx + zy * uny;
  double z1l = logavg(bl, br);
  double z5l = logavg(bl * pl, br * pr);
  Cons F;
  F.m = z5l * zn;
  double c1 = fdiv(z5b, z1b), c2 = fdiv(F.m, z1b);
  F.mx = unx * c1 + zx * c2;
  F.my = uny * c1 + zy * c2;
  F.e = fdiv(0.5 * (fdiv(fdiv(g.gamma + 1.0, g.gamma - 1.0) * F.m, z1l) + (F.mx * zx + F.my * zy)), z1b);
  return F;
}

// src/simple_avg.cc:10-46
__device__ __forceinline__ Cons avg_flux(const Gas& g, const Prim& l, const Prim& r,
                                         double unx, double uny) {
  double rhol = density(g, l), rhor = density(g, r);
  double El = 0.5 * rhol * (l.u * l.u + l.v * l.v) + fdiv(l.p, g.gamma - 1);
  double Er = 0.5 * rhor * (r.u * r.u + r.v * r.v) + fdiv(r.p, g.gamma - 1);
  double vnl = l.u * unx + l.v * uny, vnr = r.u * unx + r.v * uny;
  double flm = rhol * vnl, frm = rhor * vnr;
  double flx = unx * l.p + l.u * flm, fly = uny * l.p + l.v * flm, fle = (El + l.p) * vnl;
  double frx = unx * r.p + r.u * frm, fry = uny * r.p + r.v * frm, fre = (Er + r.p) * vnr;
  Cons F;
  F.m = (0.0 + flm * 0.5) + frm * 0.5;
  F.mx = (0.0 + flx * 0.5) + frx * 0.5;
  F.my = (0.0 + fly * 0.5) + fry * 0.5;
  F.e = (0.0 + fle * 0.5) + fre * 0.5;
  return F;
}

// Gradients of the entropy variables (f, vx, vy, l; x then y per component)
// projected on the edge vector r = x_r - x_l: the only form in which the
// reconstruction uses them (src/reconstruct.cc:13-96).
__device__ __forceinline__ void project_grad(const double* g, double rx, double ry, double* gd) {
#pragma unroll
  for (int j = 0; j < 4; ++j) gd[j] = g[2 * j] * rx + g[2 * j + 1] * ry;
}

// src/kepes_tecno_roe.cc:11-173, src/kepes_tecno_rusanov.cc:11-164.
// gdl/gdr: left/right vertex gradients projected on the edge (project_grad).
template <bool ROE, int REC>
__device__ __forceinline__ Cons tecno_flux(const Gas& g, const FluxOpts& o, double rx, double ry,
                                           const Side& L, const Side& Rr, const double* gdl,
                                           const double* gdr, double unx, double uny,
                                           double ssw) {
  // Multiply-adds fused here: the flux is bounded by FP64 issue, and fusing
  // changes each product-sum by at most an ulp.
#pragma clang fp contract(fast)
  const double gam = g.gamma;
  const Prim& l = L.w;
  const Prim& r = Rr.w;
  // 1/(gamma-1) and 1/gamma are the same for every face of a sample (the
  // compiler forms them once per thread); 1/beta serves a and the energy flux.
  const double ig1 = fdiv(1.0, gam - 1.0), igam = fdiv(1.0, gam);
  double rhol = L.rho, rhor = Rr.rho;
  double rho = logavg(rhol, rhor);
  double betal = L.beta, betar = Rr.beta;
  // beta's logarithmic mean enters only through 1/beta_hat: with
  // r = 1/(beta_l + beta_r), f = (beta_r - beta_l) r and beta_hat = 1/(2 r F(f)),
  // so 1/beta_hat = 2 F(f) r, and p_hat = (rho_l + rho_r) r / 2: one division
  // where the reference's expressions take four (same values to rounding)
  const double rb = fdiv(1.0, betal + betar);
  const double ib = 2.0 * logavg_F((betar - betal) * rb, betal, betar) * rb;
  double a = fsqrt(0.5 * gam * ib);
  double p = (0.5 * (rhol + rhor)) * rb;
  double vx = l.u * 0.5 + r.u * 0.5, vy = l.v * 0.5 + r.v * 0.5;
  double vn = vx * unx + vy * uny;
  double vel2 = 0.5 * ((l.u * l.u + l.v * l.v) + (r.u * r.u + r.v * r.v));

  // KEPEC central part with the averages above (identical values to a
  // separate kepec_flux call).
  Cons F;
  F.m = rho * vn;
  F.mx = unx * p + vx * F.m;
  F.my = uny * p + vy * F.m;
  F.e = 0.5 * (ig1 * ib - vel2) * F.m + (F.mx * vx + F.my * vy);
  if (!o.nd_active) return F;

  double vsq = vx * vx + vy * vy;
  double H = a * a * ig1 + 0.5 * vsq;
  double v1 = vx * uny - vy * unx;
  double Lam[4];
  if (ROE) {
    double vnl = l.u * unx + l.v * uny, vnr = r.u * unx + r.v * uny;
    double al = L.c, ar = Rr.c;
    double l2 = fabs(vn), l3 = fabs(vn);
    if (ssw > 0) l3 = fmax(l3, a);
    Lam[0] = (1 + o.ev_beta) * fabs(vn - a) + o.ev_alpha * fabs((vnl - al) - (vnr - ar));
    Lam[1] = l2;
    Lam[2] = l3;
    Lam[3] = (1 + o.ev_beta) * fabs(vn + a) + o.ev_alpha * fabs((vnl + al) - (vnr + ar));
  } else {
    double lr = fabs(vn) + fabs(a);
    Lam[0] = Lam[1] = Lam[2] = Lam[3] = lr;
  }
  const double* Vl = L.V;
  const double* Vr = Rr.V;
  double D[4];
  if (REC == R_ALBADA || REC == R_TVB) {
    // Limiters that are not positively homogeneous (van Albada's epsilon, the
    // TVB threshold M h^2) see the characteristic variables of the scaled
    // eigenvectors, as the reference forms them (src/kepes_tecno_roe.cc:60-73).
    double const1 = fsqrt(rho * igam);
    double psq = fsqrt(p);
    double s0 = sqrt(0.5) * const1, s1 = fsqrt(gam - 1.0) * const1;
    // scaled right eigenvectors, Rm[row][col]; rows: mass, x-mom, y-mom, energy
    double Rm[4][4];
    Rm[0][0] = 1 * s0;               Rm[0][1] = 1 * s1;     Rm[0][2] = 0 * psq;     Rm[0][3] = 1 * s0;
    Rm[1][0] = (vx - a * unx) * s0;  Rm[1][1] = vx * s1;    Rm[1][2] = uny * psq;   Rm[1][3] = (vx + a * unx) * s0;
    Rm[2][0] = (vy - a * uny) * s0;  Rm[2][1] = vy * s1;    Rm[2][2] = -unx * psq;  Rm[2][3] = (vy + a * uny) * s0;
    Rm[3][0] = (H - a * vn) * s0;    Rm[3][1] = (0.5 * vsq) * s1; Rm[3][2] = v1 * psq; Rm[3][3] = (H + a * vn) * s0;
    double h = 0.0;
    if (REC == R_TVB) h = fsqrt(rx * rx + ry * ry);
    double dZ[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      double zl = 0.0, zr = 0.0, sgl = 0.0, sgr = 0.0;
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        zl += Rm[j][i] * Vl[j];
        zr += Rm[j][i] * Vr[j];
        sgl += Rm[j][i] * gdl[j];
        sgr += Rm[j][i] * gdr[j];
      }
      double zij, zji;
      reconstruct1<REC>(zl, zr, sgl, sgr, h, o.tvbM, zij, zji);
      dZ[i] = zji - zij;
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      double d = 0.0;
#pragma unroll
      for (int j = 0; j < 4; ++j) d += Rm[i][j] * Lam[j] * dZ[j];
      D[i] = d;
    }
  } else {
    // Positively homogeneous reconstructions (first, second, MUSCL, minmod,
    // ENO2: phi(c a, c b) = c phi(a, b) for c > 0): with R = R0 diag(s),
    // s = (sqrt(rho/(2 gam)), sqrt((gam-1) rho/gam), sqrt(p), sqrt(rho/(2 gam))),
    // the reconstructed jump in the scaled characteristic variables is s times
    // the one formed with the unscaled eigenvectors R0, so
    // D = R Lam dZ = R0 (s^2 Lam dZ0): no square roots of the scalings, the
    // projections exploit R0's zeros and ones, and only the jump
    // dz0 = R0^T (V_r - V_l) enters (src/reconstruct.cc:13-96 written as
    // z_ji - z_ij).  Same values to rounding.
    const double hv = 0.5 * vsq;
    auto proj = [&](double x0, double x1, double x2, double x3, double* z) {
      const double b = x0 + vx * x1 + vy * x2;
      const double t = b + H * x3;
      const double u = a * (unx * x1 + uny * x2 + vn * x3);
      z[0] = t - u;
      z[1] = b + hv * x3;
      z[2] = uny * x1 - unx * x2 + v1 * x3;
      z[3] = t + u;
    };
    double dz[4], gl[4], gr[4];
    proj(Vr[0] - Vl[0], Vr[1] - Vl[1], Vr[2] - Vl[2], Vr[3] - Vl[3], dz);
    if (REC != R_FIRST) {
      proj(gdl[0], gdl[1], gdl[2], gdl[3], gl);
      proj(gdr[0], gdr[1], gdr[2], gdr[3], gr);
    }
    const double rg = rho * igam;
    const double q[4] = {0.5 * rg, (gam - 1.0) * rg, p, 0.5 * rg};
    double w[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      double dZi;
      if (REC == R_FIRST) {
        dZi = dz[i];
      } else if (REC == R_SECOND) {
        dZi = dz[i] - 0.5 * (gl[i] + gr[i]);
      } else if (REC == R_MUSCL) {
        const double K = 1.0 / 3.0;
        const double dl = 2.0 * gl[i] - dz[i], dr = 2.0 * gr[i] - dz[i];
        dZi = dz[i] - ((dl + dr) * (1.0 - K) + (2.0 * dz[i]) * (1.0 + K)) * 0.25;
      } else {
        const double dl = 2.0 * gl[i] - dz[i], dr = 2.0 * gr[i] - dz[i];
        dZi = dz[i] - limited_slope<REC>(dl, dz[i], 0.0, 0.0) - limited_slope<REC>(dr, dz[i], 0.0, 0.0);
      }
      w[i] = (q[i] * Lam[i]) * dZi;
    }
    const double w13 = w[0] + w[1] + w[3], d30 = w[3] - w[0];
    D[0] = w13;
    D[1] = vx * w13 + a * unx * d30 + uny * w[2];
    D[2] = vy * w13 + a * uny * d30 - unx * w[2];
    D[3] = H * (w[0] + w[3]) + hv * w[1] + a * vn * d30 + v1 * w[2];
  }
  F.m -= 0.5 * D[0];
  F.mx -= 0.5 * D[1];
  F.my -= 0.5 * D[2];
  F.e -= 0.5 * D[3];
  return F;
}

// src/material.cc:110-127
__device__ __forceinline__ Cons inviscid_bflux(const Gas& g, const Prim& w, const Prim& wb,
                                               double nx, double ny) {
  double vnb = wb.u * nx + wb.v * ny;
  double vn = w.u * nx + w.v * ny;
  double rho = density(g, w);
  Cons F;
  F.m = rho * vnb;
  F.mx = nx * wb.p + w.u * F.m;
  F.my = ny * wb.p + w.v * F.m;
  F.e = (0.5 * rho * (w.u * w.u + w.v * w.v) + fdiv(g.gamma * w.p, g.gamma - 1.0)) * vnb +
        (wb.p - w.p) * vn;
  return F;
}

}  // namespace tsm
