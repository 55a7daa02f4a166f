// rt_physics_fast.h -- the mixed-precision variant of the radiation step
// physics: FP32 transport with hardware reciprocal / reciprocal square root,
// FP32 rate fits through the hardware exp2/log2, and the chemistry sub-step
// update (x, eint, photon absorption, accept/reject decisions) in FP64 with
// reciprocal-and-Newton division.  Same equations and control flow as the
// FP64 path in rt_physics.h; only the rounding differs.
#pragma once

#include "rt_physics.h"

namespace emma {

__device__ __forceinline__ float frcp32(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ float fexp2(float x) { return __builtin_amdgcn_exp2f(x); }
__device__ __forceinline__ float flog2(float x) { return __builtin_amdgcn_logf(x); }

// FP64 reciprocal: hardware estimate refined by two Newton steps (~1 ulp)
__device__ __forceinline__ double frcp64(double b) {
  double r = __builtin_amdgcn_rcp(b);
  r = fma(fma(-b, r, 1.0), r, r);
  r = fma(fma(-b, r, 1.0), r, r);
  return r;
}

// one Newton step: for quantities that set a sub-step length or feed the FP32
// rate fits, and for absorption factors kept to FP32 accuracy downstream
__device__ __forceinline__ double frcp64_1(double b) {
  const double r = __builtin_amdgcn_rcp(b);
  return fma(fma(-b, r, 1.0), r, r);
}

__device__ __forceinline__ double fdiv64(double a, double b) {
  const double r = frcp64(b);
  const double q = a * r;
  return fma(fma(-b, q, a), r, q);
}

// Levermore closure (rad_utils.c:448-475) with fast FP32 reciprocals.
// With s = sqrt(4 - 3 f^2) (f^2 <= 4/3), chi = (3 + 4 f^2) / (5 + 2 s) equals
// (5 - 2 s) / 3 (multiply by 5 - 2 s: the denominator becomes
// 25 - 4 (4 - 3 f^2) = 3 (3 + 4 f^2)), so (1 - chi)/2 = (s - 1)/3 and
// (3 chi - 1)/2 = 2 - s need no division.  Beyond f^2 = 4/3 the reference
// clamps the root to 0 and chi = (3 + 4 f^2) / 5.
__device__ __forceinline__ Closure<float> closure_fast(float fx, float fy, float fz, float ee, float c) {
#pragma clang fp contract(fast)
  // branch-free: the isotropic case (E <= 0 or F = 0: n = 0, f = 0) by selects
  Closure<float> q;
  const float c2e = ee * c * c;
  const float f2 = fx * fx + fy * fy + fz * fz;
  const bool aniso = (ee > 0.f) & (f2 > 1e-36f);
  const float ri = aniso ? __builtin_amdgcn_rsqf(f2) : 0.f;
  q.n[0] = fx * ri;
  q.n[1] = fy * ri;
  q.n[2] = fz * ri;
  const float ff = aniso ? f2 * ri * frcp32(c * ee) : 0.f;
  const float ff2 = ff * ff;
  const float s = __builtin_amdgcn_sqrtf(fmaxf(4.f - 3.f * ff2, 0.f));
  const bool clamped = ff2 > 4.f / 3.f;
  q.diag = c2e * (clamped ? 0.2f * (1.f - 2.f * ff2) : (s - 1.f) * (1.f / 3.f));
  q.arg = c2e * (clamped ? 0.2f * (2.f + 6.f * ff2) : 2.f - s);
  return q;
}

// the rate fits of chem_utils.c:37-187 in FP32: pow(x, a) = exp2(a log2 x)
// with log2(lambda) shared, exp(-E/T) = exp2(-E log2(e) / T)
__device__ __forceinline__ Rates rates_fast(double Td) {
#pragma clang fp contract(on)
  const float T = (float)Td;
  const float rT = frcp32(T);
  const float sT = __builtin_amdgcn_sqrtf(T);
  const float sT5 = sT * 3.16227766e-3f;  // sqrt(T/1e5)
  const float inv1 = frcp32(1.f + sT5);
  const float lam = 315614.f * rT;
  const float l2 = flog2(lam);
  const float LOG2E = 1.44269504f;
  Rates r;
  const float u6 = fexp2(0.376f * (l2 - 1.16992500f));   // (lam/2.25)^0.376
  r.A6 = (double)(3.435e-30f * T * fexp2(1.970f * l2 - 3.720f * flog2(1.f + u6)));
  const float ub = fexp2(0.407f * (l2 - 1.45417589f));   // (lam/2.74)^0.407
  r.alphab = (double)(2.753e-20f * fexp2(1.5f * l2 - 2.242f * flog2(1.f + ub)));
  r.A1 = (double)(fexp2(-157809.1f * LOG2E * rT) * 2.54e-21f * sT * inv1);
  r.A3 = (double)(fexp2(-118348.f * LOG2E * rT) * 7.5e-19f * inv1);
  r.A4 = (double)(2.13e-27f * sT);
  r.beta = (double)(1.17e-16f * sT * inv1 * fexp2(-157809.f * LOG2E * rT));
  return r;
}

// chemrad for one unsplit cell, mixed-precision variant of chem_cell<G>.
// Cooling (cooling of rt_physics.h) is evaluated with its common factors
// taken out: nH'^2 x [(A1 + A3)(1 + yHe)(1 - x) + (A6 + A4 (1 + yHe)^2) x] / 10.
// The flux F of each group only gets multiplied by the accepted sub-steps'
// absorption factors; instead of carrying F through the loop the chemistry
// returns that product (unit conversions in and out included) in fscale, and
// the caller forms F * fscale.
//
// The sub-step loop is split into begin (unit conversions, EMIN test), run
// (the loop, optionally for a limited number of tries, its whole state in
// ChemFast so that another thread can continue it bit for bit) and end (the
// conversions back).  The chemistry (and its rate fits) contracts
// multiply-adds only within one expression (fp contract(on), not fast), so
// the loop rounds the same in every kernel it is inlined into: the straggler
// pass takes exactly the sub-steps the undivided loop takes.
template <int G>
struct ChemFast {
  static constexpr int GA = G > 0 ? G : kMaxGrp;
  double egy[GA];  // photon densities (SI)
  double fsc[GA];  // product of the accepted absorption factors of the flux
  double sf[GA];   // source rates
  double x0, eint, tcur, fudgecool, aexp, nh;
  int nit, tries;
};

template <int G>
__device__ __forceinline__ int chem_fast_begin(const ChemConst& K, int ng, const double* e, const double* src,
                                               double nhplus, double eint_code, double nh, ChemFast<G>& s) {
#pragma clang fp contract(on)
  constexpr int GA = ChemFast<G>::GA;
#pragma unroll
  for (int g = 0; g < GA; g++) {
    if (g < ng) {
      s.egy[g] = e[g] * K.k_e;
      s.fsc[g] = 1.;
      s.sf[g] = src[g] * K.k_src * K.factgrp[g];  // source rate of the group
    }
  }
  const double rnh = frcp64(nh);
  s.x0 = nhplus * rnh;
  s.nh = nh;
  s.eint = eint_code * K.k_eint;
  s.nit = 0;
  s.tries = 0;
  s.aexp = K.aexp;
  s.fudgecool = K.fudgecool;
  s.tcur = 0.;
  if (s.egy[0] < kEmin) return -4;
  return 0;
}

// the sub-step loop until the cell's dt is covered (0), for at most `budget`
// more tries when budget >= 0 (1: the cell still runs, its state in s), or
// until the controller fails (-5).  keep > 0: a cell that, when its budget
// runs out, has more than `keep` sub-steps to go at the pace of its accepted
// ones so far, or none accepted yet, runs on instead of stopping
template <int G>
__device__ __forceinline__ int chem_fast_run(const ChemConst& K, int ng, int chemonly, ChemFast<G>& s, int budget,
                                             double keep = 0.) {
#pragma clang fp contract(on)
  constexpr int GA = ChemFast<G>::GA;
  double et[GA], rp[GA];
  const double notchem = (double)(!chemonly);
  const double nH = s.nh * K.k_e;
  const double rnH = frcp64(nH);
  const double n22 = (nH * 1e-6) * (nH * 1e-6) * 0.1;
  const double dt = K.dt;
  const double kT = 1. / (1.5 * kKboltz);
  const double yp = 1. + kyHe;
  double tcached = -1.;
  Rates r = {};  // evaluated on the first pass (tloc != tcached)
  double cs = 0., ct = 0.;  // cooling(x) = x (cs (1-x) + ct x) at the rates' T
  double x0 = s.x0, eint = s.eint, tcur = s.tcur, fudgecool = s.fudgecool, aexp = s.aexp;
  int tries = s.tries, nit = s.nit;
  int rc = 0;
  while (tcur < dt) {
    if (budget >= 0 && budget-- == 0) {
      if (keep > 0. && (nit == s.nit || (dt - tcur) * (double)(nit - s.nit) > keep * (tcur - s.tcur))) {
        budget = -1;  // a long chain: no limit
      } else {
        rc = 1;
        break;
      }
    }
    if (++tries > kMaxTries) {
      rc = -5;
      break;
    }
    const double tloc = eint * kT * frcp64_1(1. + x0) * rnH;
    if (!(tloc == tcached)) {
      r = rates_fast(tloc);
      tcached = tloc;
      cs = (r.A1 + r.A3) * yp * n22;
      ct = (r.A6 + r.A4 * (yp * yp)) * n22;
    }
    if (fudgecool < 1e-20) {
      rc = -5;
      break;
    }
    const double omx0 = 1. - x0;
    const double Cool0 = x0 * fma(ct, x0, cs * omx0);
    double ai = 0.;
#pragma unroll
    for (int g = 0; g < GA; g++)
      if (g < ng) ai += K.heat[g] * s.egy[g];
    ai *= nH * omx0;
    const double tcool = fabs(eint * frcp64_1(ai * notchem - Cool0));
    const double dtcool = fmin(fudgecool * tcool, dt - tcur);
    // the trial sub-step; every rejection test of the reference leads to the
    // same retry with fudgecool / 10, so the tests are combined
    const double c1 = omx0 * nH * dtcool;
    bool bad = false;
#pragma unroll
    for (int g = 0; g < GA; g++) {
      if (g < ng) {
        rp[g] = frcp64_1(1. + K.alphai[g] * c1);
        et[g] = chemonly ? s.egy[g] : fma(s.sf[g], dtcool, s.egy[g]) * rp[g];
        bad = bad || (et[g] < 0) || isnan(et[g]);
      }
    }
    ai = 0.;
#pragma unroll
    for (int g = 0; g < GA; g++)
      if (g < ng) ai += K.alphai[g] * et[g];
    ai *= notchem;
    const double RECT = r.alphab * x0 * x0 * nH * dtcool;
    const double xt = 1. - fdiv64(RECT + omx0, 1. + dtcool * (r.beta * x0 * nH + ai));
    bad = bad || (xt > 1.) || (xt < 0.) || isnan(xt);
    const double omxt = 1. - xt;
    const double Cool1 = xt * fma(ct, xt, cs * omxt);
    ai = 0.;
#pragma unroll
    for (int g = 0; g < GA; g++)
      if (g < ng) ai += K.heat[g] * et[g];
    ai *= nH * omxt;
    double eintt = eint + dtcool * (ai * notchem - Cool1);
    bad = bad || (eintt < 0.) || (fabs(eintt - eint) > kFracVar * eint);  // |eintt/eint - 1| > FRAC_VAR
    if (bad) {
      fudgecool *= 0.1;
      continue;
    }
    eintt = fmax(K.emin, eintt);
#pragma unroll
    for (int g = 0; g < GA; g++) {
      if (g < ng) {
        s.egy[g] = et[g];
        if (!chemonly) s.fsc[g] *= rp[g];
      }
    }
    eint = eintt;
    if (!K.static_cosmo) {  // chem_utils.c:737-756: N, F, eint scale by (a_old/a)^3,4,5
      const double aold = aexp;
      const double hubblet = K.H0 * sqrt(K.om / aexp + K.ov * (aexp * aexp)) / aexp * (1e3 / (1e6 * kParsec));
      aexp += hubblet * dtcool * aexp;
      const double ra = aold / aexp;
      const double r3 = ra * ra * ra, r4 = r3 * ra, r5 = r4 * ra;
#pragma unroll
      for (int g = 0; g < GA; g++) {
        if (g < ng) {
          s.egy[g] *= r3;
          s.fsc[g] *= r4;
        }
      }
      eint *= r5;
    }
    x0 = xt;
    tcur += dtcool;
    fudgecool = K.fudgecool;
    nit++;
    if ((nit == K.ncvgcool) && (K.ncvgcool != 0)) break;
  }
  s.x0 = x0;
  s.eint = eint;
  s.tcur = tcur;
  s.fudgecool = fudgecool;
  s.aexp = aexp;
  s.tries = tries;
  s.nit = nit;
  return rc;
}

template <int G>
__device__ __forceinline__ void chem_fast_end(const ChemConst& K, int ng, int chemonly, const ChemFast<G>& s,
                                              double* e, double* fscale, double& nhplus, double& eint_code,
                                              double& temp) {
#pragma clang fp contract(on)
  constexpr int GA = ChemFast<G>::GA;
  const double kT = 1. / (1.5 * kKboltz);
  double ki_e = K.ki_e, ki_f = K.ki_f, ki_eint = K.ki_eint, a3f = K.pa3;
  if (!K.static_cosmo) {
    const double ra = s.aexp / K.aexp;
    a3f = K.pa3 * ra * ra * ra;
    ki_e *= ra * ra * ra;
    ki_f *= ra * ra * ra * ra;
    ki_eint *= ra * ra * ra * ra * ra;
  }
  if (!chemonly) {
#pragma unroll
    for (int g = 0; g < GA; g++) {
      if (g < ng) {
        e[g] = fmax(s.egy[g] * ki_e, K.emin_grp[g]);
        fscale[g] = s.fsc[g] * (K.k_f * ki_f);
      }
    }
  }
  nhplus = s.x0 * s.nh;
  eint_code = s.eint * ki_eint;
  // E2T: eint_code back in physical units is eint itself
  temp = s.eint * kT * frcp64((s.nh * K.unit_N / a3f) * (1. + s.x0));
}

// budget >= 0: at most `budget` sub-step tries (see chem_fast_run for
// keep); returns 1 with e, fscale, nhplus, eint_code, temp untouched
// (fscale = 1) when the cell still runs
template <int G>
__device__ int chem_cell_fast(const ChemConst& K, int ng, double* e, double* fscale, const double* src,
                              double& nhplus, double& eint_code, double nh, double& temp, int chemonly, int& nit,
                              int budget = -1, double keep = 0.) {
  constexpr int GA = ChemFast<G>::GA;
#pragma unroll
  for (int g = 0; g < GA; g++)
    if (g < ng) fscale[g] = 1.;
  ChemFast<G> s;
  int st = chem_fast_begin<G>(K, ng, e, src, nhplus, eint_code, nh, s);
  nit = 0;
  if (st != 0) return st;
  st = chem_fast_run<G>(K, ng, chemonly, s, budget, keep);
  nit = s.nit;
  if (st != 0) return st;
  chem_fast_end<G>(K, ng, chemonly, s, e, fscale, nhplus, eint_code, temp);
  return 0;
}

// the Lax-Friedrichs face flux of lf_flux with contraction allowed (mixed path)
__device__ __forceinline__ void lf_flux_fast(int d, int s, float Ue, float Ux, float Uy, float Uz,
                                             const Closure<float>& qu, float Ve, float Vx, float Vy, float Vz,
                                             const Closure<float>& qv, float hc, float F[4]) {
#pragma clang fp contract(fast)
  const float Me = s ? Ue : Ve, Pe = s ? Ve : Ue;
  const float Mf[3] = {s ? Ux : Vx, s ? Uy : Vy, s ? Uz : Vz};
  const float Pf[3] = {s ? Vx : Ux, s ? Vy : Uy, s ? Vz : Uz};
  const Closure<float>& qm = s ? qu : qv;
  const Closure<float>& qp = s ? qv : qu;
  F[0] = 0.5f * (Pf[d] + Mf[d]) + hc * (Me - Pe);
  const float wp = qp.arg * qp.n[d], wm = qm.arg * qm.n[d];
#pragma unroll
  for (int i = 0; i < 3; i++) {
    float pp = wp * qp.n[i], pm = wm * qm.n[i];
    if (i == d) {
      pp += qp.diag;
      pm += qm.diag;
    }
    F[1 + i] = 0.5f * (pp + pm) + hc * (Mf[i] - Pf[i]);
  }
}

// Twice one side's share of the Lax-Friedrichs face flux in direction D.
// With P(U) = (F_D, P_D0, P_D1, P_D2) the physical flux (P_Di = diag
// delta_Di + arg n_D n_i) and c the wave speed, the face flux of lf_flux is
// (A(M) + B(P)) / 2 where A = P + c U for the state on the minus side and
// B = P - c U for the state on the plus side.  side = +1 gives A, -1 gives B.
template <int D>
__device__ __forceinline__ void lf_share(float side, float Ue, float Ux, float Uy, float Uz, const Closure<float>& q,
                                         float c, float X[4]) {
#pragma clang fp contract(fast)
  const float Uf[3] = {Ux, Uy, Uz};
  const float sc = side * c;
  X[0] = fmaf(sc, Ue, Uf[D]);
  const float w = q.arg * q.n[D];
#pragma unroll
  for (int i = 0; i < 3; i++) {
    const float h = (i == D) ? fmaf(w, q.n[i], q.diag) : w * q.n[i];
    X[1 + i] = fmaf(sc, Uf[i], h);
  }
}

// d uniform and known after unrolling: the branches fold away
__device__ __forceinline__ void lf_share_unrolled(int d, float side, float Ue, float Ux, float Uy, float Uz,
                                                  const Closure<float>& q, float c, float X[4]) {
  if (d == 0)
    lf_share<0>(side, Ue, Ux, Uy, Uz, q, c, X);
  else if (d == 1)
    lf_share<1>(side, Ue, Ux, Uy, Uz, q, c, X);
  else
    lf_share<2>(side, Ue, Ux, Uy, Uz, q, c, X);
}

// the same with a direction that differs between lanes: components are
// picked with 0/1 masks, not by indexing (a lane-varying index, also one the
// optimiser rebuilds from selects, puts the arrays in scratch memory); all
// terms are finite
__device__ __forceinline__ void lf_share_lane(int d, float side, float Ue, float Ux, float Uy, float Uz,
                                              const Closure<float>& q, float c, float X[4]) {
#pragma clang fp contract(fast)
  const float sc = side * c;
  const float m0 = d == 0 ? 1.f : 0.f, m1 = d == 1 ? 1.f : 0.f, m2 = d == 2 ? 1.f : 0.f;
  const float Fd = fmaf(Ux, m0, fmaf(Uy, m1, Uz * m2));
  const float nd = fmaf(q.n[0], m0, fmaf(q.n[1], m1, q.n[2] * m2));
  X[0] = fmaf(sc, Ue, Fd);
  const float w = q.arg * nd;
  const float Uf[3] = {Ux, Uy, Uz};
  const float mi[3] = {m0, m1, m2};
#pragma unroll
  for (int i = 0; i < 3; i++) X[1 + i] = fmaf(sc, Uf[i], fmaf(q.diag, mi[i], w * q.n[i]));
}

}  // namespace emma
