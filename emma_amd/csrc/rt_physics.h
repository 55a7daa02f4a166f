// rt_physics.h -- device physics of the M1 + hydrogen-chemistry step (gfx950).
//
// Operation order follows the reference CPU functions cited at each function
// so that the FP64 path rounds like them; the translation units that include
// this header are compiled with -ffp-contract=off (no fused multiply-add),
// matching the unfused x86-64 reference.
#pragma once

#include <hip/hip_runtime.h>
#include <math.h>

namespace emma {

constexpr int kMaxGrp = 9;

// param.h:4-37,52-56
constexpr double kGamma = 5. / 3.;
constexpr double kFracVar = 0.25;
constexpr double kPmin = 1e-10;
constexpr double kEmin = 1e-8;
constexpr double kKboltz = 1.3806e-23;
constexpr double kParsec = 3.085677e16;
constexpr double kYHe = 0.;
constexpr double kyHe = kYHe / (1. - kYHe) / 4.002;
constexpr double kHnu0 = 13.6 * 1.6022e-19;  // chem_utils.c:196
constexpr int kMaxTries = 1 << 20;           // sub-step tries per cell and call

// Per-call constants of chemrad (chem_utils.c:266-331) that depend only on
// the run parameters and the expansion factor.  Computed once on the host
// (glibc pow), which is what the reference evaluates per cell.
struct ChemConst {
  int ngrp;
  int ncvgcool;
  int static_cosmo;  // H0 == 0: a stays fixed inside the sub-step loop
  int pad;
  double fudgecool;
  double unit_l, unit_t, unit_N, unit_n, unit_d, unit_v;
  double aexp, a2, a3;                 // aexp, aexp*aexp, aexp*aexp*aexp
  double pa2, pa3, pa4, pa5;           // pow(aexp, k)
  double pv2;                          // pow(unit_v, 2)
  double pstar;                        // unit_n*unit_d*pow(unit_v,2) (E2T)
  double dt;                           // dtnew*unit_t*pow(aexp,2)
  double emin;                         // PMIN/(GAMMA-1)/pow(aexp,5)*unit_n*unit_d*pow(unit_v,2)
  double H0, om, ov;
  double heat[kMaxGrp];                // alphae*hnu - alphai*hnu0
  double alphai[kMaxGrp];
  double factgrp[kMaxGrp];
  double emin_grp[kMaxGrp];            // EMIN*factgrp/unit_N
  float emin_grp_up[kMaxGrp];          // the smallest float >= emin_grp (mixed-path storage)
  // products of the conversion factors above, for the mixed-precision path
  // (which multiplies where the FP64 path divides)
  double k_e;      // unit_N / a^3            code N -> photons/m^3
  double k_f;      // unit_l/unit_t*unit_N/a^4
  double k_src;    // unit_N/unit_t/a^2/a^3
  double k_eint;   // unit_n*unit_d*unit_v^2/a^5
  double ki_e;     // a^3/unit_N
  double ki_f;     // a^4/unit_l*unit_t/unit_N
  double ki_eint;  // a^5/(unit_n*unit_d*unit_v^2)
};

// ---------------------------------------------------------------- Eddington
// Levermore M1 closure, rad_utils.c:448-475.  The reference recomputes the
// closure at every face; the per-state part (c^2 E, unit vector, chi) is a
// pure function of the state, so it is computed once and each tensor
// component P_ij is formed exactly as the reference forms it.
template <typename R>
struct Closure {
  R diag;  // (1-chi)/2 * c^2 E
  R arg;   // (3 chi - 1)/2 * c^2 E
  R n[3];
};

template <typename R>
__device__ __forceinline__ Closure<R> closure(R fx, R fy, R fz, R ee, R c) {
  Closure<R> q;
  R c2e = ee * c * c;
  R ff = R(0);
  q.n[0] = R(0);
  q.n[1] = R(0);
  q.n[2] = R(0);
  if (ee > R(0)) {
    ff = sqrt(fx * fx + fy * fy + fz * fz);
    if (ff > R(0)) {
      q.n[0] = fx / ff;
      q.n[1] = fy / ff;
      q.n[2] = fz / ff;
    }
    ff = ff / (c * ee);
  }
  R arg = fmax(R(4) - R(3) * ff * ff, R(0));
  R chi = (R(3) + R(4) * ff * ff) / (R(5) + R(2) * sqrt(arg));
  q.diag = (R(1) - chi) / R(2) * c2e;
  q.arg = (R(3) * chi - R(1)) / R(2) * c2e;
  return q;
}

template <typename R>
__device__ __forceinline__ R eddington_ij(const Closure<R>& q, int i, int j) {
  R res = (i == j) ? q.diag : R(0);
  res += q.arg * q.n[i] * q.n[j];
  return res;
}

// Lax-Friedrichs M1 flux through face f = 2d+s of a cell (rad_sweepX/Y/Z,
// rad_utils.c:478-1060): the cell state U with closure qu, the neighbour
// state V with closure qv; the minus side of the face is V for s=0 and U for
// s=1.  Flux in +d of (E, Fx, Fy, Fz): 0.5(fp+fm) + 0.5c(um-up), p = plus
// side, m = minus side, with f(E) = F_d and f(F_i) = P_id.
template <typename R>
__device__ __forceinline__ void lf_flux(int d, int s, R Ue, R Ux, R Uy, R Uz, const Closure<R>& qu, R Ve, R Vx,
                                        R Vy, R Vz, const Closure<R>& qv, R hc, R F[4]) {
  const R Me = s ? Ue : Ve, Pe = s ? Ve : Ue;
  const R Mf[3] = {s ? Ux : Vx, s ? Uy : Vy, s ? Uz : Vz};
  const R Pf[3] = {s ? Vx : Ux, s ? Vy : Uy, s ? Vz : Uz};
  const Closure<R>& qm = s ? qu : qv;
  const Closure<R>& qp = s ? qv : qu;
  F[0] = R(0.5) * (Pf[d] + Mf[d]) + hc * (Me - Pe);
#pragma unroll
  for (int i = 0; i < 3; i++) {
    const R fp = eddington_ij(qp, i, d);
    const R fm = eddington_ij(qm, i, d);
    F[1 + i] = R(0.5) * (fp + fm) + hc * (Mf[i] - Pf[i]);
  }
}

// ---------------------------------------------------------------- rates
// T-dependent factors of cuCompCooling (chem_utils.c:92-187 with OTSA and
// WRADTEST, CLUMPF = 1, yHE = 0), cucompute_alpha_b (37-51) and
// cucompute_beta (75-88), split so that the x- and n_H-dependent products are
// applied in the reference's left-to-right order.  They depend on T only, so
// a rejected sub-step (same T) reuses them.
struct Rates {
  double A1;      // exp(-157809.1/T)*2.54e-21*sqrt(T)/(1+sqrt(T/1e5))   collisional ionisation
  double A3;      // exp(-118348/T)*7.5e-19/(1+sqrt(T/1e5))             collisional excitation
  double A4;      // 1.42e-27*1.5*sqrt(T)                                  bremsstrahlung
  double A6;      // case-B recombination cooling
  double alphab;  // case-B recombination rate (m^3/s)
  double beta;    // collisional ionisation rate (m^3/s)
};

__device__ __forceinline__ Rates rates_of(double temp) {
  Rates r;
  const double sT = sqrt(temp);
  const double sT5 = sqrt(temp / 1e5);
  const double lam = 2e0 * 157807e0 / temp;
  r.A1 = exp(-157809.1e0 / temp) * 2.54e-21 * sT / (1. + sT5);
  r.A6 = 3.435e-30 * temp * pow(lam, 1.970e0) / pow(1. + (pow(lam / 2.250e0, 0.376e0)), 3.720e0);
  r.A3 = exp(-118348e0 / temp) * 7.5e-19 / (1 + sT5);
  r.A4 = 1.42e-27 * 1.5e0 * sT;
  double ab = 2.753e-14 * pow(lam, 1.5) / pow(1e0 + pow(lam / 2.740, 0.407), 2.242);
  r.alphab = ab * 1e-6 * 1.;
  double b = 1.17e-10 * sT / (1 + sT5) * exp(-(157809e0 / temp));
  r.beta = b * 1e-6 * 1.;
  return r;
}

// total cooling rate Lambda(T, x, n_H) in J m^-3 s^-1 (chem_utils.c:108-150,166)
__device__ __forceinline__ double cooling(const Rates& r, double x, double nH) {
  const double nh2 = nH * 1e-6;
  const double c1 = r.A1 * x * (1. - x) * nh2 * nh2 * 1. * (1 + kyHe);
  const double c6 = r.A6 * x * x * nh2 * nh2 * 1.;
  const double c3 = r.A3 * x * (1. - x) * nh2 * nh2 * 1. * (1 + kyHe);
  const double c4 = r.A4 * x * (1. + kyHe) * nh2 * x * nh2 * (1. + kyHe) * 1.;
  double lambda = c1 + 0. + c3 + c4 + 0. + c6;
  return lambda * 1e-7 * 1e6;
}

// ---------------------------------------------------------------- chemistry
// chemrad for one unsplit cell (chem_utils.c:266-791): semi-implicit
// photon absorption, ionisation and heating/cooling with adaptive sub-steps.
// In/out are the code-unit Rtype fields after transport.  Returns 0 or a
// negative status (EMMA_RT_ERR_EMIN, EMMA_RT_ERR_NOCONV); `nit` counts the
// accepted sub-steps.
template <int G>
__device__ int chem_cell(const ChemConst& K, int ng, double* e, double* fx, double* fy, double* fz,
                         const double* src, double& nhplus, double& eint_code, double nh, double& temp,
                         int chemonly, int& nit) {
  constexpr int GA = G > 0 ? G : kMaxGrp;
  double egy[GA], fl0[GA], fl1[GA], fl2[GA], srcl[GA], et[GA], p[GA];
  const double aexporg = K.aexp;
  const double a3 = K.a3;
  const double notchem = (double)(!chemonly);
#pragma unroll
  for (int g = 0; g < GA; g++) {
    if (g < ng) {
      egy[g] = e[g] / a3 * K.unit_N;
      fl0[g] = fx[g] / K.pa4 * K.unit_l / K.unit_t * K.unit_N;
      fl1[g] = fy[g] / K.pa4 * K.unit_l / K.unit_t * K.unit_N;
      fl2[g] = fz[g] / K.pa4 * K.unit_l / K.unit_t * K.unit_N;
      srcl[g] = (src[g] * K.unit_N / K.unit_t / K.a2) / K.pa3;
    }
  }
  double x0 = nhplus / nh;
  const double nH = nh / a3 * K.unit_N;
  double eint = eint_code / K.pa5 * K.unit_n * K.unit_d * K.pv2;
  const double emin = K.emin;
  nit = 0;
  if (egy[0] < kEmin) return -4;

  double aexp = aexporg;
  double fudgecool = K.fudgecool;
  double tcur = 0.;
  const double dt = K.dt;
  double tcached = -1.;
  Rates r = rates_of(1e4);
  int tries = 0;
  while (tcur < dt) {
    // the reference loops until t >= dt; a try budget keeps every wave finite
    // on pathological input (reported as non-convergence)
    if (++tries > kMaxTries) return -5;
    const double Nfree = (1. + x0) * nH;
    const double tloc = eint / (1.5 * Nfree * kKboltz);
    if (!(tloc == tcached)) {  // rates depend on T only, which changes only after an accepted sub-step
      r = rates_of(tloc);
      tcached = tloc;
    }
    const double Cool0 = cooling(r, x0, nH);
    if (fudgecool < 1e-20) return -5;
    double ai = 0.;
#pragma unroll
    for (int g = 0; g < GA; g++)
      if (g < ng) ai += K.heat[g] * egy[g] * nH * (1.0 - x0);
    const double tcool = fabs(eint / (ai * notchem - Cool0));
    const double dtcool = fmin(fudgecool * tcool, dt - tcur);
    const double alpha = r.alphab;  // OTSA (chem_utils.c:497-500)
    int test = 0;
#pragma unroll
    for (int g = 0; g < GA; g++) {
      if (g < ng) {
        const double aig = K.alphai[g] * (1. - x0);
        if (chemonly) {
          et[g] = egy[g];
        } else {
          et[g] = ((alpha - r.alphab) * x0 * x0 * nH * nH * dtcool * 0. + egy[g] + srcl[g] * dtcool * K.factgrp[g]) /
                  (1. + dtcool * (aig * nH));
        }
        if ((et[g] < 0) || isnan(et[g])) test = 1;
        p[g] = (1. + (aig * nH) * dtcool);
      }
    }
    if (test) {
      fudgecool = fudgecool / 10.;
      continue;
    }
    ai = 0.;
#pragma unroll
    for (int g = 0; g < GA; g++)
      if (g < ng) ai += K.alphai[g] * et[g] * notchem;
    const double RECT = alpha * x0 * x0 * nH * dtcool;
    const double xt = 1. - (RECT + (1.0 - x0)) / (1. + dtcool * (r.beta * x0 * nH + ai));
    if (((xt > 1.) || (xt < 0.)) || isnan(xt)) {
      fudgecool /= 10.;
      continue;
    }
    const double Cool1 = cooling(r, xt, nH);
    ai = 0.;
#pragma unroll
    for (int g = 0; g < GA; g++)
      if (g < ng) ai += K.heat[g] * et[g] * nH * (1.0 - xt);
    double eintt = (eint + dtcool * ((ai * notchem) - Cool1));
    if (eintt < 0.) {
      fudgecool = fudgecool / 10.;
      continue;
    }
    const double deltaE = fabs(eintt / eint - 1.);
    if (deltaE > kFracVar) {
      fudgecool = fudgecool / 10.;
      continue;
    }
    eintt = fmax(emin, eintt);
    if (K.static_cosmo) {
#pragma unroll
      for (int g = 0; g < GA; g++) {
        if (g < ng) {
          egy[g] = et[g];
          if (!chemonly) {
            fl0[g] = fl0[g] / p[g];
            fl1[g] = fl1[g] / p[g];
            fl2[g] = fl2[g] / p[g];
          }
        }
      }
      eint = eintt;
    } else {  // chem_utils.c:737-756
      const double aold = aexp;
      const double hubblet = K.H0 * sqrt(K.om / aexp + K.ov * (aexp * aexp)) / aexp * (1e3 / (1e6 * kParsec));
      const double da = hubblet * dtcool * aexp;
      aexp += da;
      const double r3 = pow(aold / aexp, 3), r4 = pow(aold / aexp, 4), r5 = pow(aold / aexp, 5);
#pragma unroll
      for (int g = 0; g < GA; g++) {
        if (g < ng) {
          egy[g] = et[g] * r3;
          if (!chemonly) {
            fl0[g] = fl0[g] / p[g] * r4;
            fl1[g] = fl1[g] / p[g] * r4;
            fl2[g] = fl2[g] / p[g] * r4;
          }
        }
      }
      eint = eintt * r5;
    }
    x0 = xt;
    tcur += dtcool;
    fudgecool = K.fudgecool;
    nit++;
    if ((nit == K.ncvgcool) && (K.ncvgcool != 0)) break;
  }
  // back to code units (chem_utils.c:775-791) and E2T (21-32)
  double pa3f, pa4f, pa5f;
  if (K.static_cosmo) {
    pa3f = K.pa3;
    pa4f = K.pa4;
    pa5f = K.pa5;
  } else {
    pa3f = pow(aexp, 3);
    pa4f = pow(aexp, 4);
    pa5f = pow(aexp, 5);
  }
  if (!chemonly) {
#pragma unroll
    for (int g = 0; g < GA; g++) {
      if (g < ng) {
        e[g] = fmax(egy[g] * pa3f / K.unit_N, K.emin_grp[g]);
        fx[g] = fl0[g] * pa4f / K.unit_l * K.unit_t / K.unit_N;
        fy[g] = fl1[g] * pa4f / K.unit_l * K.unit_t / K.unit_N;
        fz[g] = fl2[g] * pa4f / K.unit_l * K.unit_t / K.unit_N;
      }
    }
  }
  nhplus = x0 * nh;
  eint_code = eint * pa5f / K.unit_n / K.unit_d / K.pv2;
  {
    const double x = nhplus / nh;
    const double nHe = nh / pa3f * K.unit_N;
    const double ee = eint_code / pa5f * K.pstar;
    temp = ee / (1.5 * nHe * kKboltz * (1. + x) * (1. + kyHe));
  }
  return 0;
}

}  // namespace emma
