// rt_fill.h -- the per-cell work either side of a level step that couples the
// radiation state to the hydro state and places the sources: FillRad's
// per-cell block (src_utils.c:692-849) and RadSolver's injection back into
// the hydro fields (rad_utils.c:1941-1951, 1993-2004).  FP64 throughout with
// the reference's operation order (every file that includes this one is
// compiled without contraction).  Used by the per-level kernels of
// rt_fill.hip and by the persistent coarse-step kernel.
#pragma once

#include <hip/hip_runtime.h>

#include "rt_kernels.h"

namespace emma {

constexpr double kGammaAd = 5. / 3.;  // GAMMA, param.h:4

// E2T (chem_utils.c:21-32); pow(aexp,3), pow(aexp,5) and the pressure unit
// come from the host
__device__ __forceinline__ double e2t(double nhplus, double eint, double nh, double pa3, double pa5, double pstar,
                                      double unit_N) {
  const double x = nhplus / nh;
  const double nH = nh / pa3 * unit_N;
  const double ei = eint / pa5 * pstar;
  return ei / (1.5 * nH * kKboltz * (1. + x) * (1. + kyHe));
}

// FillRad's block for one cell (cell < f.ncell)
template <typename R>
__device__ __forceinline__ void fill_cell(const FillArgs<R>& f, int cell) {
  const int n = f.ncell;
  double nh = f.nh[cell], eint = f.eint[cell], nhplus = f.nhplus[cell];
  if (f.hydro) {  // src_utils.c:738-752 (WRADHYD)
    nh = f.d[cell] * (1. - kYHe);
    eint = f.p[cell] / (kGammaAd - 1.);
    nhplus = f.dX[cell];
    f.nh[cell] = nh;
    f.eint[cell] = eint;
    f.nhplus[cell] = nhplus;
  }
  if (f.child[cell] >= 0) {  // src_utils.c:765-772: split cells carry no source
    for (int g = 0; g < f.ng; g++) f.src[g * n + cell] = R(0);
    return;
  }
  if (f.star) {  // putsource -> stromgren_source (src_utils.c:339-412, 446-466)
    const int oct = cell >> 3, ic = cell & 7;
    const double xc = f.opos[3 * oct + 0] + (ic & 1) * f.dxcur + f.dxcur * 0.5;
    const double yc = f.opos[3 * oct + 1] + ((ic >> 1) & 1) * f.dxcur + f.dxcur * 0.5;
    const double zc = f.opos[3 * oct + 2] + (ic >> 2) * f.dxcur + f.dxcur * 0.5;
    const double rx = xc - f.xs, ry = yc - f.ys, rz = zc - f.zs;
    const double Rd = sqrt(rx * rx + ry * ry + rz * rz);
    // putsource clears every group, stromgren_source sets the frequency
    // groups of the time group tcur lies in
    const bool on = (Rd <= f.X0) && (xc > 0.) && (yc > 0.) && (zc > 0.);
    for (int g = 0; g < f.ng; g++) f.src[g * n + cell] = (on && g >= f.star_g0 && g < f.star_g1) ? (R)f.star_src : R(0);
  }
  if (f.stars) {  // putsource -> stars_sources (src_utils.c:209-335, 446-520)
    double acc[kMaxGrp];
    for (int g = 0; g < f.ng; g++) acc[g] = 0.;
    const int k0 = f.star_first ? f.star_first[cell] : 0, k1 = f.star_first ? f.star_first[cell + 1] : 0;
    for (int k = k0; k < k1; k++) {
      const int is = f.star_is[k];
      if (is < 6 || is == 100) continue;
      const double age = f.stars_tphy - f.star_age[k];
      const double t = age / f.stars_tlife;
      const double src =
          f.star_mass[k] * f.stars_unit_d * f.stars_srcint / f.stars_dx3 * f.stars_unit_t / f.stars_unit_N * f.stars_pa2;
      double v;
      if (!f.stars_decrease || age < f.stars_tlife)
        v = src * (t > 0 ? 1. : 0.);
      else
        v = src * (t < 1. ? 1. : pow(t, -4.)) * (t > 0 ? 1. : 0);
      const int g0 = f.star_rs[k] * f.nspace;
      for (int g = g0; g < g0 + f.nspace && g < f.ng; g++) acc[g] += v;
    }
    for (int g = 0; g < f.ng; g++) f.src[g * n + cell] = (R)acc[g];
  }
  if (f.hydro) f.temp[cell] = e2t(nhplus, eint, nh, f.pa3, f.pa5, f.pstar, f.unit_N);  // src_utils.c:835-838
}

// RadSolver's injection for one cell (i < h.ncell)
__device__ __forceinline__ void to_hydro_cell(const HydroArgs& h, int i) {
  const double d = h.d[i], u = h.u[i], v = h.v[i], w = h.w[i];
  const double p = (kGammaAd - 1.) * h.eint[i];
  h.p[i] = p;
  h.dX[i] = h.nhplus[i] / h.nh[i] * d * (1. - kYHe);
  h.E[i] = p / (kGammaAd - 1.) + 0.5 * d * (u * u + v * v + w * w);  // getE, hydro_utils.c:29-31
}

}  // namespace emma
