// rt_stencil.hip -- struct RGRID batch kernel: one thread per cell of each
// gathered oct computes its six face fluxes (rad_sweepX/Y/Z,
// rad_utils.c:478-1060) into New.cell[c].rflux, applies updatefieldrad
// (1538-1655) to New.cell[c].rfieldnew and runs chemrad (chem_utils.c:190-796)
// on it.  The batch stays in the reference's byte layout, so the host's
// gatherstencilrad / scatterstencilrad keep working unchanged.
#include <hip/hip_runtime.h>

#include "rt_kernels.h"
#include "rt_stencil.h"

namespace emma {

bool rgrid_layout(int ngrp, int rb, RgridLayout* out) {
  if (ngrp < 1 || ngrp > kMaxGrp || (rb != 4 && rb != 8)) return false;
  const long G = ngrp;
  long* v = out->v;
  const long rtype = (5 * G + 4) * rb;
  v[0] = rtype;
  v[1] = 0;
  v[2] = G * rb;
  v[3] = 2 * G * rb;
  v[4] = 3 * G * rb;
  v[5] = 4 * G * rb;
  v[6] = 5 * G * rb;
  v[7] = (5 * G + 1) * rb;
  v[8] = (5 * G + 2) * rb;
  v[9] = (5 * G + 3) * rb;
  v[10] = rtype + rb;  // Rtype + char split, padded to REAL alignment
  v[11] = rtype;
  v[12] = 3 * rtype + 30 * G * rb;  // deltaR, rfield, rfieldnew, rflux[6*NGRP*NVAR_R]
  v[13] = 0;
  v[14] = rtype;
  v[15] = 2 * rtype;
  v[16] = 3 * rtype;
  v[17] = 8 * v[10];
  v[18] = 8 * v[12];
  v[19] = 7 * v[17] + v[18];
  v[20] = 0;
  v[21] = 7 * v[17];
  return true;
}

struct StencilGeo {
  long rgrid, oct, octlight, celllight, split, New, cellflux, rfieldnew, rflux;
};

// transport (condadvec): run the sweeps and the update (advancerad and
// advanceradGPU skip both on the finer levels of a COARSERAD build and for
// chemonly calls there, rad_utils.c:1808-1822, rad_utils_gpu.cu:1058-1075);
// keep_split_flux: fluxes of split neighbours kept (COARSERAD, rad_utils.c:
// 518-526); update_split: split cells transport-updated (COARSERAD,
// rad_utils.c:1548-1552).  Chemistry always skips split cells.
struct StencilMode {
  int transport, keep_split_flux, update_split;
};

template <typename R, int G>
__global__ __launch_bounds__(256) void stencil_kernel(char* base, StencilGeo geo, int nread, R c, R hc, R dts,
                                                      int chemonly, StencilMode mode, ChemConst K, int* status,
                                                      unsigned long long* counters) {
  const int t = blockIdx.x * 256 + threadIdx.x;
  if (t >= nread * 8) return;
  const int i = t >> 3, ic = t & 7;
  char* rg = base + (size_t)i * geo.rgrid;
  auto rt = [&](int slot, int cc) -> const R* {
    return (const R*)(rg + geo.oct + slot * geo.octlight + cc * geo.celllight);
  };
  auto split_of = [&](int slot, int cc) -> int {
    return *(const char*)(rg + geo.oct + slot * geo.octlight + cc * geo.celllight + geo.split) != 0;
  };
  char* newc = rg + geo.New + ic * geo.cellflux;
  R* rflux = (R*)(newc + geo.rflux);
  R* rn = (R*)(newc + geo.rfieldnew);
  const bool split = split_of(6, ic);
  double e[G], fx[G], fy[G], fz[G], src[G];
  if (mode.transport) {
    const R* C = rt(6, ic);
    R Ue[G], Ux[G], Uy[G], Uz[G], De[G], Dx[G], Dy[G], Dz[G];
    Closure<R> Q[G];
#pragma unroll
    for (int g = 0; g < G; g++) {
      Ue[g] = C[g];
      Ux[g] = C[G + g];
      Uy[g] = C[2 * G + g];
      Uz[g] = C[3 * G + g];
      Q[g] = closure<R>(Ux[g], Uy[g], Uz[g], Ue[g], c);
      De[g] = Dx[g] = Dy[g] = Dz[g] = R(0);
    }
#pragma unroll
    for (int f = 0; f < 6; f++) {
      const int d = f >> 1, s = f & 1;
      const int slot = (((ic >> d) & 1) != s) ? 6 : f;  // getcellnei, oct.c:116-185
      const int nc = ic ^ (1 << d);
      const R* N = rt(slot, nc);
      const R ffact = (!mode.keep_split_flux && split_of(slot, nc)) ? R(0) : R(1);
      const R one = (f & 1) ? R(-1) : R(1);
#pragma unroll
      for (int g = 0; g < G; g++) {
        const R ve = N[g], vx = N[G + g], vy = N[2 * G + g], vz = N[3 * G + g];
        const Closure<R> qn = closure<R>(vx, vy, vz, ve, c);
        R F[4];
        lf_flux<R>(d, s, Ue[g], Ux[g], Uy[g], Uz[g], Q[g], ve, vx, vy, vz, qn, hc, F);
#pragma unroll
        for (int v = 0; v < 4; v++) {
          F[v] *= ffact;
          rflux[f * 5 * G + g * 5 + v] = F[v];
        }
        rflux[f * 5 * G + g * 5 + 4] = R(0);
        De[g] += F[0] * dts * one;
        Dx[g] += F[1] * dts * one;
        Dy[g] += F[2] * dts * one;
        Dz[g] += F[3] * dts * one;
      }
    }
    if (split && !mode.update_split) return;  // updatefieldrad and chemrad skip split cells
#pragma unroll
    for (int g = 0; g < G; g++) {
      e[g] = (double)(R)(rn[g] + De[g]);
      fx[g] = (double)(R)(rn[G + g] + Dx[g]);
      fy[g] = (double)(R)(rn[2 * G + g] + Dy[g]);
      fz[g] = (double)(R)(rn[3 * G + g] + Dz[g]);
    }
    if (e[0] < 0 || isnan(e[0]) || isnan(G > 1 ? e[G > 1 ? 1 : 0] : fx[0])) {
      if (atomicCAS(status, 0, -3) == 0) {
        status[1] = -1;
        status[2] = t;
      }
    }
    if (split) {  // COARSERAD: the split cell keeps its transported state, no chemistry
#pragma unroll
      for (int g = 0; g < G; g++) {
        rn[g] = (R)e[g];
        rn[G + g] = (R)fx[g];
        rn[2 * G + g] = (R)fy[g];
        rn[3 * G + g] = (R)fz[g];
      }
      return;
    }
  } else {
    if (split) return;
#pragma unroll
    for (int g = 0; g < G; g++) {
      e[g] = (double)rn[g];
      fx[g] = (double)rn[G + g];
      fy[g] = (double)rn[2 * G + g];
      fz[g] = (double)rn[3 * G + g];
    }
  }
#pragma unroll
  for (int g = 0; g < G; g++) src[g] = (double)rn[4 * G + g];
  double nhplus = (double)rn[5 * G], eint = (double)rn[5 * G + 1];
  const double nh = (double)rn[5 * G + 2];
  double temp = 0.;
  int nit = 0;
  const int st = chem_cell<G>(K, G, e, fx, fy, fz, src, nhplus, eint, nh, temp, chemonly, nit);
  if (st != 0 && atomicCAS(status, 0, st) == 0) {
    status[1] = -1;
    status[2] = t;
  }
#pragma unroll
  for (int g = 0; g < G; g++) {
    rn[g] = (R)e[g];
    rn[G + g] = (R)fx[g];
    rn[2 * G + g] = (R)fy[g];
    rn[3 * G + g] = (R)fz[g];
  }
  rn[5 * G] = (R)nhplus;
  rn[5 * G + 1] = (R)eint;
  rn[5 * G + 3] = (R)temp;
  if (counters != nullptr) {
    unsigned long long* slot = counters + (size_t)(blockIdx.x % kCounterSlots) * kCounterStride;
    atomicAdd(slot + 0, (unsigned long long)nit);
    atomicAdd(slot + 1, 1ull);
  }
}

template <typename R, int G>
static hipError_t launch_g(void* d, const StencilGeo& geo, int nread, double c, double dts, int chemonly,
                           StencilMode m, const ChemConst& K, int* status, unsigned long long* counters,
                           hipStream_t s) {
  const int n = nread * 8;
  hipLaunchKernelGGL((stencil_kernel<R, G>), dim3((n + 255) / 256), dim3(256), 0, s, (char*)d, geo, nread, (R)c,
                     (R)(0.5 * c), (R)dts, chemonly, m, K, status, counters);
  return hipGetLastError();
}

template <typename R>
static hipError_t launch_r(void* d, const StencilGeo& geo, int ngrp, int nread, double c, double dts, int chemonly,
                           StencilMode m, const ChemConst& K, int* status, unsigned long long* counters,
                           hipStream_t s) {
  switch (ngrp) {
    case 1: return launch_g<R, 1>(d, geo, nread, c, dts, chemonly, m, K, status, counters, s);
    case 2: return launch_g<R, 2>(d, geo, nread, c, dts, chemonly, m, K, status, counters, s);
    case 3: return launch_g<R, 3>(d, geo, nread, c, dts, chemonly, m, K, status, counters, s);
    case 4: return launch_g<R, 4>(d, geo, nread, c, dts, chemonly, m, K, status, counters, s);
    case 5: return launch_g<R, 5>(d, geo, nread, c, dts, chemonly, m, K, status, counters, s);
    case 6: return launch_g<R, 6>(d, geo, nread, c, dts, chemonly, m, K, status, counters, s);
    case 7: return launch_g<R, 7>(d, geo, nread, c, dts, chemonly, m, K, status, counters, s);
    case 8: return launch_g<R, 8>(d, geo, nread, c, dts, chemonly, m, K, status, counters, s);
    case 9: return launch_g<R, 9>(d, geo, nread, c, dts, chemonly, m, K, status, counters, s);
    default: return hipErrorInvalidValue;
  }
}

hipError_t launch_stencil(void* d, const RgridLayout& L, int ngrp, int rb, int nread, double c, double dtsurdx,
                          int chemonly, int coarserad, int transport, const ChemConst& K, int* status,
                          unsigned long long* counters, hipStream_t s) {
  StencilMode m;
  m.transport = transport;
  m.keep_split_flux = coarserad;
  m.update_split = coarserad;
  StencilGeo geo;
  geo.rgrid = L.v[19];
  geo.oct = L.v[20];
  geo.octlight = L.v[17];
  geo.celllight = L.v[10];
  geo.split = L.v[11];
  geo.New = L.v[21];
  geo.cellflux = L.v[12];
  geo.rfieldnew = L.v[15];
  geo.rflux = L.v[16];
  if (rb == 8) return launch_r<double>(d, geo, ngrp, nread, c, dtsurdx, chemonly, m, K, status, counters, s);
  return launch_r<float>(d, geo, ngrp, nread, c, dtsurdx, chemonly, m, K, status, counters, s);
}

void stencil_free(void* p) {
  if (p) (void)hipFree(p);
}

}  // namespace emma
