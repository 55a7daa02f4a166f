// rt_kernels.hip -- gfx950 kernels of the resident-level radiation step.
//
// One level step (the work of advancerad + RadSolver's final update for one
// level, rad_utils.c:1775-2006) is ONE kernel: a thread per cell reads its
// own state and its six face neighbours straight from the level's SoA arrays
// through the oct index tables (no stencil copy), forms the Lax-Friedrichs
// M1 fluxes, updates, runs the chemistry sub-stepper in FP64 and writes the
// next state; split cells write the mean of their children instead.  State is
// double-buffered (read A, write B), so the reference's rfieldnew copy
// (clean_new_rad) disappears; coarse-fine refluxes go to small per-level
// accumulators that the coarser level folds in on its next step.
#include <hip/hip_runtime.h>

#include "rt_kernels.h"
#include "rt_physics_fast.h"

namespace emma {

// ------------------------------------------------------------------ helpers
__device__ __forceinline__ int xcd_block(int nb) {
  // blocks are dealt round-robin to the 8 XCDs; give each XCD a contiguous
  // range of cells so that neighbour octs share its L2
  const int b = blockIdx.x;
  const int per = nb >> 3;
  return (b & 7) * per + (b >> 3);
}

__device__ __forceinline__ void raise_status(int* status, int code, int level, int cell) {
  if (atomicCAS(status, 0, code) == 0) {
    status[1] = level;
    status[2] = cell;
  }
}

template <typename R>
__device__ __forceinline__ void load4(const LevelArgs<R>& a, const R* e, const R* fx, const R* fy, const R* fz,
                                      int n, int j, int g, R& oe, R& ofx, R& ofy, R& ofz) {
  oe = e[g * n + j];
  ofx = fx[g * n + j];
  ofy = fy[g * n + j];
  ofz = fz[g * n + j];
}

template <typename T>
__device__ __forceinline__ T sel(int i, T a, T b, T c) {
  return i == 0 ? a : (i == 1 ? b : c);
}

// ------------------------------------------------------------------ level step
// Fast selects the mixed-precision physics (rt_physics_fast.h); the FP64
// kernel (Fast = false) keeps the reference's operation order.
template <typename R, int G, bool Fast>
__global__ __launch_bounds__(256) void level_step_kernel(LevelArgs<R> a, ChemConst K) {
  constexpr int GA = G > 0 ? G : kMaxGrp;
  const int ng = G > 0 ? G : a.ng;
  const int cell = xcd_block(a.nblocks) * 256 + (int)threadIdx.x;
  unsigned int nit_acc = 0;
  unsigned int leaf = 0;
  if (cell < a.ncell) {
    const int n = a.ncell;
    const int oct = cell >> 3, ic = cell & 7;
    const int ch = a.child[cell];

    if (a.t_first != nullptr) {  // rad_utils.c:1901-1913
      const double xion = a.nhplus[cell] / a.nh[cell];
      if (xion < 0.5) a.t_last[cell] = a.tphy;
      if ((xion >= 0.5) && (a.t_first[cell] == -1)) a.t_first[cell] = a.tphy;
    }

    if (ch >= 0) {
      // split cell: mean of the 8 children (rad_utils.c:1955-2006)
      const int nf = a.nfine;
      const int j0 = ch * 8;
      for (int g = 0; g < ng; g++) {
        R se = R(0), sx = R(0), sy = R(0), sz = R(0), ss = R(0);
        for (int k = 0; k < 8; k++) {
          const int j = g * nf + j0 + k;
          se += a.fe[j] * R(0.125);
          sx += a.ffx[j] * R(0.125);
          sy += a.ffy[j] * R(0.125);
          sz += a.ffz[j] * R(0.125);
          ss += a.fsrc[j] * R(0.125);
        }
        a.eo[g * n + cell] = se;
        a.fxo[g * n + cell] = sx;
        a.fyo[g * n + cell] = sy;
        a.fzo[g * n + cell] = sz;
        a.src[g * n + cell] = ss;
      }
      double snp = 0., sei = 0., stp = 0., snh = 0., stf = 0., stl = 0.;
      for (int k = 0; k < 8; k++) {
        const int j = j0 + k;
        if (a.t_first != nullptr) {
          stf += a.ft_first[j] * 0.125;
          stl += a.ft_last[j] * 0.125;
        }
        snp += a.fnhplus[j] * 0.125;
        sei += a.feint[j] * 0.125;
        stp += a.ftemp[j] * 0.125;
        snh += a.fnh[j] * 0.125;
      }
      a.nhpluso[cell] = snp;
      a.einto[cell] = sei;
      a.tempo[cell] = stp;
      a.nh[cell] = snh;
      if (a.t_first != nullptr) {
        a.t_first[cell] = stf;
        a.t_last[cell] = stl;
      }
    } else {
      leaf = 1;
      const R c = a.c, hc = a.hc, dts = a.dtsurdx;
      // own state and its closure, per group
      R Ue[GA], Ux[GA], Uy[GA], Uz[GA];
      R De[GA], Dx[GA], Dy[GA], Dz[GA];
      Closure<R> Q[GA];
#pragma unroll
      for (int g = 0; g < GA; g++) {
        if (g < ng) {
          Ue[g] = a.e[g * n + cell];
          Ux[g] = a.fx[g * n + cell];
          Uy[g] = a.fy[g * n + cell];
          Uz[g] = a.fz[g * n + cell];
          if constexpr (Fast)
            Q[g] = closure_fast(Ux[g], Uy[g], Uz[g], Ue[g], c);
          else
            Q[g] = closure<R>(Ux[g], Uy[g], Uz[g], Ue[g], c);
          De[g] = R(0);
          Dx[g] = R(0);
          Dy[g] = R(0);
          Dz[g] = R(0);
        }
      }
      // six faces in the reference order (rad_utils.c:1566-1574): f = 2d+s
#pragma unroll
      for (int f = 0; f < 6; f++) {
        const int d = f >> 1, s = f & 1;
        const int bit = (ic >> d) & 1;
        const int nc = ic ^ (1 << d);
        const R* Ne = a.e;
        const R* Nx = a.fx;
        const R* Ny = a.fy;
        const R* Nz = a.fz;
        int nn = n, j, split;
        int coarse_face = 0;
        if (bit != s) {  // neighbour inside the oct
          j = oct * 8 + nc;
          split = a.child[j] >= 0;
        } else {
          const int no = a.nei[oct * 6 + f];
          if (no >= 0) {
            j = no * 8 + nc;
            split = a.child[j] >= 0;
          } else {  // coarse leaf injected into the stencil
            j = -1 - no;
            split = 0;
            Ne = a.ce;
            Nx = a.cfx;
            Ny = a.cfy;
            Nz = a.cfz;
            nn = a.ncoarse;
            coarse_face = 1;
          }
        }
        const R ffact = split ? R(0) : R(1);
        const R one = (f & 1) ? R(-1) : R(1);
#pragma unroll
        for (int g = 0; g < GA; g++) {
          if (g < ng) {
            R ve, vx, vy, vz;
            load4(a, Ne, Nx, Ny, Nz, nn, j, g, ve, vx, vy, vz);
            Closure<R> qn;
            if constexpr (Fast)
              qn = closure_fast(vx, vy, vz, ve, c);
            else
              qn = closure<R>(vx, vy, vz, ve, c);
            R F[4];
            lf_flux<R>(d, s, Ue[g], Ux[g], Uy[g], Uz[g], Q[g], ve, vx, vy, vz, qn, hc, F);
#pragma unroll
            for (int v = 0; v < 4; v++) F[v] *= ffact;
            De[g] += F[0] * dts * one;
            Dx[g] += F[1] * dts * one;
            Dy[g] += F[2] * dts * one;
            Dz[g] += F[3] * dts * one;
            if (coarse_face) {  // scatterstencilrad refluxing (rad_utils.c:1738-1762)
              const int cch = a.cchild[j];
              if (cch <= -2) {
                R* acc = a.cracc + (size_t)(-2 - cch) * 4 * GA;
                const R w = -one;  // pow(-1, face+1)
                atomicAdd(acc + 0 * GA + g, F[0] * dts * w * R(0.125));
                atomicAdd(acc + 1 * GA + g, F[1] * dts * w * R(0.125));
                atomicAdd(acc + 2 * GA + g, F[2] * dts * w * R(0.125));
                atomicAdd(acc + 3 * GA + g, F[3] * dts * w * R(0.125));
              }
            }
          }
        }
      }
      // base state: rfieldnew = rfield (+ refluxes from the finer level)
      if (ch <= -2) {
        R* acc = a.racc + (size_t)(-2 - ch) * 4 * GA;
#pragma unroll
        for (int g = 0; g < GA; g++) {
          if (g < ng) {
            Ue[g] += acc[0 * GA + g];
            Ux[g] += acc[1 * GA + g];
            Uy[g] += acc[2 * GA + g];
            Uz[g] += acc[3 * GA + g];
            acc[0 * GA + g] = R(0);
            acc[1 * GA + g] = R(0);
            acc[2 * GA + g] = R(0);
            acc[3 * GA + g] = R(0);
          }
        }
      }
      double e[GA], fx[GA], fy[GA], fz[GA], src[GA];
#pragma unroll
      for (int g = 0; g < GA; g++) {
        if (g < ng) {
          const R ue = Ue[g] + De[g];
          const R ux = Ux[g] + Dx[g];
          e[g] = (double)ue;
          fx[g] = (double)ux;
          fy[g] = (double)(Uy[g] + Dy[g]);
          fz[g] = (double)(Uz[g] + Dz[g]);
          src[g] = (double)a.src[g * n + cell];
        }
      }
      // rad_utils.c:1611-1650 (the second test reads e[1], i.e. fx[0] when G=1)
      if (e[0] < 0 || isnan(e[0]) || isnan(ng > 1 ? e[1 < GA ? 1 : 0] : fx[0])) {
        raise_status(a.status, -3, a.level, cell);
      }
      double nhplus = a.nhplus[cell];
      double eint = a.eint[cell];
      const double nh = a.nh[cell];
      double temp = 0.;
      int nit = 0;
      int st;
      if constexpr (Fast)
        st = chem_cell_fast<G>(K, ng, e, fx, fy, fz, src, nhplus, eint, nh, temp, a.chemonly, nit);
      else
        st = chem_cell<G>(K, ng, e, fx, fy, fz, src, nhplus, eint, nh, temp, a.chemonly, nit);
      if (st != 0) raise_status(a.status, st, a.level, cell);
      nit_acc = (unsigned)nit;
#pragma unroll
      for (int g = 0; g < GA; g++) {
        if (g < ng) {
          a.eo[g * n + cell] = (R)e[g];
          a.fxo[g * n + cell] = (R)fx[g];
          a.fyo[g * n + cell] = (R)fy[g];
          a.fzo[g * n + cell] = (R)fz[g];
        }
      }
      a.nhpluso[cell] = nhplus;
      a.einto[cell] = eint;
      a.tempo[cell] = temp;
    }
  }
  // per-block counters: accepted chemistry sub-steps and leaf cell-updates
  if (a.counters != nullptr) {
    __shared__ unsigned int red[2][4];
    unsigned int v0 = nit_acc, v1 = leaf;
    for (int o = 32; o > 0; o >>= 1) {
      v0 += __shfl_xor(v0, o, 64);
      v1 += __shfl_xor(v1, o, 64);
    }
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
      red[0][w] = v0;
      red[1][w] = v1;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      unsigned long long* slot = a.counters + (size_t)(blockIdx.x % kCounterSlots) * kCounterStride;
      atomicAdd(slot + 0, (unsigned long long)(red[0][0] + red[0][1] + red[0][2] + red[0][3]));
      atomicAdd(slot + 1, (unsigned long long)(red[1][0] + red[1][1] + red[1][2] + red[1][3]));
    }
  }
}

// ------------------------------------------------------------------ photon budget
// FP64 sum over leaf cells of sum_g N_g * dx^3 (wavefront shuffles, one
// atomic per block)
template <typename R>
__global__ __launch_bounds__(256) void photon_sum_kernel(const R* e, const int* child, int ncell, int ng, double vol,
                                                         double* out) {
  double s = 0.;
  for (int i = blockIdx.x * 256 + threadIdx.x; i < ncell; i += gridDim.x * 256) {
    if (child[i] < 0)
      for (int g = 0; g < ng; g++) s += (double)e[g * ncell + i];
  }
  s *= vol;
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
  __shared__ double red[4];
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) atomicAdd(out, red[0] + red[1] + red[2] + red[3]);
}

// small fill kernel (used instead of hipMemsetAsync so that the work can sit
// inside a captured graph)
__global__ void fill_kernel(unsigned int* p, int n, unsigned int v) {
  int i = blockIdx.x * 256 + threadIdx.x;
  if (i < n) p[i] = v;
}

// ------------------------------------------------------------------ launchers
template <typename R>
hipError_t launch_level_step(const LevelArgs<R>& a0, const ChemConst& K, hipStream_t s) {
  LevelArgs<R> a = a0;
  int nb = (a.ncell + 255) / 256;
  nb = (nb + 7) & ~7;
  a.nblocks = nb;
  const dim3 grid(nb), block(256);
  constexpr bool fast = sizeof(R) == 4;
  switch (a.ng) {
    case 1: hipLaunchKernelGGL((level_step_kernel<R, 1, fast>), grid, block, 0, s, a, K); break;
    case 3: hipLaunchKernelGGL((level_step_kernel<R, 3, fast>), grid, block, 0, s, a, K); break;
    default: hipLaunchKernelGGL((level_step_kernel<R, 0, fast>), grid, block, 0, s, a, K); break;
  }
  return hipGetLastError();
}

template <typename R>
hipError_t launch_photon_sum(const R* e, const int* child, int ncell, int ng, double vol, double* out,
                             hipStream_t s) {
  int nb = (ncell + 255) / 256;
  if (nb > 2048) nb = 2048;
  if (nb < 1) nb = 1;
  hipLaunchKernelGGL((photon_sum_kernel<R>), dim3(nb), dim3(256), 0, s, e, child, ncell, ng, vol, out);
  return hipGetLastError();
}

hipError_t launch_fill(unsigned int* p, int n, unsigned int v, hipStream_t s) {
  hipLaunchKernelGGL(fill_kernel, dim3((n + 255) / 256), dim3(256), 0, s, p, n, v);
  return hipGetLastError();
}

template hipError_t launch_level_step<double>(const LevelArgs<double>&, const ChemConst&, hipStream_t);
template hipError_t launch_level_step<float>(const LevelArgs<float>&, const ChemConst&, hipStream_t);
template hipError_t launch_photon_sum<double>(const double*, const int*, int, int, double, double*, hipStream_t);
template hipError_t launch_photon_sum<float>(const float*, const int*, int, int, double, double*, hipStream_t);

}  // namespace emma
