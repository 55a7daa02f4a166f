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

#include <algorithm>

#include "rt_fill.h"
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

// The tile a block takes on a complete Morton-ordered cube of n = 2^m octs a
// side.  A tile is 32 Morton-consecutive octs, a 4 x 4 x 2 brick, so the
// tile index interleaves the tile coordinates (z, x, y from the lowest bit:
// tz0 tx0 ty0 tz1 ...; tz has one bit more).  Blocks are dealt round-robin
// to the 8 XCDs; XCD k takes octant k (the top bit of tx, ty, tz) and walks
// it z fastest, then y, then x: the neighbours across the largest faces
// (4 x 4 octs, +-z) are the next and previous tiles, the +-y ones n/4
// tiles away, within the ~128 tiles an XCD holds at once, so their halo
// cells are still in its L2.  In Morton order most of them were not.
__device__ __forceinline__ int raster_tile(int m) {
  const int b = blockIdx.x;
  const int k = b & 7, p = b >> 3;
  const int nz = 1 << (m - 2), ny = 1 << (m - 3);  // octant extent in tiles along z, y
  const int tz = (p & (nz - 1)) | ((k >> 2) << (m - 2));
  const int ty = ((p >> (m - 2)) & (ny - 1)) | (((k >> 1) & 1) << (m - 3));
  const int tx = (p >> (2 * m - 5)) | ((k & 1) << (m - 3));
  int t = 0;
#pragma unroll
  for (int i = 0; i < 14; i++) {
    if (i < m - 2) {
      t |= ((tz >> i) & 1) << (3 * i);
      t |= ((tx >> i) & 1) << (3 * i + 1);
      t |= ((ty >> i) & 1) << (3 * i + 2);
    }
  }
  t |= ((tz >> (m - 2)) & 1) << (3 * (m - 2));
  return t;
}

__device__ __forceinline__ void raise_status(int* status, int code, int level, int cell) {
  if (atomicCAS(status, 0, code) == 0) {
    status[1] = level;
    status[2] = cell;
  }
}

// lanes of mask m below this lane
__device__ __forceinline__ int lanes_below_mask(unsigned long long m) {
  return (int)__builtin_amdgcn_mbcnt_hi((unsigned)(m >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)m, 0u));
}

template <typename R>
__device__ __forceinline__ void load4(const LevelArgs<R>& a, const R* e, const R* fx, const R* fy, const R* fz,
                                      int n, int j, int g, R& oe, R& ofx, R& ofy, R& ofz) {
  oe = e[g * n + j];
  ofx = fx[g * n + j];
  ofy = fy[g * n + j];
  ofz = fz[g * n + j];
}

// Element i of an array whose base is the same in every lane of the wave.
// The base goes through readfirstlane (free for a value already in SGPRs) so
// that the optimiser cannot fold a per-lane index into it, and the byte
// offset is a 32-bit unsigned value: the access then takes the scalar-base
// form of the global load/store (64-bit base in SGPRs + 32-bit offset in one
// VGPR) rather than a 64-bit address per lane, and arrays indexed by the same
// cell share one offset register.  The level size limit (max_level_cells)
// keeps every offset below 4 GiB.
#define EMMA_GLOBAL __attribute__((address_space(1)))
template <typename T>
__device__ __forceinline__ EMMA_GLOBAL char* uniform_base(T* base) {
  const unsigned long long p = (unsigned long long)base;
  const unsigned long long lo = (unsigned)__builtin_amdgcn_readfirstlane((int)(unsigned)p);
  const unsigned long long hi = (unsigned)__builtin_amdgcn_readfirstlane((int)(unsigned)(p >> 32));
  return (EMMA_GLOBAL char*)((hi << 32) | lo);
}
// the 32-bit byte offset, kept opaque so that a select or mask that formed
// the index is not pushed through the zero extension (which would hide the
// 32-bit offset from the instruction selector)
template <typename T>
__device__ __forceinline__ size_t byte_offset(int i) {
  unsigned off = (unsigned)i * (unsigned)sizeof(T);
  asm("" : "+v"(off));
  return (size_t)off;
}
template <typename T>
__device__ __forceinline__ T gld(const T* base, int i) {
  return *(const EMMA_GLOBAL T*)(uniform_base(base) + byte_offset<T>(i));
}
template <typename T>
__device__ __forceinline__ void gst(T* base, int i, T v) {
  *(EMMA_GLOBAL T*)(uniform_base(base) + byte_offset<T>(i)) = v;
}

// the same with the non-temporal hint (streamed: the line is not kept in L2
// for reuse).  For a large level's next state and its once-read chemistry
// inputs, which would otherwise push out of L2 the neighbouring tiles' cells
// that halo reads are about to fetch; small levels, whose next step finds
// their state in L2, keep the plain form.
template <typename T>
__device__ __forceinline__ void gst_nt(T* base, int i, T v) {
  __builtin_nontemporal_store(v, (EMMA_GLOBAL T*)(uniform_base(base) + byte_offset<T>(i)));
}
template <typename T>
__device__ __forceinline__ T gld_nt(const T* base, int i) {
  return __builtin_nontemporal_load((const EMMA_GLOBAL T*)(uniform_base(base) + byte_offset<T>(i)));
}
template <bool NT, typename T>
__device__ __forceinline__ void gst_out(T* base, int i, T v) {
  if constexpr (NT)
    gst_nt(base, i, v);
  else
    gst(base, i, v);
}
template <bool NT, typename T>
__device__ __forceinline__ T gld_in(const T* base, int i) {
  if constexpr (NT)
    return gld_nt(base, i);
  else
    return gld(base, i);
}

template <typename T>
__device__ __forceinline__ T sel(int i, T a, T b, T c) {
  return i == 0 ? a : (i == 1 ? b : c);
}

// The level-step kernels take (LevelArgs, ChemConst) by value but read them
// through the kernarg segment pointer where they are used.  By-value kernel
// arguments are all loaded into SGPRs at kernel entry; a level step has more
// of them than there are SGPRs, and the overflow is spilled to VGPR lanes and
// read back with a v_readlane per use.  Loads at the point of use keep only
// the live ones in SGPRs.  The explicit arguments sit at offset 0 and right
// after the LevelArgs block (both 8-byte aligned).
template <typename R>
struct KernArgs {
  const LevelArgs<R>& a;
  const ChemConst& K;
};
template <typename R>
__device__ __forceinline__ KernArgs<R> kernargs() {
  static_assert(sizeof(LevelArgs<R>) % alignof(ChemConst) == 0, "ChemConst follows LevelArgs unpadded");
  const char* kp = (const char*)__builtin_amdgcn_kernarg_segment_ptr();  // constant address space -> generic
  return {*reinterpret_cast<const LevelArgs<R>*>(kp), *reinterpret_cast<const ChemConst*>(kp + sizeof(LevelArgs<R>))};
}

// ------------------------------------------------------------------ pieces
// reionisation-time bookkeeping of RadSolver (rad_utils.c:1901-1913)
template <typename R>
__device__ __forceinline__ void xion_times(const LevelArgs<R>& a, int cell) {
  const double xion = a.nhplus[cell] / a.nh[cell];
  if (xion < 0.5) a.t_last[cell] = a.tphy;
  if ((xion >= 0.5) && (a.t_first[cell] == -1)) a.t_first[cell] = a.tphy;
}

// split cell: its next state is the mean of its 8 children (rad_utils.c:1955-2006)
template <typename R>
__device__ __forceinline__ void restrict_cell(const LevelArgs<R>& a, int cell, int ch, int ng) {
  const int n = a.ncell, nf = a.nfine, j0 = ch * 8;
  for (int g = 0; g < ng; g++) {
    R se = R(0), sx = R(0), sy = R(0), sz = R(0), ss = R(0);
#pragma unroll
    for (int k = 0; k < 8; k++) {
      // plain indexing: the eight children's values merge into wide loads
      const int j = g * nf + j0 + k;
      se += a.fe[j] * R(0.125);
      sx += a.ffx[j] * R(0.125);
      sy += a.ffy[j] * R(0.125);
      sz += a.ffz[j] * R(0.125);
      ss += a.fsrc[j] * R(0.125);
    }
    gst(a.eo + g * n, cell, se);
    gst(a.fxo + g * n, cell, sx);
    gst(a.fyo + g * n, cell, sy);
    gst(a.fzo + g * n, cell, sz);
    gst(a.src + g * n, cell, ss);
  }
  // (no branch inside the loop: the 32 loads go out together)
  double snp = 0., sei = 0., stp = 0., snh = 0.;
#pragma unroll
  for (int k = 0; k < 8; k++) {
    const int j = j0 + k;
    snp += a.fnhplus[j] * 0.125;
    sei += a.feint[j] * 0.125;
    stp += a.ftemp[j] * 0.125;
    snh += a.fnh[j] * 0.125;
  }
  gst(a.nhpluso, cell, snp);
  gst(a.einto, cell, sei);
  gst(a.tempo, cell, stp);
  gst(a.nh, cell, snh);
  if (a.t_first != nullptr) {
    double stf = 0., stl = 0.;
#pragma unroll
    for (int k = 0; k < 8; k++) {
      stf += a.ft_first[j0 + k] * 0.125;
      stl += a.ft_last[j0 + k] * 0.125;
    }
    a.t_first[cell] = stf;
    a.t_last[cell] = stl;
  }
}

// reflux of one face flux into the coarse leaf behind it (scatterstencilrad,
// rad_utils.c:1738-1762): the coarse cell gains -/+ F dt/dx / 8.  The cell's
// own accumulator of the reflux pair p (sub-slot q: its place in the oct
// face), its 4 x GA values contiguous.  No other lane adds to it in this
// launch, so the sum is the same every run; the add is an atomic only so
// that the lane need not wait for the old value.
template <typename R, int GA>
__device__ __forceinline__ void reflux(const LevelArgs<R>& a, int p, int q, int f, int g, const R F[4], R dts) {
  R* acc = a.cracc + ((size_t)p * kRefluxCell + q) * 4 * GA;
  const R w = (f & 1) ? R(1) : R(-1);  // pow(-1, face+1)
  atomicAdd(acc + 0 * GA + g, F[0] * dts * w * R(0.125));
  atomicAdd(acc + 1 * GA + g, F[1] * dts * w * R(0.125));
  atomicAdd(acc + 2 * GA + g, F[2] * dts * w * R(0.125));
  atomicAdd(acc + 3 * GA + g, F[3] * dts * w * R(0.125));
}

// leaf cell after its fluxes: base state (+ refluxes from the finer level),
// update, N<0/NaN guard, chemistry, write of the next state.  Returns the
// accepted chemistry sub-steps.
// The chemistry inputs (src, x n_H, eint, n_H) come from memory, or from
// `pre` when the caller loaded them early (pre_src[GA], pre[3]).
template <typename R, int G, bool Fast, int Sub, bool NT = false>
__device__ __forceinline__ int finish_leaf(const LevelArgs<R>& a, const ChemConst& K, int cell, int ch, int ng,
                                           R* Ue, R* Ux, R* Uy, R* Uz, const R* De, const R* Dx, const R* Dy,
                                           const R* Dz, const R* pre_src = nullptr, const double* pre = nullptr) {
  constexpr int GA = G > 0 ? G : kMaxGrp;
  const int n = a.ncell;
  if (ch <= -2) {  // the refluxes of the finer octs behind its faces, in pair order
    const int code = -2 - ch;
    R* acc = a.racc + (size_t)(code >> 3) * Sub * 4 * GA;
    const int m = (code & 7) * Sub;
    // all loads before the clearing stores.  A leaf has at most six pairs:
    // the tiled kernel's loop unrolls, and its loads (an accumulator's 4G
    // values are contiguous), clamped to the leaf's last accumulator, go out
    // together instead of one pair after another
    constexpr int kMaxSub = 6 * Sub;
    R s[GA][4];
#pragma unroll
    for (int g = 0; g < GA; g++) s[g][0] = s[g][1] = s[g][2] = s[g][3] = R(0);
#pragma unroll(Sub == 1 ? kMaxSub : 1)
    for (int i = 0; i < (Sub == 1 ? kMaxSub : m); i++) {
      const R* p = acc + (size_t)min(i, m - 1) * 4 * GA;
#pragma unroll
      for (int v = 0; v < 4; v++) {
#pragma unroll
        for (int g = 0; g < GA; g++) {
          if (g < ng) {
            const R x = p[v * GA + g];
            s[g][v] += i < m ? x : R(0);
          }
        }
      }
    }
#pragma unroll(Sub == 1 ? kMaxSub : 1)
    for (int i = 0; i < kMaxSub; i++) {
      if (i < m) {
#pragma unroll
        for (int v = 0; v < 4; v++) {
#pragma unroll
          for (int g = 0; g < GA; g++)
            if (g < ng) acc[(size_t)i * 4 * GA + v * GA + g] = R(0);
        }
      }
    }
#pragma unroll
    for (int g = 0; g < GA; g++) {
      if (g < ng) {
        Ue[g] += s[g][0];
        Ux[g] += s[g][1];
        Uy[g] += s[g][2];
        Uz[g] += s[g][3];
      }
    }
  }
  double e[GA], src[GA];
  R fx[GA], fy[GA], fz[GA];  // fluxes after transport, in the storage type
#pragma unroll
  for (int g = 0; g < GA; g++) {
    if (g < ng) {
      e[g] = (double)(R)(Ue[g] + De[g]);
      fx[g] = Ux[g] + Dx[g];
      fy[g] = Uy[g] + Dy[g];
      fz[g] = Uz[g] + Dz[g];
      src[g] = (double)(pre_src != nullptr ? pre_src[g] : gld(a.src + g * n, cell));
    }
  }
  // rad_utils.c:1611-1650 (the second test reads e[1], i.e. fx[0] when G=1)
  if (e[0] < 0 || isnan(e[0]) || isnan(ng > 1 ? e[1 < GA ? 1 : 0] : (double)fx[0]))
    raise_status(a.status, -3, a.level, cell);
  double nhplus = pre != nullptr ? pre[0] : gld(a.nhplus, cell);
  double eint = pre != nullptr ? pre[1] : gld(a.eint, cell);
  const double nh = pre != nullptr ? pre[2] : gld(a.nh, cell);
  double temp = 0.;
  int nit = 0;
  int st;
  bool deferred = false;  // chemistry left to the straggler pass
  if constexpr (Fast) {
    double fscale[GA];
    st = chem_cell_fast<G>(K, ng, e, fscale, src, nhplus, eint, nh, temp, a.chemonly, nit,
                           Sub == kRefluxTiled && a.chem_budget > 0 ? a.chem_budget : -1, (double)a.chem_keep);
    if constexpr (Sub == kRefluxTiled) {
      // budget spent: the transported N, F are stored unscaled (fscale = 1)
      // and the cell goes on the straggler list, one atomic per wave
      deferred = st == 1;
      const unsigned long long dm = __ballot(deferred);
      if (dm != 0ull) {
        const int leader = __ffsll((long long)dm) - 1;
        int base = 0;
        if ((int)__lane_id() == leader) base = atomicAdd(a.defer_ctl, __popcll(dm));
        base = __shfl(base, leader);
        if (deferred) {
          const int slot = base + lanes_below_mask(dm);
          if (slot < a.defer_cap)
            a.defer_list[slot] = cell;
          else
            raise_status(a.status, -1, a.level, cell);  // list smaller than the level (host sizes it)
          st = 0;
          nit = 0;
        }
      }
    }
#pragma unroll
    for (int g = 0; g < GA; g++) {
      if (g < ng) {
        const R fs = (R)fscale[g];
        fx[g] *= fs;
        fy[g] *= fs;
        fz[g] *= fs;
      }
    }
  } else {
    double dfx[GA], dfy[GA], dfz[GA];
#pragma unroll
    for (int g = 0; g < GA; g++) {
      if (g < ng) {
        dfx[g] = (double)fx[g];
        dfy[g] = (double)fy[g];
        dfz[g] = (double)fz[g];
      }
    }
    st = chem_cell<G>(K, ng, e, dfx, dfy, dfz, src, nhplus, eint, nh, temp, a.chemonly, nit);
#pragma unroll
    for (int g = 0; g < GA; g++) {
      if (g < ng) {
        fx[g] = (R)dfx[g];
        fy[g] = (R)dfy[g];
        fz[g] = (R)dfz[g];
      }
    }
  }
  if (st != 0) raise_status(a.status, st, a.level, cell);
#pragma unroll
  for (int g = 0; g < GA; g++) {
    if (g < ng) {
      R ev = (R)e[g];
      // FP32 storage: a density the chemistry left at (or just above) its
      // floor must not round below it, or the next step would stop at the
      // floor test (chem_utils.c:382); only the rounding is corrected
      if constexpr (sizeof(R) == 4)
        if (!deferred && e[g] >= K.emin_grp[g] && ev < K.emin_grp_up[g]) ev = K.emin_grp_up[g];
      gst_out<NT>(a.eo + g * n, cell, ev);
      gst_out<NT>(a.fxo + g * n, cell, fx[g]);
      gst_out<NT>(a.fyo + g * n, cell, fy[g]);
      gst_out<NT>(a.fzo + g * n, cell, fz[g]);
    }
  }
  gst_out<NT>(a.nhpluso, cell, nhplus);
  gst_out<NT>(a.einto, cell, eint);
  gst_out<NT>(a.tempo, cell, temp);
  return nit;
}

// COARSERAD coarse level: a split cell takes the transport update of its own
// state (updatefieldrad without the split test, rad_utils.c:1548-1552) and
// keeps x, eint, T (chemrad skips it)
template <typename R, int GA>
__device__ __forceinline__ void transport_split_cell(const LevelArgs<R>& a, int cell, int ng, const R* Ue, const R* Ux,
                                                     const R* Uy, const R* Uz, const R* De, const R* Dx, const R* Dy,
                                                     const R* Dz) {
  const int n = a.ncell;
#pragma unroll
  for (int g = 0; g < GA; g++) {
    if (g < ng) {
      gst(a.eo + g * n, cell, Ue[g] + De[g]);
      gst(a.fxo + g * n, cell, Ux[g] + Dx[g]);
      gst(a.fyo + g * n, cell, Uy[g] + Dy[g]);
      gst(a.fzo + g * n, cell, Uz[g] + Dz[g]);
    }
  }
  const double e0 = (double)(R)(Ue[0] + De[0]);
  const double e1 = ng > 1 ? (double)(R)(Ue[1 < GA ? 1 : 0] + De[1 < GA ? 1 : 0]) : (double)(R)(Ux[0] + Dx[0]);
  if (e0 < 0 || isnan(e0) || isnan(e1)) raise_status(a.status, -3, a.level, cell);
  gst(a.nhpluso, cell, gld(a.nhplus, cell));
  gst(a.einto, cell, gld(a.eint, cell));
  gst(a.tempo, cell, gld(a.temp, cell));
}

// per-block counters: accepted chemistry sub-steps and leaf cell-updates,
// one atomic pair per block into a spread slot.  The wave sums come from
// ballots (scalar popcounts): the leaf flags directly, the sub-step counts
// bit plane by bit plane over their low 4 bits; a lane with 16 or more
// sub-steps adds the rest with its own atomic.
template <int NW>
__device__ __forceinline__ void block_counters(unsigned long long* counters, unsigned int nit, unsigned int leaf) {
  __shared__ unsigned int red[2][NW];
  unsigned long long* slot = counters + (size_t)(blockIdx.x % kCounterSlots) * kCounterStride;
  unsigned int s0 = 0;
#pragma unroll
  for (int b = 0; b < 4; b++) s0 += (unsigned int)__popcll(__ballot((nit >> b) & 1u)) << b;
  const unsigned int s1 = (unsigned int)__popcll(__ballot(leaf != 0u));
  if (nit >= 16u) atomicAdd(slot, (unsigned long long)(nit & ~15u));
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) {
    red[0][w] = s0;
    red[1][w] = s1;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long t0 = 0, t1 = 0;
    for (int i = 0; i < NW; i++) {
      t0 += red[0][i];
      t1 += red[1][i];
    }
    atomicAdd(slot + 0, t0);
    atomicAdd(slot + 1, t1);
  }
}

// ------------------------------------------------------------------ level step
// One thread per cell; neighbours straight from global memory.  Fast selects
// the mixed-precision physics (rt_physics_fast.h); the FP64 kernel
// (Fast = false) keeps the reference's operation order.
// The cells of block `block` (256 consecutive cells, 32 octs).
template <typename R, int G, bool Fast>
__device__ __forceinline__ void cell_step(const LevelArgs<R>& a, const ChemConst& K, int block, int lc) {
  constexpr int GA = G > 0 ? G : kMaxGrp;
  const int ng = G > 0 ? G : a.ng;
  const int cell = block * 256 + lc;
  unsigned int nit_acc = 0;
  unsigned int leaf = 0;
  if (cell < a.ncell) {
    const int n = a.ncell;
    const int oct = cell >> 3, ic = cell & 7;
    const int ch = a.child[cell];
    if (a.t_first != nullptr) xion_times(a, cell);
    if (ch >= 0 && !a.update_split) {
      restrict_cell(a, cell, ch, ng);
    } else {
      const R c = a.c, hc = a.hc, dts = a.dtsurdx;
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
          De[g] = Dx[g] = Dy[g] = Dz[g] = R(0);
        }
      }
      // six faces in the reference order (rad_utils.c:1566-1574): f = 2d+s
#pragma unroll
      for (int f = 0; f < (a.transport ? 6 : 0); f++) {
        const int d = f >> 1, s = f & 1;
        const int nc = ic ^ (1 << d);
        const R* Ne = a.e;
        const R* Nx = a.fx;
        const R* Ny = a.fy;
        const R* Nz = a.fz;
        int nn = n, j, split;
        int coarse_face = 0, pr = -1, reflect = 0;
        if (((ic >> d) & 1) != s) {  // neighbour inside the oct
          j = oct * 8 + nc;
          split = a.child[j] >= 0;
        } else {
          const int no = a.nei[oct * 6 + f];
          if (no >= 0) {
            j = no * 8 + nc;
            split = a.child[j] >= 0;
          } else if (no <= kNeiReflective) {  // domain face (rad_utils.c:1106-1213, 1376-1407)
            j = oct * 8 + trans_source(f, ic);
            split = a.child[j] >= 0;
            reflect = no == kNeiReflective;
          } else {  // coarse leaf injected into the stencil
            j = -1 - no;
            split = 0;
            Ne = a.ce;
            Nx = a.cfx;
            Ny = a.cfy;
            Nz = a.cfz;
            nn = a.ncoarse;
            coarse_face = 1;
            pr = a.rpair[oct * 6 + f];
          }
        }
        const R ffact = (split && !a.keep_split_flux) ? R(0) : R(1);
        const R one = (f & 1) ? R(-1) : R(1);
#pragma unroll
        for (int g = 0; g < GA; g++) {
          if (g < ng) {
            R ve, vx, vy, vz;
            load4(a, Ne, Nx, Ny, Nz, nn, j, g, ve, vx, vy, vz);
            if (reflect) {
              if (d == 0) vx = -vx;
              if (d == 1) vy = -vy;
              if (d == 2) vz = -vz;
            }
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
            if (coarse_face) {
              // the cell's place in the oct face: its two coordinates across d
              const int q = ((ic >> ((d + 1) % 3)) & 1) | (((ic >> ((d + 2) % 3)) & 1) << 1);
              reflux<R, GA>(a, pr, q, f, g, F, dts);
            }
          }
        }
      }
      if (ch >= 0) {
        transport_split_cell<R, GA>(a, cell, ng, Ue, Ux, Uy, Uz, De, Dx, Dy, Dz);
      } else {
        leaf = 1;
        nit_acc =
            (unsigned)finish_leaf<R, G, Fast, kRefluxCell>(a, K, cell, ch, ng, Ue, Ux, Uy, Uz, De, Dx, Dy, Dz);
      }
    }
  }
  if (a.counters != nullptr) block_counters<4>(a.counters, nit_acc, leaf);
}

template <typename R, int G, bool Fast>
__global__ __launch_bounds__(256) void level_step_kernel(LevelArgs<R> a_by_value, ChemConst K_by_value) {
  const KernArgs<R> args = kernargs<R>();
  cell_step<R, G, Fast>(args.a, args.K, xcd_block(args.a.nblocks), (int)threadIdx.x);
}

// ------------------------------------------------------------------ tiled level step
// Mixed precision, G fixed.  A 256-thread block owns 32 consecutive octs
// (Morton order: a 4x4x2 brick of octs on a complete level), one cell per
// thread.  With about 38 KB of LDS at G=3, four blocks share a CU, so the
// barrier phases of one block overlap the memory phases of the others.
//
// The Lax-Friedrichs flux through a face normal to d splits into one share
// per side: 2F = A(M) + B(P) with A = P + c U on the minus side and
// B = P - c U on the plus side, P the physical flux (lf_share).  So each
// cell computes its own A_d and B_d once per direction and a face costs one
// LDS read of the neighbour's share plus an add.  The block runs the three
// directions as three passes over one LDS share buffer.
//
// Neighbours outside the tile (same-level octs of other tiles, injected
// coarse leaves) are the halo.  The neighbour's share and whether the face
// flux is cancelled (split neighbour, RADSTEP) go to LDS slots of the wave,
// 64 per wave, so an in-tile and a halo face differ only in the LDS address
// read.  A complete level has 64 halo faces per wave (16 in x and y, 32 in
// z): each wave lists them (lane, direction) and evaluates them
// lane-parallel, one per lane, in one round before the direction passes
// (PerDir = false).  Sparse AMR levels have up to 192 (every cell's three
// leaving faces): there each lane evaluates its own halo face of direction d
// at the start of pass d into its own slot (PerDir = true), which needs no
// more LDS than the complete-level kernel.  The host picks the variant from
// the level's topology (the most halo faces in any wave).

// the value of lane (lane ^ m), m = 1, 2 or 4 (wave-uniform), through DPP
// lane moves within rows of 16 (VALU, no LDS): quad permutations for 1 and
// 2, a half-row mirror (lane ^ 7) followed by a quad reversal (lane ^ 3) for
// 4.  Every lane of the wave must be active.
__device__ __forceinline__ float lane_xor(float x, int m) {
  const int v = __float_as_int(x);
  if (m == 1) return __int_as_float(__builtin_amdgcn_mov_dpp(v, 0xB1, 0xF, 0xF, true));  // quad_perm 1,0,3,2
  if (m == 2) return __int_as_float(__builtin_amdgcn_mov_dpp(v, 0x4E, 0xF, 0xF, true));  // quad_perm 2,3,0,1
  const int t = __builtin_amdgcn_mov_dpp(v, 0x141, 0xF, 0xF, true);                       // row_half_mirror
  return __int_as_float(__builtin_amdgcn_mov_dpp(t, 0x1B, 0xF, 0xF, true));              // quad_perm 3,2,1,0
}

// The halo variants kListed, kPerDir, kLatency (rt_kernels.h) share this
// body; the kernels below instantiate it.
// one halo face: the neighbour's state, its split flag, the face direction
// and the side the neighbour lies on
template <int G>
struct HaloFace {
  float v[G][4];
  int child, d, side;
  bool lane, coarse, refl;
};

template <int G, int V>
__device__ __forceinline__ void tiled_step(const LevelArgs<float>& a, const ChemConst& K, int tile, TileLds<G>& lds,
                                           int lc) {
  constexpr bool Listed = V == kListed || V == kListedTop;
  constexpr bool PerDir = !Listed;
  constexpr bool Late = V == kPerDir;  // chemistry inputs loaded in the last direction pass
  constexpr bool NT = Listed;          // non-temporal next state and chemistry inputs (large levels)
  constexpr bool Top = V == kListedTop;  // no coarse faces on the level
  constexpr int kHaloCap = TileLds<G>::kHaloCap;  // halo slots per wave
  // shares: [g][A or B][cell] of the current direction, then the halo slots
  // [wave][slot][g]
  constexpr int kShares = G * 2 * kTileCells;
  float4* sShare = lds.share;
  // 1: the flux through the face is cancelled; [cell], then [wave][slot]
  unsigned char* sCancel = lds.cancel;
  auto& sItem = lds.item;         // halo slot -> lane*4 + direction
  auto& sItemNei = lds.item_nei;  // halo slot -> nei_oct entry of the face
  const int wv = lc >> 6, lane = lc & 63;
  const int cell = tile * kTileCells + lc;
  const int oct0 = tile * kTileOcts;
  const bool valid = cell < a.ncell;
  const int n = a.ncell;
  const float c = a.c, dts = a.dtsurdx;
  // all loads of the cell are issued first: state, child, the oct's face
  // neighbours and the chemistry inputs (their latency hides behind the
  // transport)
  const int oct = cell >> 3, ic = cell & 7;
  float Ue[G], Ux[G], Uy[G], Uz[G], Src[G];
  double pre[3] = {0., 0., 1.};  // x n_H, eint, n_H
  int noD[3] = {0, 0, 0};        // nei_oct entry of the face that leaves the oct, per direction
  int ch = -1;
  if (valid) {
    // in the order of use: topology (halo list), state, chemistry inputs
    ch = gld(a.child, cell);
#pragma unroll
    for (int d = 0; d < 3; d++) noD[d] = gld(a.nei, oct * 6 + 2 * d + ((ic >> d) & 1));
#pragma unroll
    for (int g = 0; g < G; g++) {
      Ue[g] = gld(a.e + g * n, cell);
      Ux[g] = gld(a.fx + g * n, cell);
      Uy[g] = gld(a.fy + g * n, cell);
      Uz[g] = gld(a.fz + g * n, cell);
    }
    // the chemistry inputs: early on complete levels (their latency hides
    // behind the transport); on sparse levels, whose per-direction halo
    // loads need the registers, at the start of the last direction pass
    // (load_chem below)
    if constexpr (!Late) {
#pragma unroll
      for (int g = 0; g < G; g++) Src[g] = gld_in<NT>(a.src + g * n, cell);
      pre[0] = gld_in<NT>(a.nhplus, cell);
      pre[1] = gld_in<NT>(a.eint, cell);
      pre[2] = gld_in<NT>(a.nh, cell);
    }
  }
  sCancel[lc] = (unsigned char)(ch >= 0 && !a.keep_split_flux);
  // sparse levels: the chemistry inputs of a leaf (the only cells whose step
  // runs the chemistry), issued while the last direction pass computes
  auto load_chem = [&]() {
    if (valid && ch < 0) {
#pragma unroll
      for (int g = 0; g < G; g++) Src[g] = gld(a.src + g * n, cell);
      pre[0] = gld(a.nhplus, cell);
      pre[1] = gld(a.eint, cell);
      pre[2] = gld(a.nh, cell);
    }
  };

  const bool transport = a.transport != 0;
  const bool moves = valid && transport && !(ch >= 0 && !a.update_split);
  auto in_tile = [&](int no) { return no >= oct0 && no < oct0 + kTileOcts; };
  // ---- complete levels: list the halo faces per wave
  unsigned long long hball[3] = {0ull, 0ull, 0ull};  // lanes whose face in direction d is a halo face
  int hbase[3] = {0, 0, 0};                          // first halo slot of direction d
  int nhalo = 0;
  if constexpr (!PerDir) {
#pragma unroll
    for (int d = 0; d < 3; d++) {
      const bool out = moves && !in_tile(noD[d]);
      hball[d] = __ballot(out);
      hbase[d] = nhalo;
      nhalo += __popcll(hball[d]);
      if (out) {
        const int slot = hbase[d] + lanes_below_mask(hball[d]);  // < kHaloCap (variant chosen from the topology)
        sItem[wv][slot] = (unsigned char)(lane * 4 + d);
        sItemNei[wv][slot] = noD[d];
      }
    }
    // the list is written and read by lanes of the same wave
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  }
  // the neighbour across the face in direction h.d that leaves the oct of
  // cell hcell, nei_oct entry hno: a same-level cell, an injected coarse
  // leaf, or on a domain face a copy of a cell of the same oct (its normal
  // flux negated when reflective)
  auto load_neighbour = [&](HaloFace<G>& h, int hno, int hcell) {
    // only loads here: the reflective sign and nothing else that needs the
    // loaded values is applied before store_item, so the loads of several
    // faces can be in flight together
    const int hic = hcell & 7;
    h.side = (hic >> h.d) & 1;
    const bool hbound = h.lane && hno <= kNeiReflective;
    h.coarse = h.lane && hno < 0 && !hbound;
    h.refl = hbound && hno == kNeiReflective;
    const int js = hbound ? (hcell & ~7) + trans_source(2 * h.d + h.side, hic)
                          : ((h.lane && !h.coarse) ? hno * 8 + (hic ^ (1 << h.d)) : 0);
    h.child = gld(a.child, js);
    if (Top || __ballot(h.coarse) == 0) {
#pragma unroll
      for (int g = 0; g < G; g++) {
        h.v[g][0] = gld(a.e + g * n, js);
        h.v[g][1] = gld(a.fx + g * n, js);
        h.v[g][2] = gld(a.fy + g * n, js);
        h.v[g][3] = gld(a.fz + g * n, js);
      }
    } else {  // injected coarse leaves (levels with a coarser level only): per-lane arrays
      const int nc = a.ncoarse;
      const int j = h.coarse ? -1 - hno : js;
      const int st = h.coarse ? nc : n;
      // The lane's array is this level's plus a per-lane distance (0, or the
      // coarse array's address minus this level's), both pointers read
      // unconditionally by scalar loads.  A per-lane choice between the two
      // argument fields instead becomes a per-lane vector load of the pointer
      // (one more memory round trip before the data loads), or scalar loads
      // under a branch each.
      auto lane_base = [&](const float* own, const float* coarse) {
        const unsigned long long dist = (unsigned long long)coarse - (unsigned long long)own;
        return uniform_base(own) + (h.coarse ? dist : 0ull);
      };
      const EMMA_GLOBAL char* pe = lane_base(a.e, a.ce);
      const EMMA_GLOBAL char* px = lane_base(a.fx, a.cfx);
      const EMMA_GLOBAL char* py = lane_base(a.fy, a.cfy);
      const EMMA_GLOBAL char* pz = lane_base(a.fz, a.cfz);
#pragma unroll
      for (int g = 0; g < G; g++) {
        const size_t o = byte_offset<float>(g * st + j);
        h.v[g][0] = *(const EMMA_GLOBAL float*)(pe + o);
        h.v[g][1] = *(const EMMA_GLOBAL float*)(px + o);
        h.v[g][2] = *(const EMMA_GLOBAL float*)(py + o);
        h.v[g][3] = *(const EMMA_GLOBAL float*)(pz + o);
      }
    }
  };
  // complete levels: list entry `slot` (read by every lane, used by halo lanes)
  auto load_item = [&](HaloFace<G>& h, int slot) {
    h.lane = slot < nhalo;
    const int it = sItem[wv][PerDir ? 0 : slot];
    h.d = it & 3;
    load_neighbour(h, sItemNei[wv][PerDir ? 0 : slot], (cell & ~63) + (it >> 2));
  };
  // sparse levels: this lane's own face in direction d
  auto load_face = [&](HaloFace<G>& h, int d) {
    h.lane = moves && !in_tile(noD[d]);
    h.d = d;
    load_neighbour(h, noD[d], cell);
  };
  // ... the neighbour's share (a neighbour on the plus side of the face,
  // side 1, contributes B, on the minus side A) and cancel flag
  auto store_item = [&](const HaloFace<G>& h, int slot) {
    if (h.lane) {
      sCancel[kTileCells + wv * kHaloCap + slot] = (unsigned char)(!h.coarse && !a.keep_split_flux && h.child >= 0);
      // reflective face: the normal component negated (h.d differs between lanes: no indexing by it)
      const float sx = (h.refl && h.d == 0) ? -1.f : 1.f, sy = (h.refl && h.d == 1) ? -1.f : 1.f,
                  sz = (h.refl && h.d == 2) ? -1.f : 1.f;
#pragma unroll
      for (int g = 0; g < G; g++) {
        const float vx = h.v[g][1] * sx, vy = h.v[g][2] * sy, vz = h.v[g][3] * sz;
        const Closure<float> q = closure_fast(vx, vy, vz, h.v[g][0], c);
        float X[4];
        lf_share_lane(h.d, h.side ? -1.f : 1.f, h.v[g][0], vx, vy, vz, q, c, X);
        sShare[kShares + (wv * kHaloCap + slot) * G + g] = make_float4(X[0], X[1], X[2], X[3]);
      }
    }
  };
  // ---- one halo face per lane: issue the neighbour's loads ...
  // Straight-line code (every lane loads, lanes without a halo face from
  // cell 0): a load under a divergent branch makes the wait before the
  // closures below cover it as well.
  HaloFace<G> hf[V == kLatency ? 3 : 1];
  if constexpr (V == kLatency) {
#pragma unroll
    for (int d = 0; d < 3; d++) load_face(hf[d], d);
  }
  if constexpr (!PerDir) load_item(hf[0], lane);
  // the reflux pairs of the cell's leaving faces that have a coarse leaf
  // behind them, loaded now for the refluxing in the direction passes
  int pairD[3] = {-1, -1, -1};
  if (valid && a.ncoarse > 0) {
#pragma unroll
    for (int d = 0; d < 3; d++)
      if (noD[d] < 0) pairD[d] = gld(a.rpair, oct * 6 + 2 * d + ((ic >> d) & 1));
  }
  // ---- ... compute this cell's closures meanwhile ...
  Closure<float> Q[G];
  if (valid) {
#pragma unroll
    for (int g = 0; g < G; g++) Q[g] = closure_fast(Ux[g], Uy[g], Uz[g], Ue[g], c);
  }
  if constexpr (!PerDir) store_item(hf[0], lane);

  float De[G], Dx[G], Dy[G], Dz[G];
#pragma unroll
  for (int g = 0; g < G; g++) De[g] = Dx[g] = Dy[g] = Dz[g] = 0.f;
  float ucoef = 0.f;  // coefficient of c U summed over the directions
  if (transport) {
#pragma unroll
    for (int d = 0; d < 3; d++) {
      if constexpr (V == kPerDir) {
        load_face(hf[0], d);  // latency behind this cell's shares
        if (d == 2) load_chem();
      }
      if (valid) {
#pragma unroll
        for (int g = 0; g < G; g++) {
          float X[4];
          lf_share_unrolled(d, 1.f, Ue[g], Ux[g], Uy[g], Uz[g], Q[g], c, X);
          sShare[(g * 2 + 0) * kTileCells + lc] = make_float4(X[0], X[1], X[2], X[3]);
          lf_share_unrolled(d, -1.f, Ue[g], Ux[g], Uy[g], Uz[g], Q[g], c, X);
          sShare[(g * 2 + 1) * kTileCells + lc] = make_float4(X[0], X[1], X[2], X[3]);
        }
      }
      if constexpr (PerDir) store_item(hf[V == kLatency ? d : 0], lane);  // own slot: written and read by this lane only
      __syncthreads();
      float Rf[G][4];  // flux through a coarse face (refluxing), 0 elsewhere
      int rj = -1;     // the coarse leaf on that face
#pragma unroll
      for (int g = 0; g < G; g++) Rf[g][0] = Rf[g][1] = Rf[g][2] = Rf[g][3] = 0.f;
      const int bit = (ic >> d) & 1;
      if (moves) {
        // Each cell has one face in d inside its oct and one leaving it; bit
        // says which is which (bit 1: the cell sits on the plus side of its
        // oct, so its minus face is the inside one).  The face on the cell's
        // minus side takes the neighbour's A plus the cell's B, the face on
        // its plus side the cell's A plus the neighbour's B.
        const int jin = lc ^ (1 << d);
        const int iin = (1 - bit) * kTileCells + jin;  // neighbour share, inside face
        const bool inT = in_tile(noD[d]);
        // (unsigned: a domain-face entry is near INT_MIN; jl and j are then unused)
        const int jl = (int)(((unsigned)noD[d] - (unsigned)oct0) * 8u + (unsigned)(ic ^ (1 << d)));
        const int slot = PerDir ? lane : hbase[d] + lanes_below_mask(hball[d]);
        const bool coarse_face = !Top && !inT && noD[d] < 0 && noD[d] > kNeiReflective;
        int iout, sout, cout;  // neighbour share, stride over groups, cancel flag: leaving face
        if (inT) {
          iout = bit * kTileCells + jl;
          sout = 2 * kTileCells;
          cout = jl;
        } else {
          iout = kShares + (wv * kHaloCap + slot) * G;
          sout = 1;
          cout = kTileCells + wv * kHaloCap + slot;
        }
        const int j = coarse_face ? -1 - noD[d] : (int)((unsigned)noD[d] * 8u + (unsigned)(ic ^ (1 << d)));
        const bool cancel_out = sCancel[cout] != 0;
        const bool cancel_in = sCancel[jin] != 0;
        const int im = bit ? iin : iout, sm = bit ? 2 * kTileCells : sout;
        const int ip = bit ? iout : iin, sp = bit ? sout : 2 * kTileCells;
        // The shares are doubled: the face fluxes are 2F_m = X_m + B and
        // 2F_p = A + X_p, with X the neighbours' shares and A = P + cU,
        // B = P - cU this cell's.  The update wm 2F_m + wp 2F_p then is
        // wm X_m + wp X_p + (wm + wp) P + (wp - wm) cU.  On complete levels
        // nearly every cell has wm = -wp (both faces carry flux, or
        // neither does): the P term drops and the cU term is summed over the
        // directions and applied once, and only the rare cell with one
        // cancelled face or a coarse face reads its own shares back.  Sparse
        // levels, where coarse faces are common, keep the plain form.
        const float wm = (bit ? cancel_in : cancel_out) ? 0.f : 0.5f * dts;
        const float wp = (bit ? cancel_out : cancel_in) ? 0.f : -0.5f * dts;
        const bool plain = !PerDir && wm == -wp;
        if (plain) ucoef += wp - wm;
#pragma unroll
        for (int g = 0; g < G; g++) {
          const float4 Xm = sShare[im + g * sm];
          const float4 Xp = sShare[ip + g * sp];
          if (!PerDir) {
            De[g] = fmaf(Xp.x, wp, fmaf(Xm.x, wm, De[g]));
            Dx[g] = fmaf(Xp.y, wp, fmaf(Xm.y, wm, Dx[g]));
            Dy[g] = fmaf(Xp.z, wp, fmaf(Xm.z, wm, Dy[g]));
            Dz[g] = fmaf(Xp.w, wp, fmaf(Xm.w, wm, Dz[g]));
          }
          if (!plain || coarse_face) {
            // this cell's shares are read back rather than held across the barrier
            const float4 Ag = sShare[(g * 2 + 0) * kTileCells + lc];
            const float4 Bg = sShare[(g * 2 + 1) * kTileCells + lc];
            const float Fm[4] = {Xm.x + Bg.x, Xm.y + Bg.y, Xm.z + Bg.z, Xm.w + Bg.w};
            const float Fp[4] = {Ag.x + Xp.x, Ag.y + Xp.y, Ag.z + Xp.z, Ag.w + Xp.w};
            if (PerDir) {
              De[g] = fmaf(Fp[0], wp, fmaf(Fm[0], wm, De[g]));
              Dx[g] = fmaf(Fp[1], wp, fmaf(Fm[1], wm, Dx[g]));
              Dy[g] = fmaf(Fp[2], wp, fmaf(Fm[2], wm, Dy[g]));
              Dz[g] = fmaf(Fp[3], wp, fmaf(Fm[3], wm, Dz[g]));
            } else if (!plain) {
              De[g] = fmaf(Ag.x, wp, fmaf(Bg.x, wm, De[g]));
              Dx[g] = fmaf(Ag.y, wp, fmaf(Bg.y, wm, Dx[g]));
              Dy[g] = fmaf(Ag.z, wp, fmaf(Bg.z, wm, Dy[g]));
              Dz[g] = fmaf(Ag.w, wp, fmaf(Bg.w, wm, Dz[g]));
            }
            if (coarse_face) {  // the flux through the coarse face, for refluxing
#pragma unroll
              for (int v = 0; v < 4; v++) Rf[g][v] = bit ? Fp[v] : Fm[v];
            }
          }
        }
        if (coarse_face) rj = j;
      }
      // Refluxing (scatterstencilrad, rad_utils.c:1738-1762): the four cells
      // of an oct face share its coarse neighbour, so their fluxes are summed
      // across lanes first (the same order every run) and each of the four
      // lanes then adds one flux component (E, Fx, Fy, Fz) of every group to
      // the face's reflux pair, whose accumulators no other lane writes (an
      // atomic add without return: the lane does not wait for the old value)
      const unsigned long long rmask = __ballot(rj >= 0);
      if (rmask != 0ull) {
        const int b1 = 1 << ((d + 1) % 3), b2 = 1 << ((d + 2) % 3);
#pragma unroll
        for (int g = 0; g < G; g++)
#pragma unroll
          for (int v = 0; v < 4; v++) {
            float x = Rf[g][v];
            x += lane_xor(x, b1);
            x += lane_xor(x, b2);
            Rf[g][v] = x;
          }
        const int l0 = lane & ~(b1 | b2);
        const unsigned long long grp =
            (1ull << l0) | (1ull << (l0 | b1)) | (1ull << (l0 | b2)) | (1ull << (l0 | b1 | b2));
        const unsigned long long mine = rmask & grp;
        const int pr = pairD[d];
        if (rj >= 0 && pr >= 0) {
          float* acc = a.cracc + (size_t)pr * kRefluxTiled * 4 * G;
          const float w = (bit ? 1.f : -1.f) * (0.5f * dts) * 0.125f;  // pow(-1, face+1) dt/dx / 8, shares doubled
          if (mine == grp) {  // the face's four cells all carry flux (always, with 2:1 balance)
            const int k = __popcll(mine & ((1ull << lane) - 1ull));  // component of this lane
#pragma unroll
            for (int g = 0; g < G; g++) {
              const float x = k == 0 ? Rf[g][0] : (k == 1 ? Rf[g][1] : (k == 2 ? Rf[g][2] : Rf[g][3]));
              atomicAdd(acc + k * G + g, x * w);
            }
          } else if (lane == __builtin_ctzll(mine)) {
#pragma unroll
            for (int g = 0; g < G; g++)
#pragma unroll
              for (int v = 0; v < 4; v++) atomicAdd(acc + v * G + g, Rf[g][v] * w);
          }
        }
      }
      if (d < 2) __syncthreads();  // the next direction overwrites the shares
    }
    const float cu = ucoef * c;
#pragma unroll
    for (int g = 0; g < G; g++) {
      De[g] = fmaf(cu, Ue[g], De[g]);
      Dx[g] = fmaf(cu, Ux[g], Dx[g]);
      Dy[g] = fmaf(cu, Uy[g], Dy[g]);
      Dz[g] = fmaf(cu, Uz[g], Dz[g]);
    }
  }

  if constexpr (Late)
    if (!transport) load_chem();
  unsigned int nit_acc = 0, leaf = 0;
  if (valid) {
    if (a.t_first != nullptr) xion_times(a, cell);
    if (ch >= 0 && !a.update_split) {
      restrict_cell(a, cell, ch, G);
    } else if (ch >= 0) {
      transport_split_cell<float, G>(a, cell, G, Ue, Ux, Uy, Uz, De, Dx, Dy, Dz);
    } else {
      leaf = 1;
      nit_acc = (unsigned)finish_leaf<float, G, true, kRefluxTiled, NT>(a, K, cell, ch, G, Ue, Ux, Uy, Uz, De, Dx, Dy, Dz,
                                                                    Src, pre);
    }
  }
  if (a.counters != nullptr) block_counters<kTileWaves>(a.counters, nit_acc, leaf);
}

template <int G, bool PerDir>
__global__ __launch_bounds__(kTileCells) __attribute__((amdgpu_waves_per_eu(4))) void level_step_tiled(
    LevelArgs<float> a_by_value, ChemConst K_by_value) {
  const KernArgs<float> args = kernargs<float>();
  __shared__ TileLds<G> lds;
  const int m = args.a.raster_m;
  const int tile = m > 0 ? raster_tile(m) : xcd_block(args.a.nblocks);
  if constexpr (PerDir)
    tiled_step<G, kPerDir>(args.a, args.K, tile, lds, (int)threadIdx.x);
  else if (args.a.ncoarse == 0)  // wave-uniform: the level's arguments
    tiled_step<G, kListedTop>(args.a, args.K, tile, lds, (int)threadIdx.x);
  else
    tiled_step<G, kListed>(args.a, args.K, tile, lds, (int)threadIdx.x);
}

template <int G>
__global__ __launch_bounds__(kTileCells) __attribute__((amdgpu_waves_per_eu(2))) void level_step_tiled_latency(
    LevelArgs<float> a_by_value, ChemConst K_by_value) {
  const KernArgs<float> args = kernargs<float>();
  __shared__ TileLds<G> lds;
  tiled_step<G, kLatency>(args.a, args.K, xcd_block(args.a.nblocks), lds, (int)threadIdx.x);
}

// ------------------------------------------------------------------ chemistry stragglers
// The leaves the tiled kernel put on the straggler list (sub-step loop still
// running after chem_budget tries) take their whole chemistry here, one
// thread each, from the same inputs: the transported N, F the tiled kernel
// stored into the next state, the source and x n_H, eint, n_H of the current
// state.  A wave then holds 64 long loops instead of one long loop among 63
// finished lanes.  Each level-step launch has its own list-length word
// (cleared by the host before the call), so nothing is reset here.
template <int G>
__global__ __launch_bounds__(kStragglerThreads) void chem_straggler_kernel(LevelArgs<float> a_by_value,
                                                                            ChemConst K_by_value) {
  const KernArgs<float> args = kernargs<float>();
  const LevelArgs<float>& a = args.a;
  const ChemConst& K = args.K;
  const int count = min(*(volatile const int*)a.defer_ctl, a.defer_cap);
  const int n = a.ncell;
  unsigned int nit_acc = 0, ncell_acc = 0;
  for (int i = blockIdx.x * kStragglerThreads + (int)threadIdx.x; i < count; i += gridDim.x * kStragglerThreads) {
    ncell_acc++;
    const int cell = gld(a.defer_list, i);
    double e[G], src[G], fscale[G];
#pragma unroll
    for (int g = 0; g < G; g++) {
      e[g] = (double)gld(a.eo + g * n, cell);
      src[g] = (double)gld(a.src + g * n, cell);
    }
    double nhplus = gld(a.nhplus, cell), eint = gld(a.eint, cell), temp = 0.;
    const double nh = gld(a.nh, cell);
    int nit = 0;
    const int st = chem_cell_fast<G>(K, G, e, fscale, src, nhplus, eint, nh, temp, a.chemonly, nit);
    if (st != 0) raise_status(a.status, st, a.level, cell);
    nit_acc += (unsigned)nit;
#pragma unroll
    for (int g = 0; g < G; g++) {
      float ev = (float)e[g];
      if (e[g] >= K.emin_grp[g] && ev < K.emin_grp_up[g]) ev = K.emin_grp_up[g];
      const float fs = (float)fscale[g];
      const float fx = gld(a.fxo + g * n, cell) * fs, fy = gld(a.fyo + g * n, cell) * fs,
                  fz = gld(a.fzo + g * n, cell) * fs;
      gst(a.eo + g * n, cell, ev);
      gst(a.fxo + g * n, cell, fx);
      gst(a.fyo + g * n, cell, fy);
      gst(a.fzo + g * n, cell, fz);
    }
    gst(a.nhpluso, cell, nhplus);
    gst(a.einto, cell, eint);
    gst(a.tempo, cell, temp);
  }
  if (a.counters != nullptr) {
    // one atomic per wave that did work
    for (int o = 32; o > 0; o >>= 1) {
      nit_acc += __shfl_xor(nit_acc, o, 64);
      ncell_acc += __shfl_xor(ncell_acc, o, 64);
    }
    unsigned long long* slot = a.counters + (size_t)(blockIdx.x % kCounterSlots) * kCounterStride;
    if ((threadIdx.x & 63) == 0 && ncell_acc != 0u) {
      atomicAdd(slot, (unsigned long long)nit_acc);
      atomicAdd(slot + 2, (unsigned long long)ncell_acc);
    }
  }
}

// ------------------------------------------------------------------ COARSERAD transfers
// One thread per cell of the coarse level; a split cell walks its subtree
// depth first, children in index order like the reference's recursion
// (collectE, src_utils.c:94-123; distribE, rad_utils.c:26-45).
template <typename R, bool Put>
__global__ __launch_bounds__(256) void coarse_transfer_kernel(TreeView<R> t) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= t.ncell[0]) return;
  const int ch0 = t.child[0][i];
  if (ch0 < 0 || t.nlev < 2) return;
  const int n0 = t.ncell[0];
  R acc[4][kMaxGrp];
  if (Put) {
    for (int g = 0; g < t.ng; g++) acc[0][g] = acc[1][g] = acc[2][g] = acc[3][g] = R(0);
  } else {
    for (int g = 0; g < t.ng; g++) {
      acc[0][g] = t.e[0][g * n0 + i];
      acc[1][g] = t.fx[0][g * n0 + i];
      acc[2][g] = t.fy[0][g * n0 + i];
      acc[3][g] = t.fz[0][g * n0 + i];
    }
  }
  // explicit stack: octs of levels 1..nlev-1 and the next child to visit
  int oct[kMaxLevels], next[kMaxLevels];
  int k = 1;
  oct[1] = ch0;
  next[1] = 0;
  while (k >= 1) {
    if (next[k] == 8) {
      k--;
      continue;
    }
    const int cell = oct[k] * 8 + next[k];
    next[k]++;
    const int ch = t.child[k][cell];
    if (ch >= 0 && k + 1 < t.nlev) {
      k++;
      oct[k] = ch;
      next[k] = 0;
      continue;
    }
    if (ch >= 0) continue;  // split cell on the last level (no children stored)
    const int n = t.ncell[k];
    if (Put) {
      const R fact = (R)ldexp(1.0, -3 * k);  // pow(2, 3 (lcoarse - level))
      for (int g = 0; g < t.ng; g++) {
        acc[0][g] += t.e[k][g * n + cell] * fact;
        acc[1][g] += t.fx[k][g * n + cell] * fact;
        acc[2][g] += t.fy[k][g * n + cell] * fact;
        acc[3][g] += t.fz[k][g * n + cell] * fact;
      }
    } else {
      for (int g = 0; g < t.ng; g++) {
        t.e[k][g * n + cell] = acc[0][g];
        t.fx[k][g * n + cell] = acc[1][g];
        t.fy[k][g * n + cell] = acc[2][g];
        t.fz[k][g * n + cell] = acc[3][g];
      }
    }
  }
  if (Put) {
    for (int g = 0; g < t.ng; g++) {
      t.e[0][g * n0 + i] = acc[0][g];
      t.fx[0][g * n0 + i] = acc[1][g];
      t.fy[0][g * n0 + i] = acc[2][g];
      t.fz[0][g * n0 + i] = acc[3][g];
    }
  }
}

template <typename R>
hipError_t launch_put_e2coarse(const TreeView<R>& t, hipStream_t s) {
  hipLaunchKernelGGL((coarse_transfer_kernel<R, true>), dim3((t.ncell[0] + 255) / 256), dim3(256), 0, s, t);
  return hipGetLastError();
}

template <typename R>
hipError_t launch_distrib_e(const TreeView<R>& t, hipStream_t s) {
  hipLaunchKernelGGL((coarse_transfer_kernel<R, false>), dim3((t.ncell[0] + 255) / 256), dim3(256), 0, s, t);
  return hipGetLastError();
}

template hipError_t launch_put_e2coarse<double>(const TreeView<double>&, hipStream_t);
template hipError_t launch_put_e2coarse<float>(const TreeView<float>&, hipStream_t);
template hipError_t launch_distrib_e<double>(const TreeView<double>&, hipStream_t);
template hipError_t launch_distrib_e<float>(const TreeView<float>&, hipStream_t);

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

// ------------------------------------------------------------------ persistent coarse step
// The schedule's read-only tables are read through the constant address
// space: at a wave-uniform index the loads are scalar loads, and the level
// step's arguments inside a StepDesc stay in SGPRs like kernel arguments.
#define EMMA_CONST __attribute__((address_space(4)))
template <typename T>
__device__ __forceinline__ const T& cref(const void* base, int i) {
  return *(const T*)((const EMMA_CONST T*)base + i);
}

__device__ __forceinline__ const SchedArgs& sched_args() {
  return *reinterpret_cast<const SchedArgs*>((const char*)__builtin_amdgcn_kernarg_segment_ptr());
}

// wait until ticket dt is done in this launch (epoch), bounded by wall-clock
// time: past the bound the wait is abandoned with EMMA_RT_ERR_SCHED, and
// once that status is raised no other wait starts, so that every work-group
// ends within about one bound
__device__ __forceinline__ void wait_done(const SchedArgs& s, int dt, int epoch) {
  if (__hip_atomic_load(s.done + dt, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == epoch) return;
  const long long t0 = (long long)__builtin_amdgcn_s_memrealtime();
  while (__hip_atomic_load(s.done + dt, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != epoch) {
    if (__hip_atomic_load(s.status, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == kErrSched) return;
    __builtin_amdgcn_s_sleep(2);
    if ((long long)__builtin_amdgcn_s_memrealtime() - t0 > s.wait_ticks) {
      raise_status(s.status, kErrSched, -1, dt);
      return;
    }
  }
}

template <int G>
__device__ __forceinline__ TileLds<G>* tile_lds() {
  __shared__ TileLds<G> lds;
  return &lds;
}

// The thread index as a value the optimiser cannot see through: otherwise
// every address and mask the item derives from it alone is hoisted out of
// the work-group's ticket loop and held in registers across all items.
__device__ __forceinline__ int opaque_thread() {
  int t = (int)threadIdx.x;
  asm volatile("" : "+v"(t));
  return t;
}

// one tile of an item: FillRad's block on the tile's cells, the tile's
// level step, the hydro injection of its cells.  Each thread fills and
// injects the cell it steps, so the three need no barrier between them;
// neighbours' cells are read only for their N, F and split flags, which
// FillRad does not touch.
template <typename R, int G, bool Tiled, bool Coupled>
__device__ __forceinline__ void run_item(const StepDesc<R>& D, int tile) {
  const int lc = opaque_thread();
  const int cell = tile * kTileCells + lc;
  if constexpr (Coupled)
    if (D.fill && cell < D.f.ncell) fill_cell<R>(D.f, cell);
  if constexpr (Tiled) {
    if (D.variant == kListed)
      tiled_step<G, kListed>(D.a, D.K, tile, *tile_lds<G>(), lc);
    else
      tiled_step<G, kPerDir>(D.a, D.K, tile, *tile_lds<G>(), lc);
  } else {
    cell_step<R, G, sizeof(R) == 4>(D.a, D.K, tile, lc);
  }
  if constexpr (Coupled)
    if (D.hydro && cell < D.h.ncell) to_hydro_cell(D.h, cell);
}

// Work-groups loop over tickets of their XCD's queue (the XCC_ID hardware
// register names the XCD): take the next one, wait for the tickets it
// depends on (lanes of wave 0, one dependency each), acquire at agent scope,
// run the item, release, mark the ticket done.  The release writes the L2
// back (agent scope) only for items with a dependent in another queue:
// within a queue the producer's stores are in the XCD's shared L2 once they
// have completed.  One wave of the work-group does the acquire and the
// release for all four.  Items on the schedule's critical path run at a
// raised wave priority so that the bandwidth-bound items sharing their SIMDs
// do not stretch the chain.  The last work-group to leave re-arms the queues
// (tickets 0, next epoch) for the next launch.  Coupled: FillRad's block and
// the hydro injection are compiled in (without them the kernel needs fewer
// registers).
//
// The work-group's own bookkeeping (ticket, done flag, exit count) is done
// by all of wave 0 under a wave-uniform branch, never by one lane: a
// lane-divergent block at each end of a loop that holds barriers lets the
// compiler split the loop by lane (thread 0 looping over tickets, the other
// lanes over the old ticket), which never ends.
template <typename R, int G, bool Tiled, bool Coupled>
__global__ __launch_bounds__(kTileCells) __attribute__((amdgpu_waves_per_eu(Tiled ? 4 : 1))) void persistent_step_kernel(
    SchedArgs s_by_value) {
  const SchedArgs& s = sched_args();
  __shared__ int sTicket, sEpoch;
  const int wave = __builtin_amdgcn_readfirstlane((int)threadIdx.x >> 6);
  const int lane = (int)threadIdx.x & 63;
  // the queue of this work-group's XCD
  int xcc;
  asm volatile("s_getreg_b32 %0, hwreg(HW_REG_XCC_ID)" : "=s"(xcc));
  const int q = s.xcc_queue[xcc & 15];
  // (an XCD the library did not see serves no queue)
  const int qb = q >= 0 ? cref<int>(s.qoff, q) : 0, qn = q >= 0 ? cref<int>(s.qoff, q + 1) - qb : 0;
  if (wave == 0) sEpoch = __hip_atomic_load(s.ctl + s.nq + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  int epoch = 0;
  for (;;) {
    if (wave == 0) {  // lane 0 adds one, the other lanes nothing
      const int old = q >= 0 ? atomicAdd(s.ctl + q, lane == 0 ? 1 : 0) : 0;
      sTicket = __builtin_amdgcn_readfirstlane(old);
    }
    __syncthreads();
    const int tq = __builtin_amdgcn_readfirstlane(sTicket);
    epoch = __builtin_amdgcn_readfirstlane(sEpoch);
    if (tq >= qn) break;
    const int t = qb + tq;
    const int flags = cref<unsigned char>(s.item_flags, t);
    const int d0 = cref<int>(s.dep_off, t), d1 = cref<int>(s.dep_off, t + 1);
    if (d1 > d0) {
      if (wave == 0) {
        for (int i = d0 + lane; i < d1; i += 64) wait_done(s, s.deps[i], epoch);
        // one acquire at agent scope for the work-group (its waves share
        // this CU's caches), also for dependencies on this XCD's items: an
        // L1-only invalidate (buffer_inv sc0) left same-XCD results stale in
        // measurements, the agent-scope invalidate does not
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
      }
      __syncthreads();
    }
    const int step = cref<int>(s.item_step, t), tile = cref<int>(s.item_tile, t);
    const int ntile = cref<int>(s.item_ntile, t);
    const int prio = flags & kItemPrio;
    if (prio) __builtin_amdgcn_s_setprio(2);
    for (int k = 0; k < ntile; k++) {
      if (k > 0) __syncthreads();  // the tile step's LDS is reused
      run_item<R, G, Tiled, Coupled>(cref<StepDesc<R>>(s.steps, step), tile + k);
    }
    if (prio) __builtin_amdgcn_s_setprio(0);
    // every wave's stores have reached the L2 before wave 0 publishes the
    // ticket; for readers on other XCDs it writes the L2 back first (release
    // at agent scope)
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    if (wave == 0) {
      if (flags & kItemReleaseAgent) __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
      __hip_atomic_store(s.done + t, epoch, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
  }
  if (wave == 0) {
    const int left = __builtin_amdgcn_readfirstlane(atomicAdd(s.ctl + s.nq, lane == 0 ? 1 : 0));
    if (left == (int)gridDim.x - 1) {
      for (int i = lane; i < s.nq + 1; i += 64) s.ctl[i] = 0;
      if (lane == 0) s.ctl[s.nq + 1] = epoch + 1;
    }
  }
}

// ------------------------------------------------------------------ launchers
template <typename R>
hipError_t launch_level_step(const LevelArgs<R>& a0, const ChemConst& K, hipStream_t s) {
  LevelArgs<R> a = a0;
  if constexpr (sizeof(R) == 4) {
    if (a.ng == 1 || a.ng == 3) {
      int nb = (a.ncell + kTileCells - 1) / kTileCells;
      nb = (nb + 7) & ~7;
      a.nblocks = nb;
      // sparse levels: per-direction halo faces
      const int v = a.variant;
      const bool defer = a.chem_budget > 0;
      if (defer && (a.defer_list == nullptr || a.defer_ctl == nullptr || a.defer_cap < a.ncell))
        return hipErrorInvalidValue;
      // the list holds at most the level's cells
      const dim3 sg(std::min(kStragglerBlocks, (a.ncell + kStragglerThreads - 1) / kStragglerThreads)),
          sb(kStragglerThreads);
      if (a.ng == 1) {
        if (v == kLatency)
          hipLaunchKernelGGL((level_step_tiled_latency<1>), dim3(nb), dim3(kTileCells), 0, s, a, K);
        else if (v == kPerDir)
          hipLaunchKernelGGL((level_step_tiled<1, true>), dim3(nb), dim3(kTileCells), 0, s, a, K);
        else
          hipLaunchKernelGGL((level_step_tiled<1, false>), dim3(nb), dim3(kTileCells), 0, s, a, K);
        if (defer) hipLaunchKernelGGL((chem_straggler_kernel<1>), sg, sb, 0, s, a, K);
      } else {
        if (v == kLatency)
          hipLaunchKernelGGL((level_step_tiled_latency<3>), dim3(nb), dim3(kTileCells), 0, s, a, K);
        else if (v == kPerDir)
          hipLaunchKernelGGL((level_step_tiled<3, true>), dim3(nb), dim3(kTileCells), 0, s, a, K);
        else
          hipLaunchKernelGGL((level_step_tiled<3, false>), dim3(nb), dim3(kTileCells), 0, s, a, K);
        if (defer) hipLaunchKernelGGL((chem_straggler_kernel<3>), sg, sb, 0, s, a, K);
      }
      return hipGetLastError();
    }
  }
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

__global__ void xcc_probe_kernel(int* out, int n) {
  int xcc;
  asm volatile("s_getreg_b32 %0, hwreg(HW_REG_XCC_ID)" : "=s"(xcc));
  if (threadIdx.x == 0 && (int)blockIdx.x < n) out[blockIdx.x] = xcc & 15;
}

hipError_t launch_xcc_probe(int* out, int n, hipStream_t st) {
  hipLaunchKernelGGL(xcc_probe_kernel, dim3(n), dim3(64), 0, st, out, n);
  return hipGetLastError();
}

template <typename R, int G, bool Tiled>
static void* persistent_kernel_of(bool coupled) {
  return coupled ? reinterpret_cast<void*>(persistent_step_kernel<R, G, Tiled, true>)
                 : reinterpret_cast<void*>(persistent_step_kernel<R, G, Tiled, false>);
}
template <typename R>
static void* persistent_kernel(int ng, bool coupled) {
  if constexpr (sizeof(R) == 4) {
    if (ng == 1) return persistent_kernel_of<float, 1, true>(coupled);
    if (ng == 3) return persistent_kernel_of<float, 3, true>(coupled);
    return persistent_kernel_of<float, 0, false>(coupled);
  } else {
    if (ng == 1) return persistent_kernel_of<double, 1, false>(coupled);
    if (ng == 3) return persistent_kernel_of<double, 3, false>(coupled);
    return persistent_kernel_of<double, 0, false>(coupled);
  }
}

template <typename R>
int persistent_blocks_per_cu(int ng, bool coupled) {
  int n = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, persistent_kernel<R>(ng, coupled), kTileCells, 0) != hipSuccess)
    return 0;
  return n;
}

template <typename R>
hipError_t launch_persistent(const SchedArgs& s, int ng, bool coupled, int nblocks, hipStream_t st) {
  void* args[] = {const_cast<SchedArgs*>(&s)};
  return hipLaunchKernel(persistent_kernel<R>(ng, coupled), dim3(nblocks), dim3(kTileCells), args, 0, st);
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
template int persistent_blocks_per_cu<double>(int, bool);
template int persistent_blocks_per_cu<float>(int, bool);
template hipError_t launch_persistent<double>(const SchedArgs&, int, bool, int, hipStream_t);
template hipError_t launch_persistent<float>(const SchedArgs&, int, bool, int, hipStream_t);
template hipError_t launch_photon_sum<double>(const double*, const int*, int, int, double, double*, hipStream_t);
template hipError_t launch_photon_sum<float>(const float*, const int*, int, int, double, double*, hipStream_t);

}  // namespace emma
