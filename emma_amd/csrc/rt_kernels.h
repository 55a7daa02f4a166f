// rt_kernels.h -- kernel argument blocks and launchers of the resident path.
#pragma once

#include <hip/hip_runtime.h>

#include "rt_physics.h"

namespace emma {

// Everything one level step reads and writes.  R is the storage type of the
// photon density / flux / source fields (double, or float in mixed mode).
template <typename R>
struct LevelArgs {
  int ncell, ncoarse, nfine, ng;
  int level, chemonly, nblocks, pad;
  // current state (read)
  const R* e;
  const R* fx;
  const R* fy;
  const R* fz;
  const double* nhplus;
  const double* eint;
  // fields that are not double-buffered (written only for split cells)
  R* src;
  double* nh;
  // next state (written)
  R* eo;
  R* fxo;
  R* fyo;
  R* fzo;
  double* nhpluso;
  double* einto;
  double* tempo;
  // topology: child[cell] >= 0 split, -1 leaf, <= -2 leaf with reflux slot -2-child
  const int* child;
  const int* nei;
  R* racc;  // this level's reflux slots [nslot][4][GA]
  // coarser level: current state (injection), child table, reflux slots
  const R* ce;
  const R* cfx;
  const R* cfy;
  const R* cfz;
  const int* cchild;
  R* cracc;
  // finer level: current state (restriction)
  const R* fe;
  const R* ffx;
  const R* ffy;
  const R* ffz;
  const R* fsrc;
  const double* fnhplus;
  const double* feint;
  const double* ftemp;
  const double* fnh;
  // reionisation-time bookkeeping (null when disabled)
  double* t_first;
  double* t_last;
  const double* ft_first;
  const double* ft_last;
  double tphy;
  // numerics
  R c, hc, dtsurdx;
  int* status;                   // [code, level, cell]
  unsigned long long* counters;  // kCounterSlots x [accepted sub-steps, leaf cell-updates] or null
};

// Counters are spread over kCounterSlots cache lines (block b adds into slot
// b % kCounterSlots) so that thousands of blocks do not serialise on one
// address; the host sums the slots.
constexpr int kCounterSlots = 256;
constexpr int kCounterStride = 16;  // unsigned long longs per slot (128 B)

template <typename R>
hipError_t launch_level_step(const LevelArgs<R>& a, const ChemConst& K, hipStream_t s);

template <typename R>
hipError_t launch_photon_sum(const R* e, const int* child, int ncell, int ng, double vol, double* out,
                             hipStream_t s);

hipError_t launch_fill(unsigned int* p, int n, unsigned int v, hipStream_t s);

}  // namespace emma
