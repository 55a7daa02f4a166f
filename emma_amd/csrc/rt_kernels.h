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
  int level, chemonly, nblocks;
  int nslot;   // accumulators in racc (reflux pairs x kReflux sub-slots)
  // RADSTEP: transport=1, keep_split_flux=0, update_split=0.  COARSERAD
  // (rad_utils.c:518-526,1548-1552,1808-1812): transport only on the coarse
  // level and not for chemonly calls, fluxes of split neighbours kept, split
  // cells transport-updated (their chemistry still skipped).
  int transport, keep_split_flux, update_split;
  // mixed path, 1 or 3 groups: the halo variant of the tiled kernel
  // (kListed, kPerDir or kLatency)
  int variant;
  // > 0: the level is a complete periodic cube of 2^raster_m octs a side in
  // Morton order, and the tiled kernel walks each XCD's octant of tiles in
  // raster order (z fastest) instead of Morton order (see raster_tile)
  int raster_m;
  int cnslot;  // accumulators in cracc
  // chemistry stragglers (tiled mixed kernel): a leaf whose sub-step loop is
  // still running after chem_budget tries (> 0; <= 0: no limit) stores its
  // transported N, F, appends its cell to defer_list (capacity defer_cap)
  // and leaves its chemistry to chem_straggler_kernel, which restarts it from
  // the same inputs.  defer_ctl: this launch's list length (zero before it).
  int chem_budget, defer_cap;
  float chem_keep;  // > 0: a leaf with more than this many sub-steps to go at its pace so far is not deferred
  int* defer_list;
  int* defer_ctl;
  // current state (read)
  const R* e;
  const R* fx;
  const R* fy;
  const R* fz;
  const double* nhplus;
  const double* eint;
  const double* temp;
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
  // topology: child[cell] >= 0 split, -1 leaf, <= -2 leaf with finer octs
  // behind some of its faces: -2-child = (first reflux pair << 3) | pairs.
  // A reflux pair is one (coarse leaf, face) with a finer oct behind it; a
  // leaf's pairs are consecutive, in face order.  Each pair has one
  // accumulator per flux component and group for the tiled mixed kernel
  // (its four face cells are summed across lanes first), four (one per face
  // cell) for the per-cell kernel: every accumulator has one writer per
  // launch, so refluxing needs no atomics and sums in a fixed order.
  const int* child;
  const int* nei;
  const int* rpair;  // [noct][6]: the coarser level's reflux pair of each oct face, -1 if none
  R* racc;           // this level's reflux accumulators: [nslot][4][GA], an accumulator's values contiguous
  // coarser level: current state (injection), child table, reflux accumulators
  const R* ce;
  const R* cfx;
  const R* cfy;
  const R* cfz;
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
  unsigned long long* counters;  // kCounterSlots x [accepted sub-steps, leaf cell-updates, deferred leaves] or null
};

// Counters are spread over kCounterSlots cache lines (block b adds into slot
// b % kCounterSlots) so that thousands of blocks do not serialise on one
// address; the host sums the slots.
constexpr int kCounterSlots = 256;
constexpr int kCounterStride = 16;  // unsigned long longs per slot (128 B)

// tiles of the mixed-precision level step: 256 cells = 32 Morton-consecutive
// octs per work-group of four waves
// nei_oct entries of domain-boundary faces (EMMA_RT_NEI_TRANSMISSIVE /
// EMMA_RT_NEI_REFLECTIVE): the ghost across the face is a copy of a cell of
// the same oct, its normal flux negated when reflective; no refluxing
constexpr int kNeiTransmissive = -2147483647 - 1;
constexpr int kNeiReflective = -2147483647;

// the cell of the boundary oct whose copy cell `ic` sees across the domain
// face f: itself, except on +y, where the reference's face table
// (rad_utils.c:1124-1139: face[4] = 7, face[5] = 6) makes cells 6 and 7 see
// each other
__host__ __device__ __forceinline__ int trans_source(int f, int ic) { return (f == 3 && ic >= 6) ? ic ^ 1 : ic; }

// reflux accumulators per (coarse leaf, face) pair: one per face cell for
// the per-cell kernel, one for the tiled kernel (face summed across lanes)
constexpr int kRefluxCell = 4;
constexpr int kRefluxTiled = 1;
constexpr int kTileCells = 256;
constexpr int kTileOcts = kTileCells / 8;
constexpr int kTileWaves = kTileCells / 64;
// Halo variants of the tiled level step.
//  kListed  each wave lists its halo faces (neighbours outside the tile) and
//           evaluates them lane-parallel in one round; at most 64 per wave,
//           which holds on complete levels (16 in x and y, 32 in z).
//  kPerDir  each lane evaluates its own halo face of direction d at the
//           start of pass d (up to 192 per wave: sparse AMR levels); the
//           leaves' chemistry inputs are loaded at the start of the last
//           pass (fewer live registers: four work-groups per CU).
//  kLatency as kPerDir, but the halo faces of all three directions and the
//           chemistry inputs are loaded up front: one memory round trip
//           instead of four, with the registers that takes (two waves per
//           SIMD).  For levels too small to fill the GPU, whose step is one
//           work-group's latency chain.
constexpr int kListed = 0;
constexpr int kPerDir = 1;
constexpr int kLatency = 2;
// kListed on a level with no coarser level (the tree's coarsest): no coarse
// faces and no injected leaves, so the code for them is compiled out.  The
// kernels pick it from the level's arguments (ncoarse == 0); not a caller's
// choice.
constexpr int kListedTop = 3;
// grid of the chemistry straggler pass (it loops over the list)
constexpr int kStragglerThreads = 256;
constexpr int kStragglerBlocks = 4096;

// Most cells in one level with ng groups: cell and group*cell indices are
// 32-bit ints and every byte offset into a per-group array (8-byte elements at
// most) stays below 4 GiB.
inline long long max_level_cells(int ng) {
  const long long by_index = 2147483647LL / (ng > 0 ? ng : 1);
  return by_index < (1LL << 29) ? by_index : (1LL << 29);
}

template <typename R>
hipError_t launch_level_step(const LevelArgs<R>& a, const ChemConst& K, hipStream_t s);

// Levels of a tree for the COARSERAD transfers between the coarse level and
// the leaves (putE2coarse / distribE): per level (index level - lcoarse) the
// child table and the current photon density / flux.
constexpr int kMaxLevels = 16;
template <typename R>
struct TreeView {
  int nlev, ng, lcoarse, pad;
  int ncell[kMaxLevels];
  const int* child[kMaxLevels];
  R* e[kMaxLevels];
  R* fx[kMaxLevels];
  R* fy[kMaxLevels];
  R* fz[kMaxLevels];
};

// putE2coarse (src_utils.c:127-150) on every split cell of the coarse level
template <typename R>
hipError_t launch_put_e2coarse(const TreeView<R>& t, hipStream_t s);
// distribE (rad_utils.c:26-45) from every split cell of the coarse level
template <typename R>
hipError_t launch_distrib_e(const TreeView<R>& t, hipStream_t s);

// FillRad's per-cell block (src_utils.c:738-838, init = 0) on one level's
// current state: with hydro, n_H = d (1-YHE), eint = p/(GAMMA-1), x n_H = dX;
// split cells src = 0; unsplit cells the Stromgren test source (star) and,
// with hydro, E2T.  The source term is uniform: the host evaluates its
// value, the time tests and the active time group, the kernel only selects
// the cells.
template <typename R>
struct FillArgs {
  int ncell, ng, hydro, star;
  const int* child;
  double* nh;
  double* eint;
  double* nhplus;
  double* temp;
  R* src;
  const double* d;
  const double* p;
  const double* dX;
  const double* opos;  // octs' lower corners, 3 per oct
  double dxcur, X0, xs, ys, zs;
  // groups [star_g0, star_g1) of the time group tcur lies in get star_src
  // when the lifetime test passed (stromgren_source); empty otherwise
  int star_g0, star_g1;
  double star_src;
  double pa3, pa5, pstar, unit_N;  // E2T
  // star particles (stars_sources, src_utils.c:209-335): the particles of
  // leaf cell c are first[c] .. first[c+1]-1, in the order of the cell's
  // particle list.  A particle with isStar >= 6 (other than 100, the AGN
  // case) adds to the NGRP_SPACE groups of its radiative_state the value
  // ((mass unit_d) srcint / dx^3) unit_t / unit_N a^2 while tphy > age
  // (t = (tphy - age) / tlife > 0); with decrease (the reference's
  // DECREASE_EMMISIVITY_AFTER_TLIFE) stars older than tlife keep radiating
  // at that value times t^-4.  star_first == nullptr: no particles.
  int stars, stars_decrease, nspace;
  const int* star_first;
  const double* star_mass;
  const double* star_age;
  const int* star_rs;
  const int* star_is;
  double stars_unit_d, stars_srcint, stars_dx3, stars_unit_t, stars_unit_N, stars_pa2, stars_tphy, stars_tlife;
};
template <typename R>
hipError_t launch_fill_level(const FillArgs<R>& f, hipStream_t s);

// RadSolver's injection into the hydro state (rad_utils.c:1941-1951,
// 1993-2004), every cell: p = (GAMMA-1) eint, dX = x d (1-YHE), E = getE.
struct HydroArgs {
  int ncell;
  const double* eint;
  const double* nhplus;
  const double* nh;
  const double* d;
  const double* u;
  const double* v;
  const double* w;
  double* p;
  double* dX;
  double* E;
};
hipError_t launch_rad_to_hydro(const HydroArgs& h, hipStream_t s);

// ---------------------------------------------------------------- persistent coarse step
// A coarse step (all its level steps, with FillRad's block before and the
// hydro injection after each) as ONE launch of persistent work-groups fed by
// a device work queue.  An item is a run of consecutive tiles (kTileOcts
// octs each, the cells of one level-step work-group) of one level step: one
// tile on all but the largest levels.  Tickets are handed out in the
// order the host's list schedule chose (a topological order of the items'
// dependencies); a work-group takes the next ticket, waits until the tickets
// it depends on are done, runs the item and marks its ticket done.  Every
// wait is on a smaller ticket, held by a work-group that is running, so the
// queue cannot deadlock whatever the number of resident work-groups.

// LDS of one tile step (all halo variants use the same layout, so a kernel
// that runs several variants holds it once)
template <int G>
struct TileLds {
  static constexpr int kHaloCap = 64;  // halo slots per wave
  float4 share[G * 2 * kTileCells + kTileWaves * kHaloCap * G];
  unsigned char cancel[kTileCells + kTileWaves * kHaloCap];
  unsigned char item[kTileWaves][kHaloCap];
  int item_nei[kTileWaves][kHaloCap];
};

// one level step of the schedule: its arguments, FillRad's block before it
// and the hydro injection after it (fill / hydro: 0 = skip)
template <typename R>
struct StepDesc {
  LevelArgs<R> a;
  ChemConst K;
  FillArgs<R> f;
  HydroArgs h;
  int fill, hydro, variant, pad;
};

constexpr int kErrSched = -10;  // EMMA_RT_ERR_SCHED

// item flags (as the planner writes them)
constexpr unsigned char kItemPrio = 1;          // raised wave priority
constexpr unsigned char kItemAcquireAgent = 2;  // reads results of another queue's items (every item with
                                                // dependencies acquires at agent scope, see the kernel)
constexpr unsigned char kItemReleaseAgent = 4;  // another queue's items read its results

struct SchedArgs {
  const void* steps;        // StepDesc<R>[nsteps]
  const int* qoff;          // [nq + 1]: queue q's tickets are qoff[q] .. qoff[q+1]-1
  const int* item_step;     // ticket -> level step
  const int* item_tile;     // ticket -> first tile of that level step
  const int* item_ntile;    // ticket -> consecutive tiles of the item
  const unsigned char* item_flags;  // ticket -> kItem* flags
  const int* dep_off;       // [nitems + 1]: the tickets ticket t waits for are deps[dep_off[t] .. dep_off[t+1])
  const int* deps;
  int* done;                // [nitems]: == epoch once the ticket's item is complete
  int* ctl;                 // [q < nq] queue q's next ticket, [nq] work-groups that left, [nq + 1] epoch
  int nitems;
  int nq;                   // queues: one per XCD; a work-group serves the queue of its XCD
  int xcc_queue[16];        // XCC_ID (low 4 bits) -> queue, -1: an XCD the library did not see
  int* status;              // the library's status word (a wait past its bound: EMMA_RT_ERR_SCHED)
  long long wait_ticks;     // bound of one dependency wait, s_memrealtime ticks (100 MHz)
};

// the persistent launch: R and ng select the kernel family of the level
// steps (the tiled mixed kernel for G = 1, 3; the per-cell kernel
// otherwise); coupled: the steps carry FillRad's block / the hydro injection
template <typename R>
hipError_t launch_persistent(const SchedArgs& s, int ng, bool coupled, int nblocks, hipStream_t st);
// the XCC_ID (low 4 bits) of the XCD each of `n` work-groups ran on
hipError_t launch_xcc_probe(int* out, int n, hipStream_t st);
// work-groups of the persistent kernel one CU holds at once
template <typename R>
int persistent_blocks_per_cu(int ng, bool coupled);

template <typename R>
hipError_t launch_photon_sum(const R* e, const int* child, int ncell, int ng, double vol, double* out,
                             hipStream_t s);

hipError_t launch_fill(unsigned int* p, int n, unsigned int v, hipStream_t s);

}  // namespace emma
